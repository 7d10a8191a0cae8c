"""pfh_execute_noalloc, the allocation-free variant of pfh_execute (its
workspace is the library's static device scratch), against pfh_execute
through the C ABI: bit-equal results on consecutive calls, the later ones
with smaller problems, so each call finds the previous call's contents in
the scratch; a rejected spectrum; operations whose workspace exceeds the
scratch and non-zero flags are refused."""
import numpy as np
import pytest
import torch

import pencil_fft_hip as pf

pytestmark = pytest.mark.gpu


def _noalloc(plan, op, src, dst, flags=0):
    lib = pf.load_library()
    return lib.pfh_execute_noalloc(plan._h, op, pf._ptr(src), src.numel(), pf._ptr(dst), dst.numel(), pf._stream(),
                                   flags)


def _bits(t):
    torch.cuda.synchronize()
    return t.view(torch.uint8).cpu().numpy().tobytes()


def _fresh_like(t):
    return torch.full_like(t, float("nan"))


CASES = [
    # (plan factory, input dtype): consecutive calls, later ones smaller
    (lambda: pf.R2cPlan((64, 48, 40), dtype=torch.float32, extra_shape=(2,)), torch.float32),
    (lambda: pf.R2cPlan((16, 12, 10), dtype=torch.float32), torch.float32),
    (lambda: pf.C2cPlan((24, 20, 18), dtype=torch.complex128), torch.complex128),
    (lambda: pf.R2rPlan((12, 10, 9), [pf.R2rKind.DctII, pf.R2rKind.DstI, pf.R2rKind.DctIV],
                        dtype=torch.float64), torch.float64),
    (lambda: pf.MixedR2cPlan((6, 5, 8), [pf.R2rKind.DctII, pf.AxisTransform.FFT, pf.AxisTransform.RFFT],
                             dtype=torch.float32), torch.float32),
    # a long axis (1009 points, prime): its stage runs as Bluestein pieces
    # through workspace regions, here inside the static scratch
    (lambda: pf.C2cPlan((3, 1009, 5), dtype=torch.complex64), torch.complex64),
]


def test_noalloc_matches_execute_on_consecutive_calls(gpu):
    g = torch.Generator(device="cuda")
    g.manual_seed(77)
    for make, dt in CASES:
        plan = make()
        x = torch.randn(plan.input_elements(), generator=g, device="cuda", dtype=dt)
        ref = plan.forward(x)
        got = _fresh_like(ref)
        pf._check(_noalloc(plan, pf.FORWARD, x, got))
        assert _bits(got) == _bits(ref)
        for op in (pf.INVERSE, pf.BACKWARD, pf.ADJOINT_FORWARD):
            want = plan.execute(op, ref)
            out = _fresh_like(want)
            pf._check(_noalloc(plan, op, ref, out))
            assert _bits(out) == _bits(want), (op,)


def test_noalloc_rejected_spectrum_leaves_dst(gpu):
    plan = pf.R2cPlan((16, 12, 10), dtype=torch.float64)
    x = torch.randn(plan.input_elements(), device="cuda", dtype=torch.float64)
    spec = plan.forward(x)
    bad = spec.clone()
    bad[0] += 0.5j  # DC-plane value of the reduction axis
    dst = torch.full_like(x, -1.0)
    with pytest.raises(pf.PfhError) as e:
        pf._check(_noalloc(plan, pf.INVERSE, bad, dst))
    assert e.value.variant == "InvalidSpectrum"
    torch.cuda.synchronize()
    assert torch.all(dst == -1.0)
    # the next call on the same scratch is not affected by the rejection
    back = _fresh_like(x)
    pf._check(_noalloc(plan, pf.INVERSE, spec, back))
    assert _bits(back) == _bits(plan.inverse(spec))


def test_noalloc_refuses_large_workspaces_and_flags(gpu):
    big = pf.R2cPlan((512, 512, 512), dtype=torch.float32)
    assert big.workspace_bytes(pf.FORWARD) > 128 << 20
    t = torch.zeros(1, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        pf._check(_noalloc(big, pf.FORWARD, t, t))
    assert e.value.variant == "WorkspaceTooSmall"
    assert e.value.required == big.workspace_bytes(pf.FORWARD) and e.value.actual == 128 << 20
    small = pf.R2cPlan((8, 6, 4), dtype=torch.float32)
    x = torch.zeros(small.input_elements(), device="cuda")
    y = torch.zeros(small.output_elements(), dtype=torch.complex64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        pf._check(_noalloc(small, pf.FORWARD, x, y, flags=pf.EXEC_DEFER_STATUS))
    assert e.value.variant == "InvalidArgument"
