"""GPU checks of the boundary contract through the C ABI: a workspace is
bound to its plan (FftError::WorkspaceMismatch, distributed.rs:2188-2191,
:4015-4017), overlapping buffers are refused, and an in-place operation that
fails after its first write leaves the array Poisoned
(run_in_place_transaction, distributed.rs:3776-3794; ArrayError::Poisoned,
many.rs:234) until the workspace is reset. The host emulation covers the
same rules without a GPU (tests/test_planner_emu.py)."""
import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import rel_l2, uniform_real

pytestmark = pytest.mark.gpu


def _tdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _in_place_spectrum(plan, spec):
    """An in-place R2C array holding spec (the packed output layout)."""
    data = torch.zeros(spec.numel(), dtype=spec.dtype, device="cuda")
    data.copy_(spec)
    return data


@pytest.mark.parametrize("deferred", [False, True])
def test_in_place_failure_poisons_the_workspace(gpu, deferred):
    shape = (8, 12, 16)
    rng = np.random.default_rng(41)
    x = uniform_real(rng, shape, np.float64)
    plan = pf.R2cPlan(shape, dtype=torch.float64)
    ws = plan.allocate_workspace()
    spec = plan.forward(_tdev(x), workspace=ws)
    bad = _host(spec).copy()
    bad[0] += 0.5j  # DC-plane value of the reduction axis
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_R2C, bad, O.INVERSE, shape=shape)
    data = _in_place_spectrum(plan, _tdev(bad))
    flags = pf.EXEC_DEFER_STATUS if deferred else 0
    if deferred:
        plan.execute_in_place(pf.INVERSE, data, workspace=ws, flags=flags)
        # the array was partly overwritten; a later deferred call writes nothing
        after_fail = data.clone()
        plan.execute_in_place(pf.FORWARD, data, workspace=ws, flags=flags)
        torch.cuda.synchronize()
        assert torch.equal(data, after_fail)
        with pytest.raises(pf.PfhError) as e:
            plan.take_status(ws)
        assert e.value.variant == "InvalidSpectrum"
    else:
        with pytest.raises(pf.PfhError) as e:
            plan.execute_in_place(pf.INVERSE, data, workspace=ws)
        assert e.value.variant == "InvalidSpectrum"
    after_fail = data.clone()
    for op in (pf.FORWARD, pf.INVERSE, pf.BACKWARD):
        with pytest.raises(pf.PfhError) as e:
            plan.execute_in_place(op, data, workspace=ws)
        assert e.value.variant == "Poisoned"
    assert torch.equal(data, after_fail)
    # out-of-place calls with the same workspace are unaffected
    back = plan.inverse(spec, workspace=ws)
    assert rel_l2(_host(back), x.reshape(-1)) <= 1e-12
    # a reset workspace runs in-place calls again
    plan.init_workspace(ws)
    good = _in_place_spectrum(plan, spec)
    plan.execute_in_place(pf.INVERSE, good, workspace=ws)
    got = _host(good.view(torch.float64)[: x.size])
    assert rel_l2(got, x.reshape(-1)) <= 1e-12


def test_workspace_is_bound_to_its_plan(gpu):
    shape = (6, 10, 8)
    rng = np.random.default_rng(42)
    xc = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64).reshape(-1)
    a = pf.C2cPlan(shape, dtype=torch.complex64)
    b = pf.C2cPlan(shape, dtype=torch.complex64)
    wa, wb = a.allocate_workspace(), b.allocate_workspace()
    x = _tdev(xc)
    ya = a.forward(x, workspace=wa)
    b.forward(x, workspace=wb)
    dst = torch.full_like(x, 7.0)
    with pytest.raises(pf.PfhError) as e:
        b.forward(x, dst, workspace=wa)
    assert e.value.variant == "WorkspaceMismatch"
    torch.cuda.synchronize()
    assert torch.all(dst == 7.0)
    with pytest.raises(pf.PfhError) as e:
        b.forward_in_place(dst, workspace=wa)
    assert e.value.variant == "WorkspaceMismatch"
    assert torch.all(dst == 7.0)
    # rebinding by init, and a fresh zero-filled workspace binds to its first user
    b.init_workspace(wa)
    assert torch.equal(b.forward(x, workspace=wa), ya)
    with pytest.raises(pf.PfhError):
        a.forward(x, workspace=wa)
    fresh = torch.zeros_like(wa)
    assert torch.equal(a.forward(x, workspace=fresh), ya)


def test_overlapping_buffers_are_refused(gpu):
    shape = (6, 10, 8)
    plan = pf.C2cPlan(shape, dtype=torch.complex64)
    n = int(np.prod(shape))
    big = torch.zeros(2 * n + 8, dtype=torch.complex64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.forward(big[:n], big[4: 4 + n])
    assert e.value.variant == "AliasedBuffers"
    with pytest.raises(pf.PfhError) as e:
        plan.forward(big[:n], big[:n])
    assert e.value.variant == "AliasedBuffers"
    ws = plan.allocate_workspace()
    src = torch.zeros(n, dtype=torch.complex64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.forward(src, ws.view(torch.complex64)[:n], workspace=ws)
    assert e.value.variant == "AliasedBuffers"
    # adjacent, non-overlapping halves of one allocation are fine
    plan.forward(big[:n], big[n: 2 * n])
