"""GPU parity of long lines against the oracle.

Every length that LocalC2cPlan::new accepts plans (crates/pencil-fft/src/
lib.rs:422-438). A line too long for one workgroup runs as a four-step
split into two line stages, with the twiddles applied at the first stage's
store. A length with a prime factor above 128 runs through Bluestein's
chirp convolution of 5-smooth length M >= 2n - 1. Real kinds promote the
line to a complex FFT between a pre-processing kernel and a
post-processing kernel. The lengths are 2^15, 2^16 and 2^20 (split), 1009
and 10007 (Bluestein), and the real kinds' lengths that lead there. They
run in local plans and in route plans, in every operation with a
reference counterpart, against the oracle on the same seeded inputs
(relative L2: 2e-6 f32, 1e-12 f64). At full size, the round trip also
returns the input.
"""
import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import complex_of, rel_l2, tol, uniform_complex, uniform_real

pytestmark = pytest.mark.gpu

DTYPES = [np.float32, np.float64]


def _tdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [32768, 65536, 1009, 10007, 1 << 20])
def test_local_c2c_long_lines(gpu, n, dt):
    rng = np.random.default_rng(n + 7)
    cd = complex_of(dt)
    batch = 3 if n < (1 << 20) else 1
    x = uniform_complex(rng, (batch, n), cd)
    plan = pf.LocalC2cPlan(n, dtype=torch.from_numpy(x).dtype)
    xd = _tdev(x)
    before = xd.clone()
    for op, fn in ((O.FORWARD, plan.forward), (O.BACKWARD, plan.backward), (O.INVERSE, plan.inverse)):
        got = _host(fn(xd))
        ref = O.c2c_lines(x.astype(np.complex128), op)
        err = rel_l2(got, ref)
        print(f"c2c n={n} {np.dtype(dt).name} op={op} rel_l2={err:.3g}")
        assert err <= tol(dt), (n, op, err)
    assert torch.equal(xd, before), "source must be preserved"
    back = _host(plan.inverse(plan.forward(xd)))
    assert rel_l2(back, x) <= tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [65536, 2018, 10007, 131072])
def test_local_r2c_c2r_long_lines(gpu, n, dt):
    """Even lengths whose half-length FFT is long (2^15, 2^16) or prime
    (1009), and an odd prime length: R2C forward, C2R inverse and backward."""
    rng = np.random.default_rng(n + 11)
    x = uniform_real(rng, (2, n), dt)
    plan = pf.LocalR2cPlan(n, dtype=torch.from_numpy(x).dtype)
    spec = plan.forward(_tdev(x))
    ref = O.r2c_lines(x.astype(np.float64))
    err = rel_l2(_host(spec), ref)
    print(f"r2c n={n} {np.dtype(dt).name} rel_l2={err:.3g}")
    assert err <= tol(dt), (n, err)
    # DC and Nyquist come out exactly real (the C2R preflight of the
    # reference, r2c.rs:798-826, rejects any other value there)
    hs = _host(spec).reshape(x.shape[0], -1)
    assert np.all(hs[:, 0].imag == 0) and (n % 2 or np.all(hs[:, n // 2].imag == 0))
    s = ref.astype(complex_of(dt))
    s[:, 0] = s[:, 0].real
    if n % 2 == 0:
        s[:, n // 2] = s[:, n // 2].real
    for op, fn in ((O.INVERSE, plan.inverse), (O.BACKWARD, plan.backward)):
        got = _host(fn(_tdev(s)))
        want = O.c2r_lines(s.astype(np.complex128), n, op)
        err = rel_l2(got, want)
        print(f"c2r n={n} {np.dtype(dt).name} op={op} rel_l2={err:.3g}")
        assert err <= tol(dt), (n, op, err)
    back = _host(plan.inverse(spec))
    assert rel_l2(back, x) <= tol(dt)


R2R_KINDS = [O.DCT1, O.DCT2, O.DCT3, O.DCT4, O.DST1, O.DST2, O.DST3, O.DST4]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [10007, 65536])
def test_local_r2r_and_dht_long_lines(gpu, n, dt):
    """All eight kinds and the DHT on a prime length (odd-length pieces,
    Bluestein) and on 2^16 (split): forward against the oracle, the
    normalized inverse back to the input, and the adjoint of the forward
    (its edge weights) against the oracle."""
    rng = np.random.default_rng(n + 13)
    x = uniform_real(rng, (2, n), dt)
    xd = _tdev(x)
    for kind in R2R_KINDS + [O.DHT]:
        if kind == O.DHT:
            plan = pf.LocalDhtPlan(n, dtype=torch.from_numpy(x).dtype)
        else:
            plan = pf.LocalR2rPlan(n, kind, dtype=torch.from_numpy(x).dtype)
        got = _host(plan.forward(xd))
        ref = O.r2r_lines(x.astype(np.float64), kind, O.FORWARD)
        err = rel_l2(got, ref)
        print(f"r2r kind={kind} n={n} {np.dtype(dt).name} rel_l2={err:.3g}")
        assert err <= tol(dt), (kind, n, err)
        back = _host(plan.inverse(plan.forward(xd)))
        assert rel_l2(back, x) <= tol(dt), (kind, n)
        adj = _host(plan.adjoint(O.ADJ_FORWARD, xd))
        ref = O.r2r_lines(x.astype(np.float64), kind, O.ADJ_FORWARD)
        assert rel_l2(adj, ref) <= tol(dt), (kind, n, "adjoint")


ROUTE_C2C = [((10007, 3, 4), 2), ((4, 3, 65536), 1), ((3, 1009, 5), 1), ((6, 32768, 2), 1)]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape,extra", ROUTE_C2C)
def test_route_c2c_long_axis(gpu, shape, extra, dt):
    """C2cPlan with one long axis (first, middle or last memory axis):
    forward, inverse, backward and the adjoint of the forward."""
    rng = np.random.default_rng(sum(shape) + extra)
    cd = complex_of(dt)
    x = uniform_complex(rng, (extra,) + shape, cd).reshape(-1)
    plan = pf.C2cPlan(shape, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else ())
    y = O.transform(O.KIND_C2C, x.astype(np.complex128), O.FORWARD, shape=shape, extra=extra)
    got = _host(plan.forward(_tdev(x)))
    err = rel_l2(got, y)
    print(f"route c2c {shape} {np.dtype(dt).name} rel_l2={err:.3g}")
    assert err <= tol(dt), (shape, err)
    ys = y.astype(cd)
    for op in (O.INVERSE, O.BACKWARD, O.ADJ_FORWARD):
        got = _host(plan.execute(op, _tdev(ys)))
        ref = O.transform(O.KIND_C2C, ys.astype(np.complex128), op, shape=shape, extra=extra)
        err = rel_l2(got, ref)
        assert err <= tol(dt), (shape, op, err)


ROUTE_REAL = [
    (O.KIND_R2C, (4, 6, 2018), None),
    (O.KIND_R2C, (10007, 4, 3), None),
    (O.KIND_R2R, (10007, 3, 4), [O.DCT2, O.DST1, O.DCT4]),
    (O.KIND_R2R, (6, 65536, 2), [O.DCT4, O.DHT, O.DST2]),
]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind,shape,r2r", ROUTE_REAL)
def test_route_real_long_axis(gpu, kind, shape, r2r, dt):
    """R2cPlan and R2rPlan with a long axis: forward, inverse and backward
    against the oracle."""
    rng = np.random.default_rng(sum(shape) + kind)
    x = uniform_real(rng, shape, dt).reshape(-1)
    if kind == O.KIND_R2C:
        plan = pf.R2cPlan(shape, dtype=torch.from_numpy(x).dtype)
    else:
        plan = pf.R2rPlan(shape, r2r, dtype=torch.from_numpy(x).dtype)
    kw = dict(shape=shape) if r2r is None else dict(shape=shape, r2r=r2r)
    y = O.transform(kind, x.astype(np.float64), O.FORWARD, **kw)
    got = _host(plan.forward(_tdev(x)))
    err = rel_l2(got, y)
    print(f"route kind={kind} {shape} {np.dtype(dt).name} rel_l2={err:.3g}")
    assert err <= tol(dt), (shape, err)
    ys = y.astype(got.dtype)
    for op in (O.INVERSE, O.BACKWARD):
        got = _host(plan.execute(op, _tdev(ys)))
        ref = O.transform(kind, ys.astype(np.complex128 if np.iscomplexobj(ys) else np.float64), op, **kw)
        assert rel_l2(got, ref) <= tol(dt), (shape, op)


def test_in_place_long_lines(gpu):
    """In-place operations whose stages run as pieces (the piece regions
    follow the in-place buffer assignment) give the out-of-place results:
    a C2cPlan with a 2^16 axis, LocalC2cPlan at 2^16 and 10007, and the
    packed LocalR2cPlan at 2018."""
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    plan = pf.C2cPlan((4, 3, 65536), dtype=torch.complex64)
    x = torch.randn(plan.input_elements(), generator=g, device="cuda", dtype=torch.complex64)
    want = _host(plan.forward(x))
    data = x.clone()
    got = _host(plan.forward_in_place(data))
    assert rel_l2(got, want) <= tol(np.float32)
    for n in (65536, 10007):
        lp = pf.LocalC2cPlan(n, dtype=torch.complex64)
        x = torch.randn(3 * n, generator=g, device="cuda", dtype=torch.complex64)
        want = _host(lp.forward(x))
        got = _host(lp.forward_in_place(x.clone()))
        assert rel_l2(got, want) <= tol(np.float32), n
    rp = pf.LocalR2cPlan(2018, dtype=torch.float32)
    x = torch.randn(2 * 2018, generator=g, device="cuda", dtype=torch.float32)
    want = _host(rp.forward(x))
    # packed in-place storage (r2c.rs:532-623): the reals in, 2 x 1010
    # complex values out, through the workspace
    buf = torch.zeros(2 * 2 * 1010, device="cuda", dtype=torch.float32)
    buf[: x.numel()] = x
    # the workspace holds the piece regions, then a staged copy of the input
    ws = torch.zeros(max(rp.workspace_bytes(op) for op in range(6)) + 2 * 2018 * 4 + 256, dtype=torch.uint8,
                     device="cuda")
    got = _host(rp._run_in_place(pf.FORWARD, buf, x.numel(), ws)).view(np.complex64)
    assert rel_l2(got, want) <= tol(np.float32)


def test_long_axis_with_strided_lines(gpu):
    """permute_dims = false (lines strided in place, no transposes): a
    C2cPlan with a 2^16 axis and an R2rPlan with a 10007-point axis, forward
    and inverse against the oracle."""
    rng = np.random.default_rng(12)
    shape = (4, 3, 65536)
    x = uniform_complex(rng, shape, np.complex64).reshape(-1)
    plan = pf.C2cPlan(shape, dtype=torch.complex64, permute_dims=False)
    y = O.transform(O.KIND_C2C, x.astype(np.complex128), O.FORWARD, shape=shape, permute_dims=False)
    assert rel_l2(_host(plan.forward(_tdev(x))), y) <= tol(np.float32)
    back = O.transform(O.KIND_C2C, y, O.INVERSE, shape=shape, permute_dims=False)
    assert rel_l2(_host(plan.inverse(_tdev(y.astype(np.complex64)))), back) <= tol(np.float32)
    shape = (10007, 3, 2)
    kinds = [O.DST1, O.DCT2, O.DHT]
    x = uniform_real(rng, shape, np.float64).reshape(-1)
    plan = pf.R2rPlan(shape, kinds, dtype=torch.float64, permute_dims=False)
    y = O.transform(O.KIND_R2R, x, O.FORWARD, shape=shape, r2r=kinds, permute_dims=False)
    assert rel_l2(_host(plan.forward(_tdev(x))), y) <= tol(np.float64)
    back = O.transform(O.KIND_R2R, y, O.INVERSE, shape=shape, r2r=kinds, permute_dims=False)
    assert rel_l2(_host(plan.inverse(_tdev(y))), back) <= tol(np.float64)


def test_mixed_plans_with_a_long_axis(gpu):
    """MixedC2cPlan with a 10007-point DCT-II axis on complex data (real and
    imaginary parts as real lines, Bluestein pieces), and MixedR2cPlan whose
    RFFT axis is 2018 points (half-length 1009, prime): forward, inverse
    (the mixed endpoint check on the long C2R stage) and backward against the
    oracle."""
    F, R = pf.AxisTransform.FFT, pf.AxisTransform.RFFT
    rng = np.random.default_rng(14)
    shape, codes = (10007, 4, 3), [2, F, 9]
    x = uniform_complex(rng, shape, np.complex128).reshape(-1)
    plan = pf.MixedC2cPlan(shape, codes, dtype=torch.complex128)
    y = O.transform(O.KIND_MIXED_C2C, x, O.FORWARD, shape=shape, r2r=codes)
    assert rel_l2(_host(plan.forward(_tdev(x))), y) <= tol(np.float64)
    for op in (O.INVERSE, O.BACKWARD):
        ref = O.transform(O.KIND_MIXED_C2C, y, op, shape=shape, r2r=codes)
        assert rel_l2(_host(plan.execute(op, _tdev(y))), ref) <= tol(np.float64), op
    shape, codes = (4, 2018, 3), [F, R, 2]
    x = uniform_real(rng, shape, np.float32).reshape(-1)
    plan = pf.MixedR2cPlan(shape, codes, dtype=torch.float32)
    y = O.transform(O.KIND_MIXED_R2C, x.astype(np.float64), O.FORWARD, shape=shape, r2r=codes)
    got = _host(plan.forward(_tdev(x)))
    assert rel_l2(got, y) <= tol(np.float32)
    for op in (O.INVERSE, O.BACKWARD):
        ref = O.transform(O.KIND_MIXED_R2C, got.astype(np.complex128), op, shape=shape, r2r=codes)
        assert rel_l2(_host(plan.execute(op, _tdev(got))), ref) <= tol(np.float32), op


def test_timing_entries_of_a_long_stage(gpu):
    """*_with_timing (distributed.rs:290-325) on a plan with a long axis: one
    entry per route stage, the long stage's pieces summed into its entry."""
    plan = pf.C2cPlan((4, 3, 65536), dtype=torch.complex64)
    x = torch.randn(plan.input_elements(), dtype=torch.complex64, device="cuda")
    y, t = plan.forward_with_timing(x)
    entries = t.entries()
    assert len(entries) == 3 and all(ms > 0 for _, ms, _ in entries), entries
    assert rel_l2(_host(y), _host(plan.forward(x))) == 0.0


def test_local_long_line_chunks_reuse_their_plans(gpu):
    """A local plan runs long lines in chunks of about 256 MiB of piece
    regions (819 lines of 10007 points in f32) and plans a chunk's pieces
    once per (operation, line count). 830 lines (a full chunk and 11 lines),
    5 lines, the 830 lines again (every chunk plan reused) and 821 lines (a
    new short chunk) must give every line bit-exactly as the other calls do,
    since lines never interact; the first five lines match the oracle."""
    n = 10007
    rng = np.random.default_rng(2026)
    x = uniform_complex(rng, (830, n), complex_of(np.float32)).reshape(-1)
    plan = pf.LocalC2cPlan(n, dtype=torch.complex64)
    xd = _tdev(x)
    y830 = _host(plan.forward(xd))
    y5 = _host(plan.forward(xd[: 5 * n]))
    ref5 = O.c2c_lines(x[: 5 * n].reshape(5, n).astype(np.complex128), O.FORWARD).reshape(-1)
    assert rel_l2(y5, ref5) <= tol(np.float32)
    assert np.array_equal(y5, y830[: 5 * n])
    assert np.array_equal(_host(plan.forward(xd)), y830)
    assert np.array_equal(_host(plan.forward(xd[: 821 * n])), y830[: 821 * n])
