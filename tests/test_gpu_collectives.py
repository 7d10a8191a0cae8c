"""Device reductions and broadcast pointwise combinations against the oracle
(oracle/collectives.py, the reference's folds restated on the CPU).

Bars: integer sums, min / max, any / all and every non-finite / empty /
overflow outcome bit-exact; floating sums within the sequential fold's own
error bound of the oracle and within one rounding of the exact sum
(math.fsum); norms within 1e-12 (f64) or one f32 rounding; pointwise
combinations within 1e-6 (f32) / 1e-14 (f64) of the oracle's broadcast loop.
"""
import math

import numpy as np
import pytest
import torch

import collectives as C
import pencil_fft_hip as pf

pytestmark = pytest.mark.gpu

INT_TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]
FLOAT_TYPES = [np.float32, np.float64, np.complex64, np.complex128]


def dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def rand(rng, dt, n):
    dt = np.dtype(dt)
    if dt.kind in "iu":
        info = np.iinfo(dt)
        lo, hi = max(int(info.min), -1000), min(int(info.max), 1000)
        return rng.integers(lo, hi, size=n, endpoint=True).astype(dt)
    if dt.kind == "c":
        return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt)
    return rng.standard_normal(n).astype(dt)


@pytest.mark.parametrize("dt", INT_TYPES)
@pytest.mark.parametrize("n", [0, 1, 255, 100_003])
def test_integer_reductions(gpu, dt, n):
    rng = np.random.default_rng(n + np.dtype(dt).itemsize)
    x = rand(rng, dt, n)
    d = dev(x)
    try:
        ref = C.global_sum(x)
    except C.IntegerOverflow:
        with pytest.raises(pf.CollectiveError):
            pf.global_sum(d)
    else:
        assert pf.global_sum(d) == int(ref)
    assert pf.global_min(d) == (None if n == 0 else int(C.global_min(x)))
    assert pf.global_max(d) == (None if n == 0 else int(C.global_max(x)))
    assert pf.any(d) == C.any_(x) and pf.all(d) == C.all_(x)
    norm = pf.l2_norm(d)
    ref_norm = float(C.l2_norm(x))
    assert abs(norm - ref_norm) <= 1e-12 * ref_norm


@pytest.mark.parametrize("dt", [np.int8, np.uint8, np.int16, np.int32, np.int64, np.uint64])
def test_integer_overflow_follows_the_storage_order(gpu, dt):
    info = np.iinfo(dt)
    n = 100_000
    # large positive run then an equal negative run: the total (0 + small) fits,
    # the fold overflows in the middle; swapping the halves changes nothing
    if info.min < 0:
        big = int(info.max) // 3 + 1
        x = np.concatenate([np.full(n // 2, big, dtype=dt), np.full(n // 2, -big, dtype=dt)])
        with pytest.raises(pf.CollectiveError) as e:
            pf.global_sum(dev(x))
        assert e.value.variant == "IntegerOverflow"
        with pytest.raises(C.IntegerOverflow):
            C.global_sum(x[:8].copy())  # 3 * big > MAX: the oracle fails at the third value
        # interleaved, the prefix sums stay small
        y = np.empty(n, dtype=dt)
        y[0::2], y[1::2] = big, -big
        assert pf.global_sum(dev(y)) == 0
    else:
        n = 100 if int(info.max) < 1 << 20 else n  # keep the last value representable
        x = np.full(n, int(info.max) // (n - 1), dtype=dt)
        x[-1] = int(info.max) - int(x[0]) * (n - 1) + 1  # total = MAX + 1
        with pytest.raises(pf.CollectiveError):
            pf.global_sum(dev(x))
        x[-1] -= 1
        assert pf.global_sum(dev(x)) == int(info.max)
    # the reference's own cases: [127, 1, -1] i8 fails; 35 x i8::MAX fails
    if dt == np.int8:
        with pytest.raises(pf.CollectiveError):
            pf.global_sum(dev(np.array([127, 1, -1], dtype=np.int8)))
        assert pf.global_sum(dev(np.array([127, -1, 1], dtype=np.int8))) == 127
        with pytest.raises(pf.CollectiveError):
            pf.global_sum(dev(np.full(35, 127, dtype=np.int8)))
    if dt == np.uint8:
        assert pf.global_sum(dev(np.array([200], dtype=np.uint8))) == 200


def _fold_bound(x: np.ndarray) -> float:
    """Error bound of a sequential fold of n values in the element precision."""
    eps = np.finfo(np.float32 if x.dtype in (np.float32, np.complex64) else np.float64).eps
    mag = np.abs(x.real).astype(np.float64).sum() + (np.abs(x.imag).astype(np.float64).sum() if np.iscomplexobj(x) else 0)
    return x.size * eps * mag + 1e-300


@pytest.mark.parametrize("dt", FLOAT_TYPES)
@pytest.mark.parametrize("n", [0, 1, 1000, 100_003])
def test_float_reductions(gpu, dt, n):
    rng = np.random.default_rng(7 + n)
    x = rand(rng, dt, n)
    d = dev(x)
    s = complex(pf.global_sum(d))
    ref = complex(C.global_sum(x))
    exact = complex(math.fsum(x.real.astype(np.float64)),
                    math.fsum(x.imag.astype(np.float64)) if np.iscomplexobj(x) else 0.0)
    bound = _fold_bound(x)
    assert abs(s - ref) <= bound
    ulp = np.finfo(np.float32 if dt in (np.float32, np.complex64) else np.float64).eps
    assert abs(s.real - exact.real) <= ulp * abs(exact.real) + 1e-300
    assert abs(s.imag - exact.imag) <= ulp * abs(exact.imag) + 1e-300
    norm = pf.l2_norm(d)
    rn = float(C.l2_norm(x))
    tol = 1e-12 if dt in (np.float64, np.complex128) else 2e-7
    assert abs(norm - rn) <= tol * rn
    if not np.iscomplexobj(x):
        assert pf.global_min(d) == (None if n == 0 else float(C.global_min(x)))
        assert pf.global_max(d) == (None if n == 0 else float(C.global_max(x)))
    assert pf.any(d) == C.any_(x) and pf.all(d) == C.all_(x)


def test_reference_special_value_cases(gpu):
    """tests/collectives.rs:365-438 at one rank, plus NaN / infinity rules."""
    n = 35
    w = np.zeros(n, dtype=np.complex64)
    w[0] = complex(np.inf, -np.inf)
    r = pf.global_sum(dev(w))
    assert math.isinf(r.real) and r.real > 0 and math.isinf(r.imag) and r.imag < 0
    v = np.zeros(n, dtype=np.complex64)
    v[0] = complex(np.nan, 3.0)
    r = pf.global_sum(dev(v))
    assert math.isnan(r.real) and r.imag == 3.0
    o = np.full(n, complex(np.inf, 0.0), dtype=np.complex64)
    assert pf.global_sum(dev(o)).real == math.inf
    o[1] = complex(-np.inf, 0.0)
    assert math.isnan(pf.global_sum(dev(o)).real)
    big = np.full(n, complex(np.finfo(np.float32).max, np.finfo(np.float32).max), dtype=np.complex64)
    assert math.isinf(pf.l2_norm(dev(big)))
    # f32 values whose sum overflows f32 only in the final rounding
    assert pf.global_sum(dev(np.full(4, np.finfo(np.float32).max, dtype=np.float32))) == math.inf
    for dt in (np.float32, np.float64):
        x = np.array([1.0, np.nan, 2.0], dtype=dt)
        assert math.isnan(pf.global_sum(dev(x))) and math.isnan(pf.global_min(dev(x)))
        assert math.isnan(pf.global_max(dev(x))) and math.isnan(pf.l2_norm(dev(x)))
        x = np.array([1.0, -np.inf, 2.0], dtype=dt)
        assert pf.global_sum(dev(x)) == -math.inf and pf.global_min(dev(x)) == -math.inf
        assert pf.global_max(dev(x)) == 2.0 and pf.l2_norm(dev(x)) == math.inf
    # NormState known answers (src/collectives.rs:2937-3019)
    assert pf.l2_norm(dev(np.array([3e300, 4e300]))) == pytest.approx(5e300, rel=1e-12)
    assert pf.l2_norm(dev(np.array([3e-300, 4e-300]))) == pytest.approx(5e-300, rel=1e-12)
    q = np.float64(5e-324)
    assert pf.l2_norm(dev(np.array([3 * q, 4 * q]))) == 5 * q
    assert pf.l2_norm(dev(np.array([0.0, -0.0]))) == 0.0
    assert pf.l2_norm(dev(np.array([3 + 4j]))) == 5.0
    assert pf.any(dev(np.array([0.0, np.nan]))) and pf.all(dev(np.array([1.0, np.nan])))
    assert not pf.all(dev(np.array([1 + 0j, 0j]))) and pf.any(dev(np.array([0j, 1j])))


def test_regular_array_under_a_permuted_pencil(gpu):
    """tests/collectives.rs:882-931 at one rank: the [5, 7] array with extras
    [2, 3] stored in the [1, 0] permutation's memory order."""
    exp = np.array([e0 * 100_000 + e1 * 10_000 + x * 100 + y
                    for e0 in range(2) for e1 in range(3) for x in range(5) for y in range(7)], dtype=np.int64)
    mem = np.ascontiguousarray(exp.reshape(2, 3, 5, 7).transpose(0, 1, 3, 2)).reshape(-1)
    d = dev(mem)
    assert pf.global_sum(d) == int(exp.sum())
    assert pf.global_min(d) == int(exp.min()) and pf.global_max(d) == int(exp.max())
    assert abs(pf.l2_norm(d) - math.sqrt(sum(float(v) ** 2 for v in exp))) < 1e-6
    e = dev(np.zeros(0, dtype=np.int64))
    assert pf.global_sum(e) == 0 and pf.global_min(e) is None and pf.global_max(e) is None
    assert pf.l2_norm(e) == 0.0 and pf.all(e) and not pf.any(e)


@pytest.mark.slow
def test_full_size_sums_against_the_exact_sum(gpu):
    """A 512^3 f32 field: the device sum against math.fsum (size independent:
    double-double accumulation rounds once), the norm against float64 torch."""
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    x = torch.randn(512 ** 3, generator=g, device="cuda", dtype=torch.float32)
    s = pf.global_sum(x)
    exact = math.fsum(x.double().cpu().numpy())
    assert abs(s - exact) <= np.finfo(np.float32).eps * abs(exact) + 1e-30
    n = pf.l2_norm(x)
    ref = float(torch.linalg.vector_norm(x.double()))
    assert abs(n - ref) <= 2e-7 * ref


# ---------------------------------------------------------- broadcast ops --
def _arr(rng, dt, layout, extra):
    n = layout.local_len() * int(np.prod(extra, dtype=np.int64)) if extra else layout.local_len()
    x = rand(rng, dt, n)
    return x, pf.PencilArray(dev(x), layout, extra)


@pytest.mark.parametrize("dt", FLOAT_TYPES)
def test_pointwise_linear_mul_muladd_against_the_oracle(gpu, dt):
    rng = np.random.default_rng(31)
    lay = pf.PencilLayout((6, 5, 4), perm=(2, 0, 1), decomposition=(1, 2))
    n = lay.local_len()
    shapes = [(2, 1, 3), (1, 4, 3), (2, 4, 1)]
    xs, arrs = zip(*[_arr(rng, dt, lay, s) for s in shapes])
    out_extra = (2, 4, 3)
    cplx = np.iscomplexobj(xs[0])
    odt = np.dtype(dt)
    tol = 1e-6 if odt in (np.float32, np.complex64) else 1e-14
    out = pf.PencilArray(torch.empty(n * 24, dtype=arrs[0].data.dtype, device="cuda"), lay, out_extra)
    coeffs = [0.5 + (0.25j if cplx else 0), -1.5, 2.0]
    pf.pointwise_linear(list(arrs), coeffs, out)
    ref = C.pointwise_many(list(xs), shapes, out_extra, n,
                           lambda v: coeffs[0] * v[0] + coeffs[1] * v[1] + coeffs[2] * v[2])
    got = out.data.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= tol * np.max(np.abs(ref))
    pf.pointwise_mul(arrs[0], arrs[1], out, coeff=2.0)
    ref = C.pointwise_many(list(xs[:2]), shapes[:2], out_extra, n, lambda v: 2.0 * v[0] * v[1])
    assert np.max(np.abs(out.data.cpu().numpy() - ref)) <= tol * np.max(np.abs(ref))
    pf.pointwise_mul_add(arrs[0], arrs[1], arrs[2], out, coeff_ab=-1.0, coeff_c=3.0)
    ref = C.pointwise_many(list(xs), shapes, out_extra, n, lambda v: -1.0 * v[0] * v[1] + 3.0 * v[2])
    assert np.max(np.abs(out.data.cpu().numpy() - ref)) <= tol * np.max(np.abs(ref))


def test_pointwise_in_place_real_times_complex_and_errors(gpu):
    rng = np.random.default_rng(5)
    lay = pf.PencilLayout((8, 6))
    n = lay.local_len()
    u, U = _arr(rng, np.complex128, lay, (3,))
    w, W = _arr(rng, np.float64, lay, (1,))
    before = U.data.clone()
    # in place: U = U + 0.5j * W (W broadcast over the extra axis)
    pf.pointwise_linear([U, W], [1.0, 0.5j], U)
    ref = C.pointwise_many([u, w], [(3,), (1,)], (3,), n, lambda v: v[0] + 0.5j * v[1])
    assert np.max(np.abs(U.data.cpu().numpy() - ref)) <= 1e-14 * np.max(np.abs(ref))
    # a rejected call leaves the output untouched
    keep = U.data.clone()
    other = pf.PencilArray(W.data, pf.PencilLayout((8, 6), perm=(1, 0)), (1,))
    with pytest.raises(pf.PfhError) as e:
        pf.pointwise_mul(U, other, U)
    assert e.value.variant == "IncompatiblePencils"
    bad = pf.PencilArray(torch.zeros(2 * n, dtype=torch.complex128, device="cuda"), lay, (2,))
    with pytest.raises(pf.PfhError) as e:
        pf.pointwise_mul(U, bad, U)
    assert e.value.variant == "ExtraExtentMismatch" and (e.value.index, e.value.required, e.value.actual) == (0, 3, 2)
    with pytest.raises(pf.PfhError) as e:
        pf.pointwise_linear([], [], U)
    assert e.value.variant == "EmptyInputs"
    torch.cuda.synchronize()
    assert torch.equal(U.data, keep)
    assert not torch.equal(U.data, before)


@pytest.mark.slow
def test_pointwise_full_size_matches_torch(gpu):
    """256^3 complex64 spectrum-sized arrays with extras [3]: the device
    combination against torch's own elementwise arithmetic."""
    lay = pf.PencilLayout((256, 256, 129), perm=(2, 1, 0))
    n = lay.local_len()
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    a = torch.randn(3 * n, generator=g, device="cuda", dtype=torch.complex64)
    b = torch.randn(n, generator=g, device="cuda", dtype=torch.float32)
    out = torch.empty_like(a)
    pf.pointwise_mul(pf.PencilArray(a, lay, (3,)), pf.PencilArray(b, lay, (1,)), pf.PencilArray(out, lay, (3,)),
                     coeff=1j)
    ref = (a.reshape(3, n) * b.reshape(1, n) * 1j).reshape(-1)
    assert float((out - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
