"""GPU parity of the local line plans against the oracle.

Mirrors the reference's local tests: lengths 1..13 with batches 1 and 3
(crates/pencil-fft/src/lib.rs:900-914), the peak / DC known answers
(lib.rs:916-966, r2c.rs:1195-1245), the endpoint preflight (r2c.rs:798-826),
all eight R2R kinds (r2r.rs:1064-1112) and the DHT (dht.rs:434-449, n = 1
identity dht.rs:215-224).
"""
import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import complex_of, real_of, rel_l2, tol, uniform_complex, uniform_real

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 4, 5, 7, 8, 9, 11, 12, 13, 16, 48, 60, 64, 84, 96, 112, 128, 160, 240, 256, 384, 385, 512]
DTYPES = [np.float32, np.float64]


def _tdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def test_first_plan_leaves_no_pending_device_error(gpu):
    """Device set-up (LDS limits of every kernel) happens on the first plan;
    a failed attribute call must not surface in the caller's next HIP call."""
    pf.LocalC2cPlan(64, dtype=torch.complex64)
    t = torch.ones(16, device="cuda") * 2
    torch.cuda.synchronize()
    assert float(t.sum().item()) == 32.0


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", LENGTHS)
def test_local_c2c_matches_oracle(gpu, n, dt):
    rng = np.random.default_rng(n)
    cd = complex_of(dt)
    for batch in (1, 3):
        x = uniform_complex(rng, (batch, n), cd)
        plan = pf.LocalC2cPlan(n, dtype=torch.from_numpy(x).dtype)
        xd = _tdev(x)
        before = xd.clone()
        for op, fn in ((O.FORWARD, plan.forward), (O.BACKWARD, plan.backward), (O.INVERSE, plan.inverse)):
            got = _host(fn(xd))
            ref = O.c2c_lines(x.astype(np.complex128), op)
            assert rel_l2(got, ref) <= tol(dt), (n, batch, op, rel_l2(got, ref))
        assert torch.equal(xd, before), "source must be preserved"


@pytest.mark.parametrize("dt", DTYPES)
def test_local_c2c_peak_and_dc(gpu, dt):
    n, peak = 5, 2
    cd = complex_of(dt)
    j = np.arange(n)
    x = np.exp(2j * np.pi * peak * j / n).astype(cd)
    got = _host(pf.LocalC2cPlan(n, dtype=torch.from_numpy(x).dtype).forward(_tdev(x)))
    exp = np.zeros(n, dtype=np.complex128)
    exp[peak] = n
    assert np.abs(got - exp).max() <= 1e-5 * n
    dc = np.full(n, 2.0 - 3.0j, dtype=cd)
    got = _host(pf.LocalC2cPlan(n, dtype=torch.from_numpy(dc).dtype).forward(_tdev(dc)))
    exp = np.zeros(n, dtype=np.complex128)
    exp[0] = (2.0 - 3.0j) * n
    assert np.abs(got - exp).max() <= 1e-5 * n


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", LENGTHS)
def test_local_r2c_matches_oracle(gpu, n, dt):
    rng = np.random.default_rng(100 + n)
    for batch in (1, 3):
        x = uniform_real(rng, (batch, n), dt)
        plan = pf.LocalR2cPlan(n, dtype=torch.from_numpy(x).dtype)
        spec = plan.forward(_tdev(x))
        got = _host(spec)
        ref = O.r2c_lines(x.astype(np.float64))
        assert rel_l2(got, ref) <= tol(dt), (n, batch, rel_l2(got, ref))
        # inverse / backward of the exact spectrum (endpoints made strictly real)
        s = ref.astype(complex_of(dt))
        s[:, 0] = s[:, 0].real
        if n % 2 == 0:
            s[:, -1] = s[:, -1].real
        inv = _host(plan.inverse(_tdev(s)))
        bwd = _host(plan.backward(_tdev(s)))
        ref_inv = O.c2r_lines(s.astype(np.complex128), n, O.INVERSE)
        ref_bwd = O.c2r_lines(s.astype(np.complex128), n, O.BACKWARD)
        assert rel_l2(inv, ref_inv) <= tol(dt)
        assert rel_l2(bwd, ref_bwd) <= tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
def test_local_r2c_frequency_and_normalization(gpu, dt):
    n, peak = 8, 2
    x = np.cos(2 * np.pi * peak * np.arange(n) / n).astype(dt)
    plan = pf.LocalR2cPlan(n, dtype=torch.from_numpy(x).dtype)
    spec = _host(plan.forward(_tdev(x)))
    exp = np.zeros(n // 2 + 1)
    exp[peak] = n / 2
    assert np.abs(spec.real - exp).max() <= 1e-5 * n and np.abs(spec.imag).max() <= 1e-5 * n
    dc = np.full(n, 3.5, dtype=dt)
    sd = plan.forward(_tdev(dc))
    assert abs(_host(sd)[0].real - 3.5 * n) < 1e-4
    back = _host(plan.inverse(sd))
    assert np.abs(back - dc).max() < 1e-5
    raw = _host(plan.backward(sd))
    assert np.abs(raw - dc * n).max() < 1e-4


@pytest.mark.parametrize("n", [6, 7])
def test_local_c2r_rejects_endpoints_before_writing(gpu, n):
    plan = pf.LocalR2cPlan(n, dtype=torch.float64)
    c = n // 2 + 1
    spec = np.zeros((3, c), dtype=np.complex128)
    spec[:, 1] = 1.0 + 2.0j
    spec[1, 0] = 0.0 + 1e-300j  # DC imaginary part on batch 1
    if n % 2 == 0:
        spec[2, -1] = 0.0 + 1.0j  # Nyquist on batch 2
    dst = torch.full((3 * n,), 7.0, dtype=torch.float64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.inverse(_tdev(spec), dst)
    assert e.value.variant == "InvalidSpectrumEndpoint"
    assert (e.value.batch, e.value.index) == (1, 0)
    assert torch.all(dst == 7.0), "rejected call must not write the destination"
    spec[1, 0] = -0.0 + 0.0j  # signed zero is accepted
    if n % 2 == 0:
        with pytest.raises(pf.PfhError) as e:
            plan.backward(_tdev(spec), dst)
        assert (e.value.batch, e.value.index) == (2, c - 1)
    else:
        plan.backward(_tdev(spec), dst)


def test_local_r2c_batch_errors(gpu):
    plan = pf.LocalR2cPlan(8, dtype=torch.float32)
    with pytest.raises(pf.PfhError) as e:
        plan.forward(torch.zeros(16, device="cuda"), torch.zeros(15, dtype=torch.complex64, device="cuda"))
    assert e.value.variant == "BatchCountMismatch" and (e.value.batch, e.value.index) == (2, 3)
    with pytest.raises(pf.PfhError) as e:
        plan.forward(torch.zeros(12, device="cuda"), torch.zeros(10, dtype=torch.complex64, device="cuda"))
    assert e.value.variant == "NonIntegralBatch"
    with pytest.raises(pf.PfhError) as e:
        pf.LocalR2cPlan(0)
    assert e.value.variant == "InvalidLength"


R2R_KINDS = [pf.R2rKind.DctI, pf.R2rKind.DctII, pf.R2rKind.DctIII, pf.R2rKind.DctIV, pf.R2rKind.DstI,
             pf.R2rKind.DstII, pf.R2rKind.DstIII, pf.R2rKind.DstIV]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("kind", R2R_KINDS)
def test_local_r2r_matches_oracle(gpu, kind, dt):
    rng = np.random.default_rng(kind)
    for n in [1, 2, 3, 4, 5, 6, 7, 8, 12, 13, 16, 60, 96, 240, 256, 384, 385, 386]:
        if kind == pf.R2rKind.DctI and n == 1:
            with pytest.raises(pf.PfhError):
                pf.LocalR2rPlan(1, kind)
            continue
        x = uniform_real(rng, (3, n), dt)
        plan = pf.LocalR2rPlan(n, kind, dtype=torch.from_numpy(x).dtype)
        xd = _tdev(x)
        for op, fn in ((O.FORWARD, plan.forward), (O.BACKWARD, plan.backward), (O.INVERSE, plan.inverse)):
            got = _host(fn(xd))
            ref = O.r2r_lines(x.astype(np.float64), kind, op)
            assert rel_l2(got, ref) <= 4 * tol(dt), (kind, n, op, rel_l2(got, ref))


@pytest.mark.parametrize("dt", DTYPES)
def test_local_dht_matches_oracle(gpu, dt):
    rng = np.random.default_rng(9)
    for n in [2, 3, 4, 5, 8, 12, 13, 60, 240, 256, 384, 385]:
        x = uniform_real(rng, (3, n), dt)
        plan = pf.LocalDhtPlan(n, dtype=torch.from_numpy(x).dtype)
        xd = _tdev(x)
        for op, fn in ((O.FORWARD, plan.forward), (O.INVERSE, plan.inverse)):
            got = _host(fn(xd))
            ref = O.r2r_lines(x.astype(np.float64), O.DHT, op)
            assert rel_l2(got, ref) <= 4 * tol(dt), (n, op, rel_l2(got, ref))


def test_local_dht_length_one_is_bit_exact_identity(gpu):
    x = np.array([-0.0, np.nan, 1.5, -3.25], dtype=np.float64)
    plan = pf.LocalDhtPlan(1, dtype=torch.float64)
    for fn in (plan.forward, plan.inverse, plan.backward):
        got = _host(fn(_tdev(x)))
        assert got.tobytes() == x.tobytes()


@pytest.mark.parametrize("dt", DTYPES)
def test_local_in_place_matches_out_of_place(gpu, dt):
    rng = np.random.default_rng(5)
    n = 60
    x = uniform_complex(rng, (5, n), complex_of(dt))
    plan = pf.LocalC2cPlan(n, dtype=torch.from_numpy(x).dtype)
    ref = _host(plan.forward(_tdev(x)))
    d = _tdev(x)
    plan.forward_in_place(d)
    assert np.array_equal(_host(d), ref)
    xr = uniform_real(rng, (5, n), dt)
    rp = pf.LocalR2rPlan(n, pf.R2rKind.DctII, dtype=torch.from_numpy(xr).dtype)
    ref = _host(rp.forward(_tdev(xr)))
    d = _tdev(xr)
    rp.forward_in_place(d)
    assert np.array_equal(_host(d), ref)
