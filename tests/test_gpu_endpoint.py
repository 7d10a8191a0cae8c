"""GPU parity of the C2R boundary check at its decision boundary.

The device check (k_endpoint_partial + k_endpoint_combine: chunk maxima, then
chunk sums rescaled by (m_chunk / M)^2) computes the reference's
EndpointStatistics (distributed.rs:4205-4363) in another order, so its
decision is compared with the oracle's at the points where it can flip:
||Im|| at 0.9x / 1.1x of relative * ||Re|| (and within 1e-5 / 1e-9 of it where
the boundary data reaches the check unchanged), the subnormal absolute
threshold 128 * min_subnormal * D at exactly its value and one subnormal
above, the raw backward's x prod n_a, NaN / +-Inf in one constrained value,
one failing extra batch among several, and the MixedR2cPlan error model
(mixed.rs:400-427). Every rejected call must leave dst untouched.

The spectrum fed to the inverse is built from chosen boundary data B (the
array that reaches the C2R stage): S = the forward complex transforms of B
along the complex-suffix axes, so that the normalized inverse (or, scaled by
the transverse factors, the raw backward) brings B back. Constant boundary
planes come from a single spectral value, which power-of-two FFTs carry
through exactly, so subnormal thresholds can be hit to the last bit.
"""
import math

import numpy as np
import pytest
import scipy.fft as sfft
import torch

import oracle as O
import pencil_fft_hip as pf

pytestmark = pytest.mark.gpu

MINSUB = {np.float32: 2.0 ** -149, np.float64: 2.0 ** -1074}
EPS = {np.float32: 2.0 ** -23, np.float64: 2.0 ** -52}


def _cd(dt):
    return np.complex64 if dt == np.float32 else np.complex128


def _tdt(dt):
    return torch.float32 if dt == np.float32 else torch.float64


def _to_memory(S_log, plan):
    """[extra][logical axes] -> the plan's output memory order, flat."""
    perm = plan.layout(1)[0]
    return np.ascontiguousarray(np.transpose(S_log, (0,) + tuple(1 + p for p in perm))).reshape(-1)


def _decide(plan, kind, S_mem, op, shape, extra, dt, **kw):
    """(device accepted, oracle accepted); a device rejection must leave dst
    untouched and raise InvalidSpectrum."""
    n_real = int(np.prod(shape)) * extra
    dst = torch.full((n_real,), -1.0, dtype=_tdt(dt), device="cuda")
    src = torch.from_numpy(np.ascontiguousarray(S_mem)).cuda()
    try:
        plan.execute(op, src, dst)
        torch.cuda.synchronize()
        dev = True
    except pf.PfhError as e:
        assert e.variant == "InvalidSpectrum", e
        torch.cuda.synchronize()
        assert torch.all(dst == -1.0), "a rejected spectrum wrote dst"
        dev = False
    try:
        O.transform(kind, S_mem, op, shape=shape, extra=extra, **kw)
        orc = True
    except O.InvalidSpectrum:
        orc = False
    return dev, orc


def _boundary(rng, extra, n0, n1, nc, dt, dc_ratio, ny_ratio, relative):
    """Boundary data [extra][n0][n1][nc]: DC and Nyquist planes with
    ||Im|| = ratio * relative * ||Re|| per batch, interior bins random."""
    B = rng.standard_normal((extra, n0, n1, nc)) + 1j * rng.standard_normal((extra, n0, n1, nc))
    for e in range(extra):
        for k, ratios in ((0, dc_ratio), (nc - 1, ny_ratio)):
            re = rng.standard_normal((n0, n1)) * 10.0 ** rng.uniform(-3, 3, size=(n0, 1))
            im = rng.standard_normal((n0, n1))
            im *= ratios[e] * relative * np.linalg.norm(re) / np.linalg.norm(im)
            B[e, :, :, k] = re + 1j * im
    return B


# (DC ratios per batch, Nyquist ratios per batch, accepted)
RATIO_CASES = [
    ((0.9, 0.9, 0.9), (0.9, 0.9, 0.9), True),
    ((0.0, 0.5, 0.9), (0.9, 0.0, 0.5), True),
    ((0.5, 1.1, 0.5), (0.5, 0.5, 0.5), False),  # one batch's DC plane fails
    ((0.5, 0.5, 0.5), (0.5, 0.5, 1.1), False),  # one batch's Nyquist plane fails
    ((1.1, 1.1, 1.1), (0.9, 0.9, 0.9), False),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dc,ny,accepted", RATIO_CASES)
def test_relative_branch_near_the_threshold(gpu, dc, ny, accepted, dt, permute):
    shape, extra = (8, 12, 16), 3
    n0, n1, n2 = shape
    nc = n2 // 2 + 1
    rel, _, depth = O.endpoint_thresholds_c(O.KIND_R2C, dt, shape, O.INVERSE)
    assert depth == 8  # 1 + ceil_log2(8) + ceil_log2(12)
    rng = np.random.default_rng(hash((dc, ny)) % 2 ** 32)
    B = _boundary(rng, extra, n0, n1, nc, dt, dc, ny, rel)
    S = np.fft.fft2(B, axes=(1, 2))
    plan = pf.R2cPlan(shape, dtype=_tdt(dt), extra_shape=(extra,), permute_dims=permute)
    S_mem = _to_memory(S, plan).astype(_cd(dt))
    dev, orc = _decide(plan, O.KIND_R2C, S_mem, pf.INVERSE, shape, extra, dt, permute_dims=permute)
    assert dev == orc == accepted
    # the raw backward checks the same planes with the same relative threshold
    S_raw = _to_memory(S / (n0 * n1), plan).astype(_cd(dt))
    dev, orc = _decide(plan, O.KIND_R2C, S_raw, pf.BACKWARD, shape, extra, dt, permute_dims=permute)
    assert dev == orc == accepted


def _spike_spectrum(shape, extra, dt, t_dc, t_ny, scale):
    """Spectrum whose boundary DC / Nyquist planes are the constants i*t
    (Re = 0) after the complex suffix: one value scale * i * t at
    (k0, k1) = (0, 0) of each plane, zero elsewhere on the planes; interior
    bins random."""
    n0, n1, n2 = shape
    nc = n2 // 2 + 1
    rng = np.random.default_rng(9)
    S = (rng.standard_normal((extra, n0, n1, nc)) + 1j * rng.standard_normal((extra, n0, n1, nc))).astype(_cd(dt))
    for k, t in ((0, t_dc), (nc - 1, t_ny)):
        S[:, :, :, k] = 0
        S[:, 0, 0, k] = 1j * (dt(scale) * dt(t))
        assert S[0, 0, 0, k].imag == scale * t  # exact in the element type
    return S


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_absolute_branch_at_the_exact_subnormal_threshold(gpu, dt):
    """absolute = 128 * min_subnormal * D (x n0 * n1 for the raw backward):
    a boundary plane with Re = 0 and Im = absolute passes, Im = absolute plus
    one subnormal fails. D = 1 + 3 + 4 = 8 for 8 x 16 x 16."""
    shape, extra = (8, 16, 16), 2
    n0, n1, _ = shape
    ms = MINSUB[dt]
    rel, ab_inv, _ = O.endpoint_thresholds_c(O.KIND_R2C, dt, shape, O.INVERSE)
    _, ab_raw, _ = O.endpoint_thresholds_c(O.KIND_R2C, dt, shape, O.BACKWARD)
    assert ab_inv == 128 * 8 * ms and ab_raw == ab_inv * n0 * n1
    assert ab_raw < np.finfo(dt).tiny  # both thresholds are subnormal
    plan = pf.R2cPlan(shape, dtype=_tdt(dt), extra_shape=(extra,))
    cases = [
        # (op, t on DC, t on Nyquist, accepted)
        (pf.INVERSE, ab_inv, 0.0, True),
        (pf.INVERSE, ab_inv + ms, 0.0, False),
        (pf.INVERSE, 0.0, ab_inv, True),
        (pf.INVERSE, 0.0, ab_inv + ms, False),
        (pf.INVERSE, ab_raw, 0.0, False),  # the raw threshold is n0 * n1 times larger
        (pf.BACKWARD, ab_raw, ab_raw, True),
        (pf.BACKWARD, ab_raw + ms, 0.0, False),
        (pf.BACKWARD, 0.0, ab_raw + ms, False),
        (pf.BACKWARD, ab_inv * 2, ab_inv, True),
    ]
    for op, t_dc, t_ny, accepted in cases:
        scale = n0 * n1 if op == pf.INVERSE else 1
        S = _spike_spectrum(shape, extra, dt, t_dc, t_ny, scale)
        S_mem = _to_memory(S, plan)
        dev, orc = _decide(plan, O.KIND_R2C, S_mem, op, shape, extra, dt)
        assert dev == orc == accepted, (op, t_dc / ms, t_ny / ms, dev, orc)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_constrained_value_rejects(gpu, dt, bad):
    shape, extra = (8, 12, 16), 3
    rng = np.random.default_rng(17)
    x = rng.standard_normal((extra,) + shape)
    plan = pf.R2cPlan(shape, dtype=_tdt(dt), extra_shape=(extra,))
    S = np.asarray(O.transform(O.KIND_R2C, x.reshape(-1).astype(dt), O.FORWARD, shape=shape, extra=extra))
    Sl = S.reshape((extra,) + tuple(np.array(plan.layout(1)[1])[list(plan.layout(1)[0])]))
    # memory order [kz][ky][kx] (permute_dims): DC plane = Sl[:, 0], Nyquist = Sl[:, -1]
    assert plan.layout(1)[0] == (2, 1, 0)
    for plane, part in ((0, 1.0), (-1, 1j)):
        T = Sl.copy()
        T[1, plane, 3, 5] = (bad if part == 1.0 else complex(0.0, bad))
        dev, orc = _decide(plan, O.KIND_R2C, T.reshape(-1), pf.INVERSE, shape, extra, dt)
        assert not dev and not orc
    # an interior bin is not checked: accepted by both
    T = Sl.copy()
    T[0, 2, 3, 5] = bad
    dev, orc = _decide(plan, O.KIND_R2C, T.reshape(-1), pf.INVERSE, shape, extra, dt)
    assert dev and orc


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_relative_branch_to_the_last_digits(gpu, dt):
    """Selection [2]: no complex suffix (D = 1), so the spectrum is the
    boundary data itself and ||Im|| / ||Re|| can be set within 1e-5 (f32) /
    1e-9 (f64) of relative. Plane maxima differ by up to 1e20 between rows,
    and the 64 x 48 planes span several device chunks, so the rescaled chunk
    sums must reproduce the reference's globally scaled sums."""
    shape = (64, 48, 10)
    n0, n1, n2 = shape
    nc = n2 // 2 + 1
    rel, ab, depth = O.endpoint_thresholds_c(O.KIND_R2C, dt, shape, O.INVERSE, selected=[2])
    assert depth == 1
    plan = pf.R2cPlan(shape, dtype=_tdt(dt), selection=[2])
    margin = 1e-5 if dt == np.float32 else 1e-9
    rng = np.random.default_rng(23)
    seen = set()
    for case in range(12):
        B = rng.standard_normal((1, n0, n1, nc)) + 1j * rng.standard_normal((1, n0, n1, nc))
        for k in (0, nc - 1):
            re = rng.standard_normal((n0, n1)) * 10.0 ** rng.uniform(-10, 10, size=(n0, 1))
            im = rng.standard_normal((n0, n1)) * 10.0 ** rng.uniform(-2, 2, size=(n0, 1))
            want_pass = case % 2 == 0 if k == 0 else True
            target = rel * (1 - margin if want_pass else 1 + margin)
            # the ratio is measured on the values as stored in the element type
            for _ in range(3):
                im *= target * np.linalg.norm(re.astype(dt).astype(np.float64)) / np.linalg.norm(
                    im.astype(dt).astype(np.float64))
            B[0, :, :, k] = re + 1j * im
        B = B.astype(_cd(dt))
        S_mem = _to_memory(B, plan)
        dev, orc = _decide(plan, O.KIND_R2C, S_mem, pf.INVERSE, shape, 1, dt, selected=[2])
        assert dev == orc == (case % 2 == 0), (case, dev, orc)
        seen.add(dev)
    assert seen == {True, False}


def _mixed_forward_suffix(B, n0, codes):
    """Forward complex suffix of the MixedR2cPlan (DCT-II on axis 0 applied to
    re and im separately, unnormalized FFTW REDFT10; FFT on axis 1)."""
    assert codes[:2] == (O.DCT2, O.AXIS_FFT)
    X = sfft.dct(B.real, type=2, axis=1) + 1j * sfft.dct(B.imag, type=2, axis=1)
    return np.fft.fft(X, axis=2)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_mixed_r2c_thresholds(gpu, dt):
    """MixedR2cPlan (DCT-II, FFT, RFFT) on 8 x 12 x 16: D = 1 + (5 + 2) +
    (4 + 0) = 12, raw factor 2 * 8 * 12 = 192 (mixed.rs:400-427). Relative
    branch at 0.9x / 1.1x; absolute branch bracketed between the thresholds
    a slip in D or in the factor would give."""
    shape, extra = (8, 12, 16), 2
    codes = (O.DCT2, O.AXIS_FFT, O.AXIS_RFFT)
    n0, n1, n2 = shape
    nc = n2 // 2 + 1
    rel, ab_inv, depth = O.endpoint_thresholds_c(O.KIND_MIXED_R2C, dt, shape, O.INVERSE, r2r=codes)
    _, ab_raw, _ = O.endpoint_thresholds_c(O.KIND_MIXED_R2C, dt, shape, O.BACKWARD, r2r=codes)
    ms = MINSUB[dt]
    assert depth == 12 and ab_inv == 1536 * ms and ab_raw == 1536 * 192 * ms
    plan = pf.MixedR2cPlan(shape, list(codes), dtype=_tdt(dt), extra_shape=(extra,))
    kw = dict(r2r=codes)
    rng = np.random.default_rng(31)
    for dc, ny, accepted in (((0.9, 0.9), (0.9, 0.9), True), ((0.9, 1.1), (0.9, 0.9), False),
                             ((0.5, 0.5), (1.1, 0.5), False)):
        B = _boundary(rng, extra, n0, n1, nc, dt, dc, ny, rel)
        S = _mixed_forward_suffix(B, n0, codes)
        dev, orc = _decide(plan, O.KIND_MIXED_R2C, _to_memory(S, plan).astype(_cd(dt)), pf.INVERSE, shape, extra,
                           dt, **kw)
        assert dev == orc == accepted, (dc, ny)
        S_raw = S / (2 * n0 * n1)
        dev, orc = _decide(plan, O.KIND_MIXED_R2C, _to_memory(S_raw, plan).astype(_cd(dt)), pf.BACKWARD, shape,
                           extra, dt, **kw)
        assert dev == orc == accepted, (dc, ny, "raw")
    # absolute branch, Re = 0: D = 12 gives 1536 subnormals; D without the
    # allowance or with n for E_a gives 1280, with 8n for E_a 1664
    for op, t, accepted in ((pf.INVERSE, 1450, True), (pf.INVERSE, 1600, False),
                            (pf.BACKWARD, 280000, True), (pf.BACKWARD, 310000, False)):
        scale = 2 * n0 * n1 if op == pf.INVERSE else 1
        S = _spike_spectrum(shape, extra, dt, t * ms, 0.0, scale)
        dev, orc = _decide(plan, O.KIND_MIXED_R2C, _to_memory(S, plan), op, shape, extra, dt, **kw)
        assert dev == orc == accepted, (op, t, dev, orc)


def test_c2r_plan_forward_threshold(gpu):
    """C2rPlan forward is the raw C2R: its absolute threshold carries the
    transverse factor (reduction axis 0 of 16 x 8 x 8: D = 1 + 3 + 3 = 7,
    factor 64)."""
    dt = np.float64
    shape = (16, 8, 8)
    rel, ab_raw, depth = O.endpoint_thresholds_c(O.KIND_C2R, dt, shape, O.BACKWARD)
    assert depth == 7 and ab_raw == 128 * 7 * 64 * MINSUB[dt]
    plan = pf.C2rPlan(shape, dtype=torch.float64)
    perm, lshape, _, _ = plan.layout(0)
    rng = np.random.default_rng(5)
    nc = shape[0] // 2 + 1
    for t, accepted in ((ab_raw, True), (ab_raw + MINSUB[dt], False)):
        # complex canonical layout [k0][x1][x2] after the suffix; spectral spike
        # at (k1, k2) = (0, 0) of the k0 = 0 and k0 = nc - 1 planes
        S = (rng.standard_normal((nc, 8, 8)) + 1j * rng.standard_normal((nc, 8, 8)))
        for k in (0, nc - 1):
            S[k] = 0
            S[k, 0, 0] = 1j * t
        S_mem = np.ascontiguousarray(np.transpose(S[None], (0,) + tuple(1 + p for p in perm))).reshape(-1)
        dst = torch.full((int(np.prod(shape)),), -1.0, dtype=torch.float64, device="cuda")
        try:
            plan.forward(torch.from_numpy(S_mem).cuda(), dst)
            torch.cuda.synchronize()
            dev = True
        except pf.PfhError as e:
            assert e.variant == "InvalidSpectrum"
            assert torch.all(dst == -1.0)
            dev = False
        try:
            O.transform(O.KIND_C2R, S_mem, O.FORWARD, shape=shape)
            orc = True
        except O.InvalidSpectrum:
            orc = False
        assert dev == orc == accepted, (t / MINSUB[dt], dev, orc)
