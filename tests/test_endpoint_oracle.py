"""Pins the oracle's C2R boundary check (CPU): its thresholds against the
reference's closed forms, written here as literal numbers (normalization_depth
/ raw_absolute_threshold_for_shape, distributed/r2c.rs:4841-4878; the
MixedR2cPlan error model, distributed/mixed.rs:400-427, :4085-4117), the
Python and C restatements against each other, and its accept / reject
decisions at the exact threshold and against a direct restatement of
EndpointStatistics (distributed.rs:4205-4363)."""
import math

import numpy as np
import pytest

import oracle as O

F32_EPS, F64_EPS = 2.0 ** -23, 2.0 ** -52          # f32::EPSILON, f64::EPSILON
F32_MINSUB, F64_MINSUB = 2.0 ** -149, 2.0 ** -1074  # f32::from_bits(1), f64::from_bits(1)


def _both(kind, dt, shape, op, selected=None, r2r=None):
    """(relative, absolute) of the C and the Python restatements."""
    rel_c, abs_c, _ = O.endpoint_thresholds_c(kind, dt, shape, op, selected=selected, r2r=r2r)
    if kind == O.KIND_MIXED_R2C:
        ra = list(r2r).index(O.AXIS_RFFT)
        rel_p, abs_p = O.endpoint_thresholds(dt, shape, None, ra, op == O.INVERSE, mixed=r2r)
    else:
        sel = list(range(len(shape))) if selected is None else list(selected)
        ra = max(sel) if kind == O.KIND_R2C else min(sel)
        rel_p, abs_p = O.endpoint_thresholds(dt, shape, sel, ra, op == O.INVERSE)
    assert rel_c == rel_p and abs_c == abs_p, (rel_c, rel_p, abs_c, abs_p)
    return rel_c, abs_c


# (kind, shape, selection, depth D, transverse prod n_a) worked out by hand
# from the reference's formulas: D = 1 + sum ceil_log2(n_a) over the selected
# non-reduction axes with n_a > 1; the raw backward threshold multiplies by
# prod n_a over the same axes (extent 1 included).
CLOSED_FORMS = [
    (O.KIND_R2C, (512, 512, 512), None, 19, 512 * 512),  # 1 + 9 + 9 (the 512^3 workload)
    (O.KIND_R2C, (4, 6, 8), None, 6, 24),                 # 1 + 2 + 3
    (O.KIND_R2C, (5, 3, 8), None, 6, 15),                 # 1 + 3 + 2
    (O.KIND_R2C, (1, 6, 8), None, 4, 6),                  # extent 1: ceil_log2 0, factor 1
    (O.KIND_R2C, (2, 9, 8), None, 6, 18),                 # 1 + 1 + 4
    (O.KIND_R2C, (4, 6, 8), [0, 2], 3, 4),                # only axis 0 besides the reduction axis 2
    (O.KIND_R2C, (4, 6, 8), [0, 1], 3, 4),                # reduction axis 1 (largest selected)
    (O.KIND_R2C, (7, 6, 8), [2], 1, 1),                   # no other selected axis
    (O.KIND_C2R, (6, 8, 12), None, 8, 96),                # reduction axis 0: 1 + 3 + 4
    (O.KIND_R2C, (3, 5, 7, 8), None, 9, 105),             # 1 + 2 + 3 + 3 (4D)
]


@pytest.mark.parametrize("kind,shape,sel,depth,transverse", CLOSED_FORMS)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_thresholds_match_the_closed_forms(kind, shape, sel, depth, transverse, dt):
    eps, minsub = (F32_EPS, F32_MINSUB) if dt == np.float32 else (F64_EPS, F64_MINSUB)
    rel, ab = _both(kind, dt, shape, O.INVERSE, selected=sel)
    assert rel == 128.0 * eps * depth
    # 128 * min_subnormal * D is exact in f64 (a multiple of 2^-149 / 2^-1074)
    assert ab == math.ldexp(128 * depth, -149 if dt == np.float32 else -1074)
    rel_b, ab_b = _both(kind, dt, shape, O.BACKWARD, selected=sel)
    assert rel_b == rel
    assert ab_b == ab * transverse
    _, _, d = O.endpoint_thresholds_c(kind, dt, shape, O.INVERSE, selected=sel)
    assert d == depth


def test_thresholds_of_the_512_cube_workload():
    """The numbers the north-star workload runs with: relative 2.9e-4 (f32)
    and 5.4e-13 (f64) for D = 19; the absolute thresholds are subnormal
    multiples (a FLT_MIN / DBL_MIN slip would be 2^126 / 2^1022 times larger)."""
    r32, a32 = _both(O.KIND_R2C, np.float32, (512, 512, 512), O.INVERSE)
    r64, a64 = _both(O.KIND_R2C, np.float64, (512, 512, 512), O.INVERSE)
    assert r32 == pytest.approx(2.899e-4, rel=1e-3) and r64 == pytest.approx(5.40e-13, rel=1e-3)
    assert a32 == pytest.approx(3.41e-42, rel=1e-2) and a32 < np.finfo(np.float32).tiny
    assert a64 == 2432 * 2.0 ** -1074 and a64 < np.finfo(np.float64).tiny


# (code, n) -> (E_a, factor_a, allowance_a) from the MixedR2cPlan contract
# (mixed.rs:408-417) and the R2R normalization factors (r2r.rs:1222-1228).
AXIS_MODEL = [
    (O.AXIS_FFT, 5, (5, 5, 0)), (O.AXIS_FFT, 8, (8, 8, 0)),
    (O.DCT1, 5, (8, 8, 2)), (O.DCT1, 8, (14, 14, 2)),
    (O.DST1, 5, (12, 12, 2)), (O.DST1, 8, (18, 18, 2)),
    (O.DCT2, 5, (20, 10, 2)), (O.DCT3, 5, (20, 10, 2)), (O.DST2, 8, (32, 16, 2)), (O.DST3, 8, (32, 16, 2)),
    (O.DCT4, 5, (40, 10, 2)), (O.DST4, 8, (64, 16, 2)),
    (O.DHT, 5, (5, 5, 2)), (O.DHT, 8, (8, 8, 2)),
]


@pytest.mark.parametrize("code,n,model", AXIS_MODEL)
def test_mixed_axis_error_model(code, n, model):
    assert O.mixed_axis_model(code, n) == model
    assert O.mixed_axis_model_c(code, n) == model
    assert O.mixed_axis_model(0, n) is None and O.mixed_axis_model_c(0, n) is None


# MixedR2cPlan depth: D = 1 + sum over the non-identity axes before the RFFT
# axis of (ceil_log2(E_a) + allowance_a); raw backward x prod factor_a.
MIXED = [
    # DCT-IV n=5: E 40 -> 6 + 2; FFT n=6: 3 + 0  => D = 12; factors 10 * 6
    ((5, 6, 8), (O.DCT4, O.AXIS_FFT, O.AXIS_RFFT), 12, 60),
    # DCT-II n=8: E 32 -> 5 + 2; FFT n=12: 4 + 0 => D = 12; factors 16 * 12
    ((8, 12, 16), (O.DCT2, O.AXIS_FFT, O.AXIS_RFFT), 12, 192),
    # identity axis skipped; DST-I n=5: E 12 -> 4 + 2 => D = 7; factor 12
    ((4, 5, 10), (0, O.DST1, O.AXIS_RFFT), 7, 12),
    # DHT n=6: 3 + 2; DCT-I n=5: E 8 -> 3 + 2 => D = 11; factors 6 * 8
    ((6, 5, 9), (O.DHT, O.DCT1, O.AXIS_RFFT), 11, 48),
    # axes after the RFFT axis (real prefix) do not count: D = 1 + (3 + 0)
    ((8, 9, 6), (O.AXIS_FFT, O.AXIS_RFFT, O.DCT2), 4, 8),
]


@pytest.mark.parametrize("shape,codes,depth,factor", MIXED)
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_mixed_thresholds_match_the_closed_forms(shape, codes, depth, factor, dt):
    eps, e_min = (F32_EPS, -149) if dt == np.float32 else (F64_EPS, -1074)
    rel, ab = _both(O.KIND_MIXED_R2C, dt, shape, O.INVERSE, r2r=codes)
    assert rel == 128.0 * eps * depth
    assert ab == math.ldexp(128 * depth, e_min)
    _, ab_b = _both(O.KIND_MIXED_R2C, dt, shape, O.BACKWARD, r2r=codes)
    assert ab_b == ab * factor


# ------------------------------------------------------------ decisions --

def _accept_direct(planes, relative, absolute):
    """EndpointStatistics restated directly (distributed.rs:4270-4350) for a
    list of constrained planes of one batch (complex numpy arrays): all values
    finite; per plane max|Im| <= absolute or sqrt(sum (im/M)^2) <=
    relative * sqrt(sum (re/M)^2), M = max over the plane of max(|re|,|im|)."""
    ok = True
    for p in planes:
        re, im = p.real.astype(np.float64), p.imag.astype(np.float64)
        if not (np.all(np.isfinite(re)) and np.all(np.isfinite(im))):
            return False
        m = float(np.max(np.maximum(np.abs(re), np.abs(im)))) if p.size else 0.0
        mi = float(np.max(np.abs(im))) if p.size else 0.0
        sr = si = 0.0
        if m != 0.0:
            sr = float(np.sum((re / m) ** 2))
            si = float(np.sum((im / m) ** 2))
        if not (mi <= absolute or math.sqrt(si) <= relative * math.sqrt(sr)):
            ok = False
    return ok


def _planes_array(dc, ny, n_real, interior_seed=0):
    """[x][y][kz] identity-layout complex array with the given DC and Nyquist
    planes (interior bins random)."""
    rng = np.random.default_rng(interior_seed)
    c = n_real // 2 + 1
    shape = dc.shape
    a = rng.standard_normal(shape + (c,)) + 1j * rng.standard_normal(shape + (c,))
    a = a.astype(dc.dtype)
    a[..., 0] = dc
    if n_real % 2 == 0:
        a[..., c - 1] = ny
    return a


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_absolute_branch_at_the_exact_threshold(dt):
    """max|Im| == absolute passes, the next representable value fails (Re = 0,
    so only the absolute branch can accept a nonzero Im)."""
    cd = np.complex64 if dt == np.float32 else np.complex128
    shape, n = (4, 6), 8
    rel, ab = _both(O.KIND_R2C, dt, shape + (n,), O.INVERSE)
    t = dt(ab)
    assert float(t) == ab  # 128 * D min-subnormals is representable
    up = np.nextafter(t, dt(1))
    for im, want in ((t, True), (up, False), (-t, True), (-up, False)):
        dc = np.zeros(shape, dtype=cd)
        dc[1, 2] = 1j * im
        ny = np.zeros(shape, dtype=cd)
        arr = _planes_array(dc, ny, n)
        got = O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)
        assert got == want == _accept_direct([dc, ny], rel, ab)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_relative_branch_agrees_with_the_direct_restatement(dt):
    """Random planes whose ||Im|| / ||Re|| sits within 1e-5 of the relative
    threshold, with plane maxima far apart (the scaled sums matter): the C
    restatement's decision equals the direct one on every case."""
    cd = np.complex64 if dt == np.float32 else np.complex128
    rng = np.random.default_rng(12)
    shape, n = (5, 7), 10
    rel, ab = _both(O.KIND_R2C, dt, shape + (n,), O.INVERSE)
    seen = set()
    for case in range(200):
        planes = []
        for _ in range(2):
            scale = 10.0 ** rng.uniform(-30, 30)
            re = rng.standard_normal(shape) * scale
            im = rng.standard_normal(shape)
            ratio = rel * (1.0 + rng.choice([-1, 1]) * 10.0 ** rng.uniform(-5, -1))
            im *= ratio * np.linalg.norm(re) / np.linalg.norm(im)
            planes.append((re + 1j * im).astype(cd))
        arr = _planes_array(planes[0], planes[1], n, case)
        want = _accept_direct(planes, rel, ab)
        got = O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)
        assert got == want, case
        seen.add(want)
    assert seen == {True, False}


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_constrained_values_reject(bad):
    shape, n = (3, 4), 6
    rel, ab = _both(O.KIND_R2C, np.float64, shape + (n,), O.INVERSE)
    dc = np.ones(shape, dtype=np.complex128)
    ny = np.ones(shape, dtype=np.complex128)
    for which, part in ((0, 1.0), (1, 1j)):
        d, y = dc.copy(), ny.copy()
        (d if which == 0 else y)[2, 1] = bad * part if part == 1.0 else complex(1.0, bad)
        arr = _planes_array(d, y, n)
        assert not O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)
        assert not _accept_direct([d, y], rel, ab)
    # a non-finite interior bin is not checked
    arr = _planes_array(dc, ny, n)
    arr[0, 0, 1] = bad
    assert O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)


def test_odd_length_has_no_nyquist_plane():
    shape, n = (3, 4), 7
    rel, ab = _both(O.KIND_R2C, np.float64, shape + (n,), O.INVERSE)
    arr = _planes_array(np.ones(shape, dtype=np.complex128), None, n)
    arr[..., -1] = 5j  # the last stored bin of an odd length is unconstrained
    assert O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)
    arr[..., 0] = 1 + 5j
    assert not O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), 1, 2, n, rel, ab)


def test_one_failing_extra_batch_rejects():
    shape, n, extra = (4, 6), 8, 3
    rel, ab = _both(O.KIND_R2C, np.float64, shape + (n,), O.INVERSE)
    rng = np.random.default_rng(3)
    batches = []
    for b in range(extra):
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
        ratio = 1.1 if b == 1 else 0.5
        im *= ratio * rel * np.linalg.norm(re) / np.linalg.norm(im)
        batches.append(_planes_array(re + 1j * im, np.ones(shape, dtype=np.complex128), n, b))
    arr = np.stack(batches)
    assert not O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), extra, 2, n, rel, ab)
    arr[1, ..., 0] = arr[0, ..., 0]
    assert O.endpoint_accept(arr, (0, 1, 2), shape + (n // 2 + 1,), extra, 2, n, rel, ab)


def test_planted_threshold_slip_is_caught():
    """A plausible slip (FLT_MIN in place of the smallest subnormal, or the
    FFT length in place of the error-model length E_a) changes the thresholds
    these tests pin."""
    _, ab = _both(O.KIND_R2C, np.float32, (4, 6, 8), O.INVERSE)
    assert ab != 128.0 * float(np.finfo(np.float32).tiny) * 6
    _, ab_m = _both(O.KIND_MIXED_R2C, np.float32, (5, 6, 8), O.INVERSE, r2r=(O.DCT4, O.AXIS_FFT, O.AXIS_RFFT))
    d_wrong = 1 + ((5 - 1).bit_length() + 2) + 3
    assert ab_m != math.ldexp(128 * d_wrong, -149)
