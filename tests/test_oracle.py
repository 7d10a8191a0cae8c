"""CPU tests that pin the oracle (no GPU).

Each check is independent of the oracle's own formulas: numpy / scipy
goldens, direct O(n^2) sums in extended precision, the reference's
definitional `direct` R2R sums (crates/pencil-fft/src/r2r.rs:1064-1112) and
cas sums (dht.rs:434-449), its known-answer tests (lib.rs:916-966,
r2c.rs:1195-1245), Parseval, numpy transposes of integer-coded arrays
(pencil-array/tests/local_transpose.rs:78-91), hand-built spectra around the
endpoint thresholds (distributed/r2c.rs:4841-4878, distributed.rs:4330-4349)
and adjoint dot-product identities.
"""
import itertools
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLDEN, allow_pickle=False)


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d else 1.0)


# ---------------------------------------------------------------- lines --

def test_c2c_lines_match_numpy_goldens(gold):
    for key in gold.files:
        if key.startswith("c2c_") and key.endswith("_in"):
            base = key[:-3]
            x = gold[key]
            assert rel(O.c2c_lines(x, O.FORWARD), gold[base + "_fwd"]) < 1e-14, base
            assert rel(O.c2c_lines(x, O.INVERSE), gold[base + "_inv"]) < 1e-14, base


def test_r2c_lines_match_numpy_goldens(gold):
    for key in gold.files:
        if key.startswith("r2c_") and key.endswith("_in"):
            base = key[:-3]
            x = gold[key]
            n = x.shape[-1]
            y = O.r2c_lines(x)
            assert rel(y, gold[base + "_fwd"]) < 1e-14, base
            assert rel(O.c2r_lines(y, n, O.INVERSE), x) < 1e-14, base


def _direct_dft(x, sign):
    x = np.asarray(x, dtype=np.clongdouble)
    n = x.shape[-1]
    j = np.arange(n)
    pi = np.arctan(np.longdouble(1)) * 4
    ang = (np.outer(j, j) % n).astype(np.longdouble) * (2 * pi / n) * sign
    w = np.cos(ang) + 1j * np.sin(ang).astype(np.longdouble)
    return x @ w.T


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 11, 12, 13, 17, 31])
def test_c2c_lines_match_direct_extended_sum(n):
    """The reference's dft_oracle (lib.rs:774-799) in long double."""
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (3, n)) + 1j * rng.uniform(-1, 1, (3, n))
    assert rel(O.c2c_lines(x, O.FORWARD), _direct_dft(x, -1).astype(np.complex128)) < 1e-15
    assert rel(O.c2c_lines(x, O.BACKWARD), _direct_dft(x, +1).astype(np.complex128)) < 1e-15
    # the extended-precision oracle path agrees too
    assert rel(O.c2c_lines(x, O.FORWARD, extended=True), _direct_dft(x, -1).astype(np.complex128)) < 4e-16


def test_c2c_peak_sign_and_dc_scale():
    n, peak = 5, 2
    x = np.exp(2j * np.pi * peak * np.arange(n) / n)[None, :]
    y = O.c2c_lines(x)
    exp = np.zeros(n, complex)
    exp[peak] = n
    assert np.abs(y - exp).max() < 1e-13
    dc = np.full((1, n), 2 - 3j)
    y = O.c2c_lines(dc)
    assert abs(y[0, 0] - (2 - 3j) * n) < 1e-13 and np.abs(y[0, 1:]).max() < 1e-13


def test_r2c_frequency_dc_and_normalization():
    n, peak = 8, 2
    x = np.cos(2 * np.pi * peak * np.arange(n) / n)[None, :]
    y = O.r2c_lines(x)
    exp = np.zeros(n // 2 + 1)
    exp[peak] = n / 2
    assert np.abs(y - exp).max() < 1e-13
    dc = np.full((1, n), 3.5)
    y = O.r2c_lines(dc)
    assert abs(y[0, 0] - 3.5 * n) < 1e-13
    assert np.abs(O.c2r_lines(y, n, O.INVERSE) - dc).max() < 1e-14
    assert np.abs(O.c2r_lines(y, n, O.BACKWARD) - dc * n).max() < 1e-12


@pytest.mark.parametrize("n", [4, 9, 64, 385])
def test_parseval(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    y = O.c2c_lines(x)
    assert abs(np.sum(np.abs(y) ** 2) / n - np.sum(np.abs(x) ** 2)) < 1e-12 * np.sum(np.abs(x) ** 2)


def _direct_r2r(kind, x):
    """The reference's `direct` definitions (r2r.rs:1064-1112)."""
    x = np.asarray(x, dtype=np.longdouble)
    n = x.shape[-1]
    pi = np.arctan(np.longdouble(1)) * 4
    out = np.zeros_like(x)
    for k in range(n):
        for j in range(n):
            if kind == 1:
                c = (-1.0 if (j == n - 1 and k % 2) else 1.0) if j in (0, n - 1) else 2 * np.cos(pi * j * k / (n - 1))
            elif kind == 2:
                c = 2 * np.cos(pi * (j + 0.5) * k / n)
            elif kind == 3:
                c = 1.0 if j == 0 else 2 * np.cos(pi * j * (k + 0.5) / n)
            elif kind == 4:
                c = 2 * np.cos(pi * (j + 0.5) * (k + 0.5) / n)
            elif kind == 5:
                c = 2 * np.sin(pi * (j + 1) * (k + 1) / (n + 1))
            elif kind == 6:
                c = 2 * np.sin(pi * (j + 0.5) * (k + 1) / n)
            elif kind == 7:
                c = ((1.0 if k % 2 == 0 else -1.0) if j == n - 1 else 2 * np.sin(pi * (j + 1) * (k + 0.5) / n))
            else:
                c = 2 * np.sin(pi * (j + 0.5) * (k + 0.5) / n)
            out[..., k] += x[..., j] * c
    return out.astype(np.float64)


@pytest.mark.parametrize("kind", range(1, 9))
def test_r2r_lines_match_reference_direct_definitions(kind):
    rng = np.random.default_rng(kind)
    for n in [1, 2, 3, 4, 5, 8, 13]:
        if kind == 1 and n == 1:
            with pytest.raises(ValueError):
                O.r2r_lines(np.zeros((1, 1)), kind)
            continue
        x = rng.uniform(-1, 1, (3, n))
        assert rel(O.r2r_lines(x, kind), _direct_r2r(kind, x)) < 1e-14, (kind, n)


def test_r2r_and_dht_lines_match_scipy_goldens(gold):
    for key in gold.files:
        if key.startswith("r2r_") and key.endswith("_in"):
            base = key[:-3]
            kind = int(base.split("_")[1][1:])
            x = gold[key]
            assert rel(O.r2r_lines(x, kind), gold[base + "_fwd"]) < 1e-13, base


@pytest.mark.parametrize("kind", range(1, 10))
def test_r2r_inverse_and_backward_pairing(kind):
    """backward = paired kind; inverse = backward / normalization factor
    (r2r.rs:66-78, :798-848)."""
    rng = np.random.default_rng(30 + kind)
    n = 12
    x = rng.uniform(-1, 1, (2, n))
    y = O.r2r_lines(x, kind)
    assert rel(O.r2r_lines(y, kind, O.INVERSE), x) < 1e-14
    pair = O.R2R_BACKWARD_KIND[kind]
    assert rel(O.r2r_lines(y, kind, O.BACKWARD), O.r2r_lines(y, pair, O.FORWARD)) == 0.0


def test_dht_matches_cas_sum_and_length_one_identity():
    rng = np.random.default_rng(3)
    for n in [2, 3, 7, 12]:
        x = rng.uniform(-1, 1, (2, n))
        j = np.arange(n)
        ang = 2 * np.pi * np.outer(j, j) / n
        exp = x @ (np.cos(ang) + np.sin(ang)).T
        assert rel(O.r2r_lines(x, O.DHT), exp) < 1e-14
    x = np.array([[1.5], [-0.0]])
    assert O.r2r_lines(x, O.DHT).tobytes() == x.tobytes()


# --------------------------------------------------------------- routes --

def _meta(m):
    m = list(m)
    cut = m.index(-1)
    return tuple(m[: cut - 1]), m[cut - 1], m[cut + 1:]


def test_routes_match_numpy_goldens(gold):
    for kind, prefix in ((O.KIND_C2C, "route_c2c"), (O.KIND_R2C, "route_r2c"), (O.KIND_R2R, "route_r2r"),
                         (O.KIND_C2R, "route_c2r"), (O.KIND_MIXED_C2C, "route_mixc2c"),
                         (O.KIND_MIXED_R2C, "route_mixr2c")):
        i = 0
        while f"{prefix}{i}_in" in gold.files:
            shape, extra, tail = _meta(gold[f"{prefix}{i}_meta"])
            x = gold[f"{prefix}{i}_in"]
            kw = {"r2r": tail} if kind in (O.KIND_R2R, O.KIND_MIXED_C2C, O.KIND_MIXED_R2C) else {"selected": tail}
            y = O.transform(kind, x, O.FORWARD, shape=shape, extra=extra, **kw)
            assert rel(y, gold[f"{prefix}{i}_fwd"]) < 1e-13, (prefix, i)
            back = O.transform(kind, y, O.INVERSE, shape=shape, extra=extra, **kw)
            assert rel(back, x) < 1e-13, (prefix, i)
            i += 1
        assert i > 0


def test_route_memory_orders_follow_build_route():
    assert O.route(3) == [(0, 1, 2), (0, 2, 1), (2, 1, 0)]
    assert O.route(4) == [(0, 1, 2, 3), (0, 1, 3, 2), (0, 3, 2, 1), (3, 2, 1, 0)]
    assert O.route(3, permute_dims=False) == [(0, 1, 2)] * 3


def test_c2c_per_axis_directions():
    rng = np.random.default_rng(8)
    shape = (3, 4, 5)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    y = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shape, backward_dir=[0, 1, 0])
    exp = np.fft.fft(np.fft.ifft(np.fft.fft(x, axis=0), axis=1) * shape[1], axis=2)
    assert rel(y, np.transpose(exp, (2, 1, 0))) < 1e-14


def _dot(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.sum(a.real * b.real) + np.sum(np.imag(a) * np.imag(b)))


@pytest.mark.parametrize("kind,shape,kw", [
    (O.KIND_C2C, (4, 6, 5), {}),
    (O.KIND_R2C, (4, 6, 8), {}),
    (O.KIND_R2C, (5, 6, 7), {}),
    (O.KIND_R2C, (4, 6, 8), {"selected": [0, 1]}),
    (O.KIND_R2R, (5, 6, 7), {"r2r": [2, 6, 1]}),
    (O.KIND_R2R, (6, 5, 8), {"r2r": [3, 7, 9]}),
    (O.KIND_R2R, (6, 5, 8), {"r2r": [4, 5, 8]}),
])
def test_oracle_adjoints_satisfy_dot_identity(kind, shape, kw):
    rng = np.random.default_rng(len(shape) + sum(shape))
    cplx = kind == O.KIND_C2C
    x = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    ax = O.transform(kind, x, O.FORWARD, shape=shape, **kw)
    y = rng.standard_normal(ax.shape) + (1j * rng.standard_normal(ax.shape) if np.iscomplexobj(ax) else 0)
    aty = O.transform(kind, y, O.ADJ_FORWARD, shape=shape, **kw)
    assert abs(_dot(ax, y) - _dot(x, aty)) <= 1e-13 * abs(_dot(ax, y))
    z = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    for op, adj in ((O.INVERSE, O.ADJ_INVERSE), (O.BACKWARD, O.ADJ_BACKWARD)):
        bx = O.transform(kind, ax, op, shape=shape, **kw)
        btz = O.transform(kind, z, adj, shape=shape, **kw)
        assert abs(_dot(bx, z) - _dot(ax, btz)) <= 1e-13 * abs(_dot(bx, z))


@pytest.mark.parametrize("kind,shape,codes", [
    (O.KIND_MIXED_C2C, (4, 5, 6), [10, 2, 10]),
    (O.KIND_MIXED_C2C, (6, 4, 5), [1, 9, 10]),
    (O.KIND_MIXED_C2C, (5, 6, 7), [3, 7, 5]),
    (O.KIND_MIXED_R2C, (4, 6, 5), [10, 11, 2]),
    (O.KIND_MIXED_R2C, (6, 5, 8), [6, 10, 11]),
    (O.KIND_MIXED_R2C, (5, 7, 6), [4, 11, 1]),
])
def test_oracle_mixed_adjoints_satisfy_dot_identity(kind, shape, codes):
    rng = np.random.default_rng(sum(shape) + codes[0])
    cplx = kind == O.KIND_MIXED_C2C
    x = rng.standard_normal(int(np.prod(shape))) + (1j * rng.standard_normal(int(np.prod(shape))) if cplx else 0)
    ax = O.transform(kind, x, O.FORWARD, shape=shape, r2r=codes)
    y = rng.standard_normal(ax.shape) + 1j * rng.standard_normal(ax.shape)
    aty = O.transform(kind, y, O.ADJ_FORWARD, shape=shape, r2r=codes)
    assert abs(_dot(ax, y) - _dot(x, aty)) <= 1e-13 * abs(_dot(ax, y))
    z = rng.standard_normal(x.shape) + (1j * rng.standard_normal(x.shape) if cplx else 0)
    for op, adj in ((O.INVERSE, O.ADJ_INVERSE), (O.BACKWARD, O.ADJ_BACKWARD)):
        bx = O.transform(kind, ax, op, shape=shape, r2r=codes)
        btz = O.transform(kind, z, adj, shape=shape, r2r=codes)
        assert abs(_dot(bx, z) - _dot(ax, btz)) <= 1e-13 * abs(_dot(bx, z))


def test_oracle_mixed_graph_rules():
    """validate_r2c_graph (mixed.rs:1602-1629): one RFFT axis, no FFT after it;
    no RFFT in a mixed C2C plan."""
    x = np.zeros(4 * 5 * 6)
    for codes in ([10, 2, 10], [11, 11, 0], [0, 11, 10]):
        with pytest.raises(Exception):
            O.transform(O.KIND_MIXED_R2C, x, O.FORWARD, shape=(4, 5, 6), r2r=codes)
    with pytest.raises(Exception):
        O.transform(O.KIND_MIXED_C2C, x.astype(complex), O.FORWARD, shape=(4, 5, 6), r2r=[11, 0, 0])


@pytest.mark.parametrize("shape,kw", [((6, 5, 4), {}), ((5, 6, 7), {}), ((4, 6, 5), {"selected": [1, 2]})])
def test_oracle_c2r_orientation_adjoints_satisfy_dot_identity(shape, kw):
    """C2rPlan orientation: FORWARD (raw C2R) from a valid canonical spectrum,
    then the R2C-direction ops (INVERSE, BACKWARD) from the real side."""
    rng = np.random.default_rng(sum(shape))
    r = rng.standard_normal(int(np.prod(shape)))
    x = O.transform(O.KIND_C2R, r, O.BACKWARD, shape=shape, **kw)
    ax = O.transform(O.KIND_C2R, x, O.FORWARD, shape=shape, **kw)
    y = rng.standard_normal(ax.shape)
    aty = O.transform(O.KIND_C2R, y, O.ADJ_FORWARD, shape=shape, **kw)
    assert abs(_dot(ax, y) - _dot(x, aty)) <= 1e-13 * abs(_dot(ax, y))
    z = rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)
    for op, adj in ((O.INVERSE, O.ADJ_INVERSE), (O.BACKWARD, O.ADJ_BACKWARD)):
        bx = O.transform(O.KIND_C2R, ax, op, shape=shape, **kw)
        btz = O.transform(O.KIND_C2R, z, adj, shape=shape, **kw)
        assert abs(_dot(bx, z) - _dot(ax, btz)) <= 1e-13 * abs(_dot(bx, z))


def test_a_planted_sign_error_is_caught_by_the_goldens(gold):
    """A transposed-sign oracle (backward instead of forward on one axis)
    must fail the numpy golden; guards against goldens that pin nothing."""
    shape, extra, sel = _meta(gold["route_c2c1_meta"])
    x = gold["route_c2c1_in"]
    # axis 2 has length 5 (for length 2 both signs coincide)
    wrong = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shape, extra=extra, selected=sel,
                        backward_dir=[0, 0, 1])
    assert rel(wrong, gold["route_c2c1_fwd"]) > 1e-3


# ----------------------------------------------------------- transposes --

def _coded(shape, extra):
    idx = np.indices(shape).reshape(len(shape), -1)
    code = np.zeros(idx.shape[1], dtype=np.uint64)
    for a in range(len(shape)):
        code = code * 100 + idx[a].astype(np.uint64)
    return np.concatenate([code + np.uint64(e * 1_000_000) for e in range(extra)])


@pytest.mark.parametrize("shape", [(3, 4, 5), (2, 3, 4, 5), (1, 4, 1)])
def test_local_transpose_matches_numpy_for_every_permutation(shape):
    nd = len(shape)
    extra = 2
    logical = _coded(shape, extra).reshape((extra,) + shape)
    perms = list(itertools.permutations(range(nd)))
    for sp in perms:
        src = np.ascontiguousarray(np.transpose(logical, (0,) + tuple(a + 1 for a in sp))).reshape(-1)
        for dp in perms[:: max(1, len(perms) // 6)]:
            exp = np.ascontiguousarray(np.transpose(logical, (0,) + tuple(a + 1 for a in dp))).reshape(-1)
            got = O.local_transpose(src, shape, extra, sp, dp)
            assert np.array_equal(got, exp), (sp, dp)
            got2, payload = O.alltoallv_1x1(src, shape, extra, sp, dp)
            assert np.array_equal(got2, exp)
            # payload order = global row-major of the region per extra block
            assert np.array_equal(payload, logical.reshape(-1))


# --------------------------------------------------- endpoint statistics --

def _spectrum_planes(n_real=8, shape=(4, 6), im_scale=0.0, seed=0):
    """A valid half-spectrum [x][y][kz] (identity layout) with the DC and
    Nyquist planes' imaginary parts set to im_scale * N(0,1)."""
    rng = np.random.default_rng(seed)
    c = n_real // 2 + 1
    s = rng.standard_normal(shape + (c,)) + 1j * rng.standard_normal(shape + (c,))
    for k in (0, c - 1):
        s[..., k] = s[..., k].real + 1j * im_scale * rng.standard_normal(shape)
    return s


def test_endpoint_statistics_thresholds():
    shape, n = (4, 6), 8
    rshape = shape + (n // 2 + 1,)
    perm = (0, 1, 2)
    sel = [0, 1, 2]
    relative, absolute = O.endpoint_thresholds(np.float64, shape + (n,), sel, 2, True)
    # depth = 1 + ceil(log2 4) + ceil(log2 6) = 1 + 2 + 3
    assert relative == pytest.approx(128 * np.finfo(np.float64).eps * 6)
    s = _spectrum_planes(n, shape, 0.0)
    assert O.endpoint_accept(s, perm, rshape, 1, 2, n, relative, absolute)

    def with_ratio(ratio):
        t = s.copy()
        for k in (0, n // 2):
            re = t[..., k].real
            im = np.ones_like(re)
            im *= ratio * relative * np.linalg.norm(re) / np.linalg.norm(im)
            t[..., k] = re + 1j * im
        return t

    assert O.endpoint_accept(with_ratio(0.9), perm, rshape, 1, 2, n, relative, absolute)
    assert not O.endpoint_accept(with_ratio(1.1), perm, rshape, 1, 2, n, relative, absolute)
    t = s.copy()
    t[0, 0, 0] = np.nan
    assert not O.endpoint_accept(t, perm, rshape, 1, 2, n, relative, absolute)
    # raw backward scales the absolute threshold by the transverse lengths
    r2, a2 = O.endpoint_thresholds(np.float64, shape + (n,), sel, 2, False)
    assert a2 == pytest.approx(absolute * 4 * 6) and r2 == relative


def test_invalid_spectrum_is_rejected_by_the_route():
    shape = (4, 6, 8)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    y = O.transform(O.KIND_R2C, x, O.FORWARD, shape=shape)
    y[0] += 0.25j
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_R2C, y, O.INVERSE, shape=shape)
