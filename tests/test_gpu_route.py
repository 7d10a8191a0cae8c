"""GPU parity of the one-rank distributed plans and the layout permutations.

Shapes follow the reference's fixture case list
(tools/fftw-reference/README.md:271-294) and its input formulas
(tools/fftw-reference/generate.jl:201-250): full-axis C2C / R2C, partial
selections, extras, mixed R2R / DHT kinds. Every operation, including the
adjoints, is compared with the oracle; permutations are bit-exact with
integer-coded values (pencil-array/tests/local_transpose.rs:78-91).
"""
import itertools

import numpy as np
import pytest
import torch

import oracle as O
import pencil_fft_hip as pf
from util import complex_of, real_of, rel_l2, tol, uniform_complex, uniform_real

pytestmark = pytest.mark.gpu

OPS = [O.FORWARD, O.INVERSE, O.BACKWARD, O.ADJ_FORWARD, O.ADJ_INVERSE, O.ADJ_BACKWARD]


def _tdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def complex_value(coords, seed):
    """generate.jl:201-220 (coordinates in Rust logical order)."""
    re = 0.173 + 0.019 * seed
    im = -0.271 - 0.013 * seed
    for axis, c in enumerate(coords, start=1):
        re += (0.071 + 0.009 * axis) * (c + 1)
        im += (0.053 + 0.011 * axis) * (c + 2)
        re += 0.017 * np.sin((axis + 2) * (c + 1) + seed * 0.07)
        im += 0.023 * np.cos((axis + 3) * (c + 2) - seed * 0.05)
    for left in range(1, len(coords)):
        for right in range(left + 1, len(coords) + 1):
            re += 0.0067 * left * (right + 1) * (coords[left - 1] + 1) * (coords[right - 1] + 2)
            im -= 0.0049 * (left + 1) * right * (coords[left - 1] + 2) * (coords[right - 1] + 1)
    return re + 1j * im


def real_value(coords, seed):
    """generate.jl:222-234."""
    v = 0.417 + 0.023 * seed
    for axis, c in enumerate(coords, start=1):
        v += (0.087 + 0.013 * axis) * (c + 1)
        v += 0.031 * np.sin((axis + 1) * (c + 2) + seed * 0.09)
    for left in range(1, len(coords)):
        for right in range(left + 1, len(coords) + 1):
            v += 0.0083 * (left + 1) * (right + 2) * (coords[left - 1] + 1) * (coords[right - 1] + 2)
    return v


def fixture_field(shape, extra, seed, complex_):
    out = np.empty((extra,) + tuple(shape), dtype=np.complex128 if complex_ else np.float64)
    for e in range(extra):
        for idx in itertools.product(*[range(n) for n in shape]):
            out[(e,) + idx] = (complex_value if complex_ else real_value)(list(idx), seed + e)
    return out


def run_all_ops(plan, kind, shape, extra, x_in, x_out, dt, **kw):
    """Checks all six operations of a plan against the oracle. x_in / x_out
    are inputs in the plan's input / output layouts (flat)."""
    for op in OPS:
        fwd_like = op in (O.FORWARD, O.ADJ_INVERSE, O.ADJ_BACKWARD)
        src = x_in if fwd_like else x_out
        ref = O.transform(kind, src.astype(np.complex128 if np.iscomplexobj(src) else np.float64), op,
                          shape=shape, extra=extra, **kw)
        got = _host(plan.execute(op, _tdev(src)))
        err = rel_l2(got, ref)
        assert err <= 4 * tol(dt), (shape, extra, kw, op, err)


C2C_CASES = [
    ((3, 4), 1, None),
    ((3, 2, 5), 6, None),
    ((2, 1, 3, 4), 2, None),
    ((2, 3, 2, 3), 1, []),
    ((2, 3, 2, 3), 1, [0, 3]),
    ((8, 12, 10), 2, None),
    ((16, 48, 60), 1, [0, 2]),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extra,sel", C2C_CASES)
def test_c2c_plan_matches_oracle(gpu, shape, extra, sel, dt, permute):
    cd = complex_of(dt)
    x = fixture_field(shape, extra, 1.0, True).astype(cd)
    plan = pf.C2cPlan(shape, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else (),
                      selection=sel, permute_dims=permute)
    y = O.transform(O.KIND_C2C, x.astype(np.complex128), O.FORWARD, shape=shape, extra=extra, selected=sel,
                    permute_dims=permute).astype(cd)
    run_all_ops(plan, O.KIND_C2C, shape, extra, x.reshape(-1), y, dt, selected=sel, permute_dims=permute)


def test_c2c_plan_directions(gpu):
    shape = (4, 6, 10)
    rng = np.random.default_rng(3)
    x = uniform_complex(rng, shape, np.complex128)
    dirs = [False, True, False]
    plan = pf.C2cPlan(shape, dtype=torch.complex128, directions=dirs)
    y = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shape, backward_dir=dirs)
    run_all_ops(plan, O.KIND_C2C, shape, 1, x.reshape(-1), y, np.float64, backward_dir=dirs)


R2C_CASES = [
    ((3, 1), 1, None),
    ((3, 2), 1, None),
    ((3, 1, 4), 1, None),
    ((3, 1, 5), 6, None),
    ((2, 1, 3, 3), 2, None),
    ((2, 3, 2, 3), 1, [0]),
    ((2, 3, 2, 3), 1, [0, 2]),
    ((2, 3, 2, 3), 1, [0, 3]),
    ((2, 3, 4), 2, [0, 2]),
    ((8, 12, 16), 1, None),
    ((16, 10, 9), 3, None),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extra,sel", R2C_CASES)
def test_r2c_plan_matches_oracle(gpu, shape, extra, sel, dt, permute):
    x = fixture_field(shape, extra, 2.0, False).astype(dt)
    plan = pf.R2cPlan(shape, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else (),
                      selection=sel, permute_dims=permute)
    y = O.transform(O.KIND_R2C, x.astype(np.float64), O.FORWARD, shape=shape, extra=extra, selected=sel,
                    permute_dims=permute).astype(complex_of(dt))
    run_all_ops(plan, O.KIND_R2C, shape, extra, x.reshape(-1), y, dt, selected=sel, permute_dims=permute)


def test_r2c_inverse_rejects_invalid_spectrum_without_writing(gpu):
    shape = (4, 6, 8)
    rng = np.random.default_rng(4)
    x = uniform_real(rng, shape, np.float64)
    plan = pf.R2cPlan(shape, dtype=torch.float64)
    spec = plan.forward(_tdev(x))
    host = _host(spec).copy()
    # find the DC plane element (kz = 0) of the output layout [kz][ky][kx] and corrupt it
    host[3] += 0.5j
    dst = torch.full((x.size,), -1.0, dtype=torch.float64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.inverse(_tdev(host), dst)
    assert e.value.variant == "InvalidSpectrum"
    assert torch.all(dst == -1.0)
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_R2C, host, O.INVERSE, shape=shape)
    # the untouched spectrum passes and round-trips
    back = _host(plan.inverse(spec))
    assert rel_l2(back, x.reshape(-1)) <= 1e-12


C2R_CASES = [
    ((3, 4), 1, None),
    ((3, 2, 5), 2, None),
    ((4, 3, 5), 1, None),
    ((2, 3, 2, 3), 1, [1, 3]),
    ((5, 4, 6), 1, [0, 2]),
    ((16, 12, 8), 1, None),
    ((9, 10, 16), 3, None),
    ((12, 10, 16), 1, [1, 2]),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extra,sel", C2R_CASES)
def test_c2r_plan_matches_oracle(gpu, shape, extra, sel, dt, permute):
    """C2rPlan (distributed/c2r.rs:72-361): all six operations against the
    oracle's C2R orientation; the forward input is the (valid) spectrum of a
    fixture field in the canonical memory order."""
    kw = dict(shape=shape, extra=extra, selected=sel, permute_dims=permute)
    # a fixture field as the plan's real array (its memory order does not
    # matter for the comparison), and its normalized spectrum as the forward input
    real_in = fixture_field(shape, extra, 4.0, False).reshape(-1)
    spec = O.transform(O.KIND_C2R, real_in, O.INVERSE, **kw)
    plan = pf.C2rPlan(shape, dtype=torch.float32 if dt == np.float32 else torch.float64,
                      extra_shape=(extra,) if extra > 1 else (), selection=sel, permute_dims=permute)
    x_in = spec.astype(complex_of(dt))
    x_out = real_in.astype(real_of(dt))
    run_all_ops(plan, O.KIND_C2R, shape, extra, x_in, x_out, dt, selected=sel, permute_dims=permute)


def test_c2r_plan_forward_rejects_invalid_spectrum_without_writing(gpu):
    shape = (6, 4, 8)
    rng = np.random.default_rng(5)
    plan = pf.C2rPlan(shape, dtype=torch.float64)
    real = uniform_real(rng, shape, np.float64).reshape(-1)
    spec = _host(plan.backward(_tdev(real))).copy()
    spec[5] += 0.25j  # a DC-plane value of the reduction axis 0 (canonical order: axis 0 slowest)
    dst = torch.full((real.size,), -1.0, dtype=torch.float64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.forward(_tdev(spec), dst)
    assert e.value.variant == "InvalidSpectrum"
    assert torch.all(dst == -1.0)
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_C2R, spec, O.FORWARD, shape=shape)
    with pytest.raises(ValueError):
        plan.execute_in_place(pf.FORWARD, _tdev(spec))


F, R = pf.AxisTransform.FFT, pf.AxisTransform.RFFT
MIXED_C2C_CASES = [
    ((4, 5, 6), 1, [F, 2, F]),
    ((5, 6, 4), 2, [9, F, 4]),
    ((3, 4, 5), 1, [0, F, 5]),
    ((6, 4, 5), 1, [1, 9, F]),
    ((4, 3, 2, 5), 1, [F, 6, 0, 8]),
    ((16, 12, 20), 3, [F, 3, 7]),
]
MIXED_R2C_CASES = [
    ((4, 6, 5), 1, [F, R, 2]),
    ((5, 4, 6), 2, [9, R, 0]),
    ((6, 5, 8), 1, [2, F, R]),
    ((4, 6, 5, 3), 1, [F, 6, R, 4]),
    ((12, 16, 10), 1, [3, R, 9]),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128])
@pytest.mark.parametrize("shape,extra,codes", MIXED_C2C_CASES)
def test_mixed_c2c_plan_matches_oracle(gpu, shape, extra, codes, dt, permute):
    """MixedC2cPlan (distributed/mixed.rs:392-439): FFT / R2R / identity per
    axis on complex data, all six operations against the oracle."""
    x = fixture_field(shape, extra, 5.0, True).astype(dt).reshape(-1)
    plan = pf.MixedC2cPlan(shape, codes, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else (),
                           permute_dims=permute)
    y = O.transform(O.KIND_MIXED_C2C, x.astype(np.complex128), O.FORWARD, shape=shape, extra=extra, r2r=codes,
                    permute_dims=permute).astype(dt)
    run_all_ops(plan, O.KIND_MIXED_C2C, shape, extra, x, y, dt, r2r=codes, permute_dims=permute)


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extra,codes", MIXED_R2C_CASES)
def test_mixed_r2c_plan_matches_oracle(gpu, shape, extra, codes, dt, permute):
    """MixedR2cPlan (distributed/mixed.rs:441-446): real R2R prefix, the RFFT
    boundary, complex FFT / R2R suffix; all six operations against the oracle
    (the inverse with the mixed endpoint thresholds)."""
    x = fixture_field(shape, extra, 6.0, False).astype(dt).reshape(-1)
    plan = pf.MixedR2cPlan(shape, codes, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else (),
                           permute_dims=permute)
    y = O.transform(O.KIND_MIXED_R2C, x.astype(np.float64), O.FORWARD, shape=shape, extra=extra, r2r=codes,
                    permute_dims=permute).astype(complex_of(dt))
    run_all_ops(plan, O.KIND_MIXED_R2C, shape, extra, x, y, dt, r2r=codes, permute_dims=permute)


def test_mixed_plans_reject_invalid_graphs(gpu):
    for codes in ([F, 2, F], [R, R, 0], [0, R, F]):
        with pytest.raises(pf.PfhError) as e:
            pf.MixedR2cPlan((4, 5, 6), codes)
        assert e.value.variant == "InvalidGraph"
    with pytest.raises(pf.PfhError) as e:
        pf.MixedC2cPlan((4, 5, 6), [R, 0, 0])
    assert e.value.variant == "InvalidGraph"
    with pytest.raises(pf.PfhError) as e:
        pf.MixedC2cPlan((4, 5, 6), [2, F, 0], directions=[1, 0, 0])
    assert e.value.variant == "PreparationFailed"


def test_mixed_r2c_inverse_rejects_invalid_spectrum(gpu):
    shape, codes = (6, 4, 8), [F, R, 2]
    rng = np.random.default_rng(6)
    plan = pf.MixedR2cPlan(shape, codes, dtype=torch.float64)
    x = uniform_real(rng, shape, np.float64).reshape(-1)
    spec = _host(plan.forward(_tdev(x))).copy()
    # output memory order [axis2][axis1][axis0]: element 1 is axis1 = 0 (the DC plane), axis0 = 1
    spec[1] += 0.5j
    dst = torch.full((x.size,), -1.0, dtype=torch.float64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.inverse(_tdev(spec), dst)
    assert e.value.variant == "InvalidSpectrum"
    assert torch.all(dst == -1.0)
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_MIXED_R2C, spec, O.INVERSE, shape=shape, r2r=codes)


def test_r2c_partial_selection_inverse_rejects_without_writing(gpu):
    """With selection [0, 1] the C2R stage is followed by a real copy stage
    (axis 2); a rejected spectrum must leave dst untouched all the same."""
    shape = (6, 8, 5)
    rng = np.random.default_rng(8)
    plan = pf.R2cPlan(shape, dtype=torch.float64, selection=[0, 1])
    x = uniform_real(rng, shape, np.float64).reshape(-1)
    spec = _host(plan.forward(_tdev(x))).copy()
    _, out_shape, _, _ = plan.layout(1)
    # memory [axis2][axis1][axis0]; element 1: axis1 = 0 (the DC plane of the reduced axis 1)
    spec[1] += 0.5j
    dst = torch.full((x.size,), -1.0, dtype=torch.float64, device="cuda")
    with pytest.raises(pf.PfhError) as e:
        plan.inverse(_tdev(spec), dst)
    assert e.value.variant == "InvalidSpectrum"
    assert torch.all(dst == -1.0)
    with pytest.raises(O.InvalidSpectrum):
        O.transform(O.KIND_R2C, spec, O.INVERSE, shape=shape, selected=[0, 1])


R2R_CASES = [
    ((3, 4), 1, [pf.R2rKind.Dht, pf.R2rKind.Dht]),
    ((3, 2, 4), 6, [pf.R2rKind.Dht, None, pf.R2rKind.Dht]),
    ((2, 3, 2, 3), 1, [pf.R2rKind.Dht, None, None, pf.R2rKind.Dht]),
    ((2, 3, 2, 3), 1, [None, None, None, None]),
    ((2, 3, 2, 3), 1, [pf.R2rKind.DctI, pf.R2rKind.DctII, pf.R2rKind.DctIII, pf.R2rKind.DctIV]),
    ((2, 3, 2, 3), 1, [pf.R2rKind.DstI, pf.R2rKind.DstII, pf.R2rKind.DstIII, pf.R2rKind.DstIV]),
    ((16, 24, 20), 1, [pf.R2rKind.DctII, pf.R2rKind.DstI, pf.R2rKind.DctIV]),
    ((12, 10, 9), 2, [pf.R2rKind.DstII, pf.R2rKind.DctIII, pf.R2rKind.DstIV]),
    # even DHT lines whose half length has a compile-time plan (R2C kernels with a Hartley store)
    ((128, 96, 40), 1, [pf.R2rKind.Dht, pf.R2rKind.Dht, pf.R2rKind.Dht]),
    ((6, 128), 3, [pf.R2rKind.Dht, pf.R2rKind.Dht]),
    ((128, 5, 96), 1, [pf.R2rKind.Dht, None, pf.R2rKind.Dht]),
    # R2R kinds whose FFT length has a compile-time plan (first pass gathers the pre-processed values)
    ((128, 120, 63), 1, [pf.R2rKind.DctII, pf.R2rKind.DctIV, pf.R2rKind.DstI]),
    ((120, 64, 128), 2, [pf.R2rKind.DstII, pf.R2rKind.DctIII, pf.R2rKind.DstIV]),
    ((65, 6, 128), 1, [pf.R2rKind.DctI, None, pf.R2rKind.DctIII]),
    ((128, 127, 2), 1, [pf.R2rKind.DstIII, pf.R2rKind.DstI, None]),
    # 385-point FFTs (the 11 x 7 x 5 compile-time plan in f32): odd-length
    # DCT-II / DCT-III / DST-II / DST-III / DHT lines of length 385, DST-I of
    # length 384 and DCT-I of length 386
    ((385, 4, 6), 1, [pf.R2rKind.DctII, None, pf.R2rKind.DstIII]),
    ((6, 385, 4), 1, [pf.R2rKind.DctIII, pf.R2rKind.DstII, None]),
    ((4, 6, 385), 1, [None, pf.R2rKind.Dht, pf.R2rKind.Dht]),
    ((384, 3, 386), 1, [pf.R2rKind.DstI, None, pf.R2rKind.DctI]),
    # 384-point DST-I lines: an odd number of lines, a middle axis, an extra
    # batch
    ((384, 3, 5), 1, [pf.R2rKind.DstI, pf.R2rKind.DctII, None]),
    ((5, 384, 7), 1, [None, pf.R2rKind.DstI, None]),
    ((3, 384), 2, [pf.R2rKind.DstII, pf.R2rKind.DstI]),
    # the compile-time DCT-II (n = 256) and DCT-IV (n = 240) kernels, whose
    # loads scatter each value to its FFT-input slot, under the input
    # reversal / alternating sign of DST-II / DST-IV and the edge weights of
    # the DCT-III / DST-III adjoints
    ((256, 3, 4), 1, [pf.R2rKind.DstII, None, None]),
    ((4, 240, 3), 1, [None, pf.R2rKind.DstIV, None]),
    ((3, 4, 256), 2, [None, None, pf.R2rKind.DctIII]),
    ((256, 2, 3), 1, [pf.R2rKind.DstIII, None, pf.R2rKind.DctIV]),
]


@pytest.mark.parametrize("permute", [True, False])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("shape,extra,kinds", R2R_CASES)
def test_r2r_plan_matches_oracle(gpu, shape, extra, kinds, dt, permute):
    x = fixture_field(shape, extra, 3.0, False).astype(dt)
    plan = pf.R2rPlan(shape, kinds, dtype=torch.from_numpy(x).dtype, extra_shape=(extra,) if extra > 1 else (),
                      permute_dims=permute)
    r2r = [k or 0 for k in kinds]
    y = O.transform(O.KIND_R2R, x.astype(np.float64), O.FORWARD, shape=shape, extra=extra, r2r=r2r,
                    permute_dims=permute).astype(dt)
    run_all_ops(plan, O.KIND_R2R, shape, extra, x.reshape(-1), y, dt, r2r=r2r, permute_dims=permute)


def _coded(shape, extra, dtype):
    """Integer-coded values x*10000 + y*100 + z (+ extra * 1e6)."""
    idx = np.indices(shape).reshape(len(shape), -1)
    code = np.zeros(idx.shape[1], dtype=np.int64)
    for a in range(len(shape)):
        code = code * 100 + idx[a]
    vals = np.concatenate([code + e * 1_000_000 for e in range(extra)])
    return vals.astype(dtype)


@pytest.mark.parametrize("plan_cls", [pf.LocalTransposePlan, pf.AllToAllvTransposePlan])
def test_transpose_all_3d_permutations_bit_exact(gpu, plan_cls):
    shape, extra = (5, 7, 6), 2
    perms = list(itertools.permutations(range(3)))
    base = _coded(shape, extra, np.uint64)
    for sp in perms:
        src = O.local_transpose(base, shape, extra, (0, 1, 2), sp)  # logical -> sp memory order
        for dp in perms:
            exp = O.local_transpose(src, shape, extra, sp, dp)
            if plan_cls is pf.AllToAllvTransposePlan:
                exp2, _ = O.alltoallv_1x1(src, shape, extra, sp, dp)
                assert np.array_equal(exp, exp2)
            plan = plan_cls(shape, sp, dp, 8, extra_shape=(extra,))
            d = torch.empty(src.size, dtype=torch.int64, device="cuda")
            plan.execute(_tdev(src.view(np.int64)), d)
            assert np.array_equal(_host(d).view(np.uint64), exp), (sp, dp)


@pytest.mark.parametrize("elem", [1, 2, 4, 8, 12, 16, 24])
def test_transpose_element_sizes_bit_exact(gpu, elem):
    shape, extra = (3, 4, 5, 6), 1
    rng = np.random.default_rng(elem)
    raw = rng.integers(0, 256, size=np.prod(shape) * elem, dtype=np.uint8)
    for sp, dp in [((0, 1, 2, 3), (3, 2, 1, 0)), ((2, 0, 3, 1), (1, 3, 0, 2)), ((0, 1, 3, 2), (0, 1, 2, 3))]:
        # bytes of element i are raw[i*elem:(i+1)*elem]
        src = raw.copy()
        exp_idx = O.local_transpose(np.arange(np.prod(shape), dtype=np.int64), shape, 1, sp, dp)
        exp = src.reshape(-1, elem)[exp_idx].reshape(-1)
        plan = pf.LocalTransposePlan(shape, sp, dp, elem)
        d = torch.empty(raw.size, dtype=torch.uint8, device="cuda")
        plan.execute(_tdev(src), d)
        assert np.array_equal(_host(d), exp), (elem, sp, dp)


def test_config1_64cubed_c2c_f64_roundtrip_and_transpose(gpu):
    """64^3 complex f64, re/im uniform[-1,1) from a seed: forward, normalized
    inverse, and the LocalTransposePlan z,y,x -> x,z,y permutation."""
    shape = (64, 64, 64)
    rng = np.random.default_rng(1)
    x = uniform_complex(rng, shape, np.complex128)
    plan = pf.C2cPlan(shape, dtype=torch.complex128)
    y = plan.forward(_tdev(x))
    ref = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shape)
    assert rel_l2(_host(y), ref) <= 1e-12
    back = _host(plan.inverse(y))
    assert rel_l2(back, x.reshape(-1)) <= 1e-12
    tp = pf.LocalTransposePlan(shape, (2, 1, 0), (0, 2, 1), 16)
    src = _host(y)
    d = torch.empty_like(y)
    tp.execute(y, d)
    exp = O.local_transpose(src.view(np.complex128), shape, 1, (2, 1, 0), (0, 2, 1))
    assert _host(d).tobytes() == exp.tobytes()


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_in_place_matches_out_of_place(gpu, dt):
    shape = (8, 12, 16)
    rng = np.random.default_rng(11)
    x = uniform_real(rng, shape, dt)
    plan = pf.R2cPlan(shape, dtype=torch.from_numpy(x).dtype)
    ref = _host(plan.forward(_tdev(x)))
    n_out = plan.output_elements()
    data = torch.zeros(n_out, dtype=torch.from_numpy(ref).dtype, device="cuda")
    data.view(torch.from_numpy(x).dtype)[: x.size].copy_(_tdev(x).reshape(-1))
    plan.execute_in_place(O.FORWARD, data)
    assert np.array_equal(_host(data), ref)
    cplan = pf.C2cPlan(shape, dtype=torch.from_numpy(ref).dtype)
    xc = uniform_complex(rng, shape, complex_of(dt))
    refc = _host(cplan.forward(_tdev(xc)))
    d = _tdev(xc.reshape(-1))
    cplan.execute_in_place(O.FORWARD, d)
    assert np.array_equal(_host(d), refc)


def test_empty_extra_shape_is_a_no_op(gpu):
    """ExtraShape [0] (crates/pencil-fft/tests/distributed_c2c.rs:150-156,
    distributed/r2c tests :873-894, distributed_mixed.rs:869-893): plans build,
    and every operation on the empty arrays succeeds without device work."""
    e = torch.empty(0, device="cuda")
    ec = torch.empty(0, dtype=torch.complex64, device="cuda")
    plans = [
        (pf.C2cPlan((3, 1, 4), extra_shape=(0,)), ec, ec),
        (pf.R2cPlan((2, 3, 2), dtype=torch.float32, extra_shape=(0,)), e, ec),
        (pf.C2rPlan((3, 2, 4), dtype=torch.float32, extra_shape=(0,)), ec, e),
        (pf.R2rPlan((3, 4, 5), [pf.R2rKind.DctII, pf.R2rKind.DstI, None], dtype=torch.float32, extra_shape=(0,)),
         e, e),
        (pf.MixedR2cPlan((2, 3), [pf.AxisTransform.FFT, pf.AxisTransform.RFFT], dtype=torch.float32,
                         extra_shape=(0,)), e, ec),
    ]
    for plan, src, out in plans:
        assert plan.input_elements() == 0 and plan.output_elements() == 0
        y = plan.forward(src)
        assert y.numel() == 0
        back = plan.inverse(y)
        assert back.numel() == 0
        if not isinstance(plan, pf.C2rPlan):
            plan.execute_in_place(pf.FORWARD, torch.empty(0, dtype=out.dtype, device="cuda"))
    tp = pf.LocalTransposePlan((2, 3, 4), (0, 1, 2), (2, 1, 0), element_bytes=8, extra_shape=(0,))
    tp.execute(torch.empty(0, dtype=torch.float64, device="cuda"), torch.empty(0, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()


def test_unit_extents_and_ragged_shapes(gpu):
    """Extent-1 axes and odd lengths on every stage (tests/distributed_c2c.rs:130-166
    use [3, 1, 4] and [3, 2, 3, 4]): a 4D C2C f64 plan against the oracle."""
    rng = np.random.default_rng(21)
    for shape, extra in (((3, 1, 4), (2,)), ((3, 2, 3, 4), ()), ((5, 1, 1, 7), (3,))):
        x = uniform_complex(rng, tuple(extra) + shape, np.complex128).reshape(-1)
        plan = pf.C2cPlan(shape, dtype=torch.complex128, extra_shape=extra)
        got = _host(plan.forward(_tdev(x)))
        ref = O.transform(O.KIND_C2C, x, O.FORWARD, shape=shape, extra=int(np.prod(extra, dtype=np.int64)))
        assert rel_l2(got, ref) <= 1e-12
        back = _host(plan.inverse(_tdev(got)))
        assert rel_l2(back, x) <= 1e-12


def test_reference_method_names(gpu):
    """forward / inverse _in_place, _with_timing and _with_overlap under the
    reference's names give the plain operations' results."""
    rng = np.random.default_rng(31)
    shape = (6, 10, 8)
    x = _tdev(uniform_complex(rng, shape, np.complex128).reshape(-1))
    plan = pf.C2cPlan(shape, dtype=torch.complex128)
    y = plan.forward(x)
    assert torch.equal(plan.forward_with_overlap(x), y)
    yt, timing = plan.forward_with_timing(x)
    assert torch.equal(yt, y) and len(timing.entries()) == 3
    d = x.clone()
    plan.forward_in_place(d)
    assert torch.equal(d, y)
    plan.inverse_in_place(d)
    assert rel_l2(_host(d), _host(x)) <= 1e-12
    assert torch.equal(plan.inverse_with_overlap(y), plan.inverse(y))
