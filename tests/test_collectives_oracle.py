"""CPU tests that pin the reduction / pointwise oracle (oracle/collectives.py)
and the C ABI's validation of those calls (no GPU: every call below fails
validation before any device work, or only loads the library).

The oracle is pinned by the reference's own known answers:
crates/pencil-array/tests/collectives.rs (the regular [5, 7] array with
extras [2, 3] under a permuted pencil :42-73, :882-951; the i8 / u8 overflow
cases :262-298; the complex witness, per-component NaN and opposing-infinity
cases and the finite-large norm :365-438; empty extras :920-931), the
NormState unit tests (src/collectives.rs:2937-3019), the checked-sum example
of the module documentation (src/collectives.rs:14-17), and the broadcast and
error-precedence tests of tests/array_ops.rs:10-96, :356-425; plus closed
forms (arithmetic series, Pythagorean triples) and brute force.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

import collectives as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------- reductions --
def regular_expected():
    """regular_expected (tests/collectives.rs:63-73), logical order."""
    return np.array([e0 * 100_000 + e1 * 10_000 + x * 100 + y
                     for e0 in range(2) for e1 in range(3) for x in range(5) for y in range(7)], dtype=np.int64)


def regular_memory_order():
    """The same values in the storage order of the [1, 0]-permuted pencil
    (memory [extra][y][x])."""
    v = regular_expected().reshape(2, 3, 5, 7)
    return np.ascontiguousarray(v.transpose(0, 1, 3, 2)).reshape(-1)


def test_regular_array_reductions():
    mem = regular_memory_order()
    exp = regular_expected()
    assert C.global_sum(mem) == exp.sum()
    assert C.global_min(mem) == exp.min()
    assert C.global_max(mem) == exp.max()
    expected_norm = math.sqrt(sum(float(v) * float(v) for v in exp))
    assert abs(C.l2_norm(mem) - expected_norm) < 1e-6
    empty = np.zeros(0, dtype=np.int64)
    assert C.global_sum(empty) == 0
    assert C.global_min(empty) is None and C.global_max(empty) is None
    assert C.l2_norm(empty) == 0.0
    assert C.all_(empty) is True and C.any_(empty) is False
    # any_by(v < 0) is false and all_by(v >= 0) true on these values
    assert C.any_(mem < 0) is False and C.all_(mem >= 0) is True


def test_checked_integer_sum():
    # tests/collectives.rs:270-279: one rank's 35 values of i8::MAX overflow
    with pytest.raises(C.IntegerOverflow):
        C.global_sum(np.full(35, 127, dtype=np.int8))
    # the module documentation's example: [127, 1, -1] fails at the second value
    with pytest.raises(C.IntegerOverflow):
        C.global_sum(np.array([127, 1, -1], dtype=np.int8))
    # the same values in another order never leave the range
    assert C.global_sum(np.array([127, -1, 1], dtype=np.int8)) == 127
    # tests/collectives.rs:283-297 at one rank: a single 200u8 is fine
    assert C.global_sum(np.array([200], dtype=np.uint8)) == 200
    with pytest.raises(C.IntegerOverflow):
        C.global_sum(np.array([200, 56], dtype=np.uint8))
    assert C.global_sum(np.array([200, 55], dtype=np.uint8)) == 255
    # brute force against exact prefix sums on random i16 data
    rng = np.random.default_rng(0)
    for _ in range(50):
        v = rng.integers(-20000, 20000, size=8, dtype=np.int16)
        pref = np.cumsum(v.astype(np.int64))
        if pref.min() < -32768 or pref.max() > 32767:
            with pytest.raises(C.IntegerOverflow):
                C.global_sum(v)
        else:
            assert C.global_sum(v) == pref[-1]


def test_complex_sum_rules():
    n = 35
    w = np.zeros(n, dtype=np.complex64)
    w[0] = complex(np.inf, -np.inf)
    r = C.global_sum(w)
    assert np.isposinf(r.real) and np.isneginf(r.imag)
    v = np.zeros(n, dtype=np.complex64)
    v[0] = complex(np.nan, 3.0)
    r = C.global_sum(v)
    assert np.isnan(r.real) and r.imag == 3.0
    o = np.full(n, complex(np.inf, 0.0), dtype=np.complex64)
    r = C.global_sum(o)
    assert np.isposinf(r.real)
    # opposite input infinities give NaN
    o[1] = complex(-np.inf, 0.0)
    assert np.isnan(C.global_sum(o).real)
    # finite values whose norm overflows f32
    big = np.full(n, complex(np.finfo(np.float32).max, np.finfo(np.float32).max), dtype=np.complex64)
    assert np.isinf(C.l2_norm(big))


def test_float_sum_closed_form_and_flags():
    n = 1000
    v = np.arange(n, dtype=np.float64)
    assert C.global_sum(v) == n * (n - 1) / 2
    assert C.global_sum(v.astype(np.float32)) == np.float32(n * (n - 1) / 2)
    assert np.isnan(C.global_sum(np.array([1.0, np.nan, np.inf])))
    assert np.isneginf(C.global_sum(np.array([1.0, -np.inf])))
    assert C.global_min(np.array([3.0, -2.0, 5.0])) == -2.0
    assert np.isnan(C.global_max(np.array([3.0, np.nan, 5.0])))
    assert C.global_max(np.array([3.0, np.inf])) == np.inf


def test_norm_known_answers():
    assert abs(C.norm_state([1.0, 2.0, 3.0, 4.0]) - math.sqrt(30)) / math.sqrt(30) < 1e-12
    assert abs(C.norm_state([4.0, 3.0, 2.0, 1.0]) - math.sqrt(30)) / math.sqrt(30) < 1e-12
    for vals, exp in (([3e300, 4e300], 5e300), ([3e-300, 4e-300], 5e-300)):
        for f in (C.norm_state, lambda v: float(C.l2_norm(np.array(v)))):
            r = f(vals)
            assert math.isfinite(r) and r != 0.0 and abs(r - exp) / exp < 1e-12
    q = np.float64(5e-324)
    for vals in ([3 * q, 4 * q], [4 * q, 3 * q]):
        assert C.norm_state(vals) == 5 * q
        assert C.l2_norm(np.array(vals)) == 5 * q
    assert C.norm_state([0.0, -0.0]) == 0.0
    assert C.norm_state([complex(3.0, 4.0)]) == 5.0
    assert C.l2_norm(np.array([3 + 4j])) == 5.0
    assert np.isnan(C.l2_norm(np.array([1.0, np.nan, 2.0])))
    assert np.isposinf(C.l2_norm(np.array([1.0, -np.inf, 2.0])))
    assert C.l2_norm(np.zeros(4, dtype=np.float32)).dtype == np.float32


def test_truth():
    assert C.any_(np.array([0.0, np.nan])) is True
    assert C.all_(np.array([1.0, np.nan])) is True
    assert C.all_(np.array([1 + 0j, 0j])) is False
    assert C.any_(np.array([0j, 1j])) is True


# ---------------------------------------------------------- broadcast ops --
def test_broadcast_known_answer():
    """tests/array_ops.rs:10-44: a (extra [1, 3]) + b (extra [2, 1]) + 7 into
    extra [2, 3]."""
    p = C.Pencil((4, 3))
    n = p.local_len()
    a = np.arange(3 * n, dtype=np.int64)
    b = 100 + np.arange(2 * n, dtype=np.int64)
    out = C.pointwise_many([a, b], [(1, 3), (2, 1)], (2, 3), n, lambda xs: xs[0] + xs[1] + 7)
    for i in range(2):
        for j in range(3):
            for k in range(n):
                assert out[(i * 3 + j) * n + k] == a[j * n + k] + b[i * n + k] + 7
    # empty output extras visit nothing
    assert C.pointwise_many([a[:0]], [(0,)], (0,), n, lambda xs: 1 / 0).size == 0


def test_error_precedence():
    """tests/array_ops.rs:356-425 and the many-input order."""
    p = C.Pencil((4, 3), decomposition=(1,))
    other = C.Pencil((4, 3), decomposition=(0,))
    with pytest.raises(C.PointwiseError) as e:
        C.validate_pointwise2((p, (2, 2)), (other, (2, 3)), (p, (2, 3)))
    assert e.value.variant == "IncompatiblePencils"
    with pytest.raises(C.PointwiseError) as e:
        C.validate_pointwise2((p, (2,)), (p, (2, 2)), (p, (2, 3)))
    assert e.value.variant == "ExtraRankMismatch" and e.value.fields == {"expected": 2, "actual": 1}
    with pytest.raises(C.PointwiseError) as e:
        C.validate_pointwise2((p, (2, 3)), (p, (2, 2)), (p, (2, 3)))
    assert e.value.variant == "ExtraExtentMismatch" and e.value.fields == {"axis": 1, "output": 3, "input": 2}
    with pytest.raises(C.PointwiseError) as e:
        C.validate_many([], (p, (2, 3)))
    assert e.value.variant == "EmptyInputs"
    with pytest.raises(C.PointwiseError) as e:
        C.validate_many([(p, (1, 3)), (p, (2, 1)), (other, (2, 3))], (p, (2, 3)))
    assert e.value.variant == "IncompatiblePencils" and e.value.index == 2
    perm = C.Pencil((4, 3), perm=(1, 0), decomposition=(1,))
    with pytest.raises(C.PointwiseError) as e:
        C.validate_many([(p, (2, 3)), (perm, (2, 3))], (p, (2, 3)))
    assert e.value.index == 1


# ------------------------------------------------ C ABI (validation only) --
LIB = os.path.join(ROOT, "pencil-rs_amd", "libpencilfft_hip.so")


@pytest.fixture(scope="module")
def pf():
    import pencil_fft_hip as pf

    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "pencil-rs_amd"), "-j8"], check=True)
    pf.load_library()
    return pf


def test_struct_layouts_match_the_header(pf, tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "pencil_fft_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(pfh_array), sizeof(pfh_pencil_layout),"
        " sizeof(pfh_reduce_result), offsetof(pfh_array, pencil), offsetof(pfh_pencil_layout, decomposition),"
        " offsetof(pfh_reduce_result, outcome)); return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [ctypes.sizeof(pf._Array), ctypes.sizeof(pf._PencilLayout), 32, pf._Array.pencil.offset,
                   pf._PencilLayout.decomposition.offset, 16]


def _fake(pf, shape, extra, perm=None, decomposition=(), dtype_code=9, ptr=0x1000):
    a = pf._Array()
    a.data = ptr
    a.scalar_type = dtype_code
    a.extra_ndim = len(extra)
    for i, v in enumerate(extra):
        a.extra_shape[i] = v
    a.pencil = pf.PencilLayout(shape, perm, decomposition)._c()
    return a


def _call(pf, op, inputs, out, ncoef=3):
    arr = (pf._Array * max(1, len(inputs)))(*inputs)
    co = (ctypes.c_double * (2 * ncoef))(*([1.0, 0.0] * ncoef))
    return pf._lib.pfh_pointwise(op, arr, len(inputs), ctypes.byref(out), co, None)


def _err(pf):
    info = pf._ErrorInfo()
    pf._lib.pfh_last_error(ctypes.byref(info))
    return info


def test_pointwise_validation_precedence_through_the_abi(pf):
    sh = (4, 3)
    out = _fake(pf, sh, (2, 3), decomposition=(1,), ptr=0x9000)
    left_other = _fake(pf, sh, (2, 2), decomposition=(0,))
    st = _call(pf, pf.POINTWISE_MUL, [_fake(pf, sh, (2, 2), decomposition=(1,)),
                                      _fake(pf, sh, (2, 3), decomposition=(0,))], out)
    assert pf.STATUS_NAMES[st] == "IncompatiblePencils"
    st = _call(pf, pf.POINTWISE_MUL, [_fake(pf, sh, (2,), decomposition=(1,)),
                                      _fake(pf, sh, (2, 2), decomposition=(1,))], out)
    info = _err(pf)
    assert pf.STATUS_NAMES[st] == "ExtraRankMismatch" and (info.required, info.actual) == (2, 1)
    st = _call(pf, pf.POINTWISE_MUL, [_fake(pf, sh, (2, 3), decomposition=(1,)),
                                      _fake(pf, sh, (2, 2), decomposition=(1,))], out)
    info = _err(pf)
    assert pf.STATUS_NAMES[st] == "ExtraExtentMismatch"
    assert (info.index, info.required, info.actual, info.batch) == (1, 3, 2, 1)
    st = _call(pf, pf.POINTWISE_LINEAR, [], out)
    assert pf.STATUS_NAMES[st] == "EmptyInputs"
    st = _call(pf, pf.POINTWISE_MUL_ADD, [_fake(pf, sh, (1, 3), decomposition=(1,)),
                                          _fake(pf, sh, (2, 1), decomposition=(1,)), left_other], out)
    info = _err(pf)
    assert pf.STATUS_NAMES[st] == "IncompatiblePencils" and info.batch == 2
    # types: a complex input needs a complex output; precisions must agree
    st = _call(pf, pf.POINTWISE_LINEAR, [_fake(pf, sh, (2, 3), decomposition=(1,), dtype_code=11)], out)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"
    st = _call(pf, pf.POINTWISE_LINEAR, [_fake(pf, sh, (2, 3), decomposition=(1,), dtype_code=10)], out)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"
    # an input that is the output must not broadcast
    st = _call(pf, pf.POINTWISE_LINEAR, [_fake(pf, sh, (1, 3), decomposition=(1,), ptr=0x9000)], out)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"
    # an empty output extra shape visits nothing and succeeds without a device
    empty = _fake(pf, sh, (0,), decomposition=(1,), ptr=0x9000)
    st = _call(pf, pf.POINTWISE_MUL, [_fake(pf, sh, (0,), decomposition=(1,)), _fake(pf, sh, (1,), decomposition=(1,))],
               empty)
    assert st == 0


def test_reduce_argument_checks(pf):
    r = ctypes.c_void_p(0x2000)
    ws = ctypes.c_void_p(0x3000)
    wsb = pf.reduce_workspace_bytes()
    assert wsb >= 1024 * 32
    st = pf._lib.pfh_reduce(pf.REDUCE_MIN, 11, ctypes.c_void_p(0x1000), 4, ws, wsb, r, None)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"  # complex values have no order
    st = pf._lib.pfh_reduce(9, 9, ctypes.c_void_p(0x1000), 4, ws, wsb, r, None)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"
    st = pf._lib.pfh_reduce(pf.REDUCE_SUM, 13, ctypes.c_void_p(0x1000), 4, ws, wsb, r, None)
    assert pf.STATUS_NAMES[st] == "InvalidArgument"
    st = pf._lib.pfh_reduce(pf.REDUCE_SUM, 9, ctypes.c_void_p(0x1000), 4, ws, 100, r, None)
    info = _err(pf)
    assert pf.STATUS_NAMES[st] == "WorkspaceTooSmall" and (info.required, info.actual) == (wsb, 100)
