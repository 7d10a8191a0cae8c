"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker or the timed CPU
baseline. The product path (``pencil-rs_amd``) never imports it.

Arrays are numpy arrays. Complex data is ``complex128`` / ``complex64``
(interleaved re, im, as in the reference). Plan arrays are laid out
``[extra][memory-ordered spatial axes]``; ``route(ndim, permute_dims)``
returns the per-stage memory orders of the reference's ``build_route``
(crates/pencil-fft/src/distributed.rs:2306-2373).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OK, ERR_ARG, ERR_INVALID_SPECTRUM = 0, 1, 2
DCT1, DCT2, DCT3, DCT4, DST1, DST2, DST3, DST4, DHT = range(1, 10)
KIND_C2C, KIND_R2C, KIND_R2R, KIND_C2R, KIND_MIXED_C2C, KIND_MIXED_R2C = 1, 2, 3, 4, 5, 6
AXIS_FFT, AXIS_RFFT = 10, 11
FORWARD, INVERSE, BACKWARD, ADJ_FORWARD, ADJ_INVERSE, ADJ_BACKWARD = range(6)
MAXD = 4

R2R_BACKWARD_KIND = {DCT1: DCT1, DCT2: DCT3, DCT3: DCT2, DCT4: DCT4, DST1: DST1,
                     DST2: DST3, DST3: DST2, DST4: DST4, DHT: DHT}


class _Desc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_size_t * MAXD),
        ("extra", ctypes.c_size_t),
        ("selected", ctypes.c_uint),
        ("r2r", ctypes.c_int * MAXD),
        ("backward_dir", ctypes.c_int * MAXD),
        ("permute_dims", ctypes.c_int),
    ]


def build():
    """Compiles liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        _lib.pfo_execute_f64.argtypes = [ctypes.POINTER(_Desc), ctypes.c_int, vp, vp, ctypes.c_int]
        _lib.pfo_execute_f32.argtypes = [ctypes.POINTER(_Desc), ctypes.c_int, vp, vp, ctypes.c_int]
        for name in ("pfo_c2c_lines_f64", "pfo_c2c_lines_f32", "pfo_c2c_lines_f80",
                     "pfo_r2c_lines_f64", "pfo_r2c_lines_f32"):
            getattr(_lib, name).argtypes = [vp, vp, sz, sz, ctypes.c_int]
        for name in ("pfo_r2r_lines_f64", "pfo_r2r_lines_f32"):
            getattr(_lib, name).argtypes = [vp, vp, sz, sz, ctypes.c_int, ctypes.c_int]
        _lib.pfo_local_transpose.argtypes = [sz, ctypes.c_int, vp, sz, vp, vp, vp, vp]
        _lib.pfo_alltoallv_1x1.argtypes = [sz, ctypes.c_int, vp, sz, vp, vp, vp, vp, vp]
        for name in ("pfo_endpoint_accept_f64", "pfo_endpoint_accept_f32"):
            getattr(_lib, name).argtypes = [ctypes.c_int, vp, vp, sz, ctypes.c_int, sz, vp,
                                            ctypes.c_double, ctypes.c_double]
        _lib.pfo_build_route.argtypes = [ctypes.c_int, ctypes.c_int, vp]
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.pfo_endpoint_thresholds.argtypes = [ctypes.POINTER(_Desc), ctypes.c_int, ctypes.c_int, dp, dp, dp]
        _lib.pfo_mixed_axis_model.argtypes = [ctypes.c_int, sz, ctypes.POINTER(sz), ctypes.POINTER(sz),
                                              ctypes.POINTER(ctypes.c_int)]
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _real_dtype(dtype) -> np.dtype:
    d = np.dtype(dtype)
    return np.dtype(np.float32) if d in (np.float32, np.complex64) else np.dtype(np.float64)


def _sfx(dtype) -> str:
    return "f32" if _real_dtype(dtype) == np.float32 else "f64"


def route(ndim: int, permute_dims: bool = True):
    perm = (ctypes.c_int * (MAXD * MAXD))()
    lib().pfo_build_route(ndim, int(permute_dims), ctypes.cast(perm, ctypes.c_void_p))
    return [tuple(perm[s * MAXD + p] for p in range(ndim)) for s in range(ndim)]


# ---------------------------------------------------------------- lines --

def c2c_lines(x: np.ndarray, op: int = FORWARD, extended: bool = False) -> np.ndarray:
    """Batched contiguous C2C lines (last axis). extended=True computes in
    long double (complex128 storage)."""
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    y = np.empty_like(x)
    if extended:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.empty_like(x)
        st = lib().pfo_c2c_lines_f80(_ptr(x), _ptr(y), n, batch, op)
    else:
        st = getattr(lib(), "pfo_c2c_lines_" + _sfx(x.dtype))(_ptr(x), _ptr(y), n, batch, op)
    if st != OK:
        raise ValueError(f"oracle c2c_lines status {st}")
    return y


def r2c_lines(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n
    cd = np.complex64 if x.dtype == np.float32 else np.complex128
    y = np.empty(x.shape[:-1] + (n // 2 + 1,), dtype=cd)
    st = getattr(lib(), "pfo_r2c_lines_" + _sfx(x.dtype))(_ptr(x), _ptr(y), n, batch, FORWARD)
    if st != OK:
        raise ValueError(f"oracle r2c_lines status {st}")
    return y


def c2r_lines(x: np.ndarray, n: int, op: int = INVERSE) -> np.ndarray:
    x = np.ascontiguousarray(x)
    batch = x.size // (n // 2 + 1)
    rd = _real_dtype(x.dtype)
    y = np.empty(x.shape[:-1] + (n,), dtype=rd)
    st = getattr(lib(), "pfo_r2c_lines_" + _sfx(x.dtype))(_ptr(x), _ptr(y), n, batch, op)
    if st != OK:
        raise ValueError(f"oracle c2r_lines status {st}")
    return y


def r2r_lines(x: np.ndarray, kind: int, op: int = FORWARD) -> np.ndarray:
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n if n else 0
    y = np.empty_like(x)
    st = getattr(lib(), "pfo_r2r_lines_" + _sfx(x.dtype))(_ptr(x), _ptr(y), n, batch, kind, op)
    if st != OK:
        raise ValueError(f"oracle r2r_lines status {st}")
    return y


# ---------------------------------------------------------------- plans --

class InvalidSpectrum(Exception):
    """The oracle rejected a C2R boundary (R2cError::InvalidSpectrum)."""


def make_desc(kind, shape, extra=1, selected=None, r2r=None, backward_dir=None, permute_dims=True):
    d = _Desc()
    d.kind = kind
    d.ndim = len(shape)
    for a, n in enumerate(shape):
        d.shape[a] = n
    d.extra = extra
    if selected is None:
        selected = range(len(shape))
    mask = 0
    for a in selected:
        mask |= 1 << a
    d.selected = mask
    for a in range(len(shape)):
        d.r2r[a] = (r2r[a] or 0) if r2r is not None else 0
        d.backward_dir[a] = int(bool(backward_dir[a])) if backward_dir is not None else 0
    d.permute_dims = int(bool(permute_dims))
    return d


def plan_shapes(kind, shape, selected=None, r2r=None):
    """(input logical shape, output logical shape) of the forward operation
    (for KIND_C2R the forward operation is the C2R: complex in, real out;
    for KIND_MIXED_R2C the reduced axis is the AXIS_RFFT one)."""
    shape = tuple(shape)
    if kind == KIND_MIXED_R2C:
        red = list(r2r).index(AXIS_RFFT)
        half = list(shape)
        half[red] = shape[red] // 2 + 1
        return shape, tuple(half)
    if kind not in (KIND_R2C, KIND_C2R):
        return shape, shape
    sel = list(range(len(shape))) if selected is None else list(selected)
    red = max(sel) if kind == KIND_R2C else min(sel)
    half = list(shape)
    half[red] = shape[red] // 2 + 1
    return (shape, tuple(half)) if kind == KIND_R2C else (tuple(half), shape)


def execute(desc: _Desc, op: int, src: np.ndarray, out_dtype, out_count: int, threads: int = 0):
    src = np.ascontiguousarray(src)
    dst = np.empty(out_count, dtype=out_dtype)
    fn = lib().pfo_execute_f32 if _real_dtype(out_dtype) == np.float32 else lib().pfo_execute_f64
    st = fn(ctypes.byref(desc), op, _ptr(src), _ptr(dst), threads)
    if st == ERR_INVALID_SPECTRUM:
        raise InvalidSpectrum()
    if st != OK:
        raise ValueError(f"oracle execute status {st}")
    return dst


def transform(kind, x: np.ndarray, op=FORWARD, shape=None, extra=1, selected=None, r2r=None,
              backward_dir=None, permute_dims=True, threads=0):
    """Runs one plan operation on a flat array laid out as the plan's input for
    `op` and returns the flat output array."""
    assert shape is not None
    desc = make_desc(kind, shape, extra, selected, r2r, backward_dir, permute_dims)
    real = _real_dtype(x.dtype)
    cd = np.complex64 if real == np.float32 else np.complex128
    shp_in, shp_out = plan_shapes(kind, shape, selected, r2r)
    forward_like = op in (FORWARD, ADJ_INVERSE, ADJ_BACKWARD)
    if kind in (KIND_C2C, KIND_MIXED_C2C):
        count, dt = int(np.prod(shape)) * extra, cd
    elif kind == KIND_R2R:
        count, dt = int(np.prod(shape)) * extra, real
    elif kind in (KIND_R2C, KIND_MIXED_R2C):
        if forward_like:
            count, dt = int(np.prod(shp_out)) * extra, cd
        else:
            count, dt = int(np.prod(shp_in)) * extra, real
    else:  # C2R orientation: the forward operation is complex -> real
        if forward_like:
            count, dt = int(np.prod(shp_out)) * extra, real
        else:
            count, dt = int(np.prod(shp_in)) * extra, cd
    return execute(desc, op, x.reshape(-1), dt, count, threads)


# ------------------------------------------------------------- layouts --

def local_transpose(x: np.ndarray, shape, extra, src_perm, dst_perm) -> np.ndarray:
    x = np.ascontiguousarray(x).reshape(-1)
    y = np.empty_like(x)
    nd = len(shape)
    shp = (ctypes.c_size_t * MAXD)(*shape)
    sp = (ctypes.c_int * MAXD)(*src_perm)
    dp = (ctypes.c_int * MAXD)(*dst_perm)
    st = lib().pfo_local_transpose(x.itemsize, nd, ctypes.cast(shp, ctypes.c_void_p), extra,
                                   ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(dp, ctypes.c_void_p),
                                   _ptr(x), _ptr(y))
    if st != OK:
        raise ValueError("oracle local_transpose failed")
    return y


def alltoallv_1x1(x: np.ndarray, shape, extra, src_perm, dst_perm):
    """Returns (destination, payload) of the one-peer Alltoallv transpose."""
    x = np.ascontiguousarray(x).reshape(-1)
    y = np.empty_like(x)
    pay = np.empty_like(x)
    nd = len(shape)
    shp = (ctypes.c_size_t * MAXD)(*shape)
    sp = (ctypes.c_int * MAXD)(*src_perm)
    dp = (ctypes.c_int * MAXD)(*dst_perm)
    st = lib().pfo_alltoallv_1x1(x.itemsize, nd, ctypes.cast(shp, ctypes.c_void_p), extra,
                                 ctypes.cast(sp, ctypes.c_void_p), ctypes.cast(dp, ctypes.c_void_p),
                                 _ptr(x), _ptr(y), _ptr(pay))
    if st != OK:
        raise ValueError("oracle alltoallv failed")
    return y, pay


def endpoint_accept(data: np.ndarray, perm, rshape, extra, axis, real_len, relative, absolute) -> bool:
    data = np.ascontiguousarray(data)
    nd = len(rshape)
    shp = (ctypes.c_size_t * MAXD)(*rshape)
    pp = (ctypes.c_int * MAXD)(*perm)
    fn = lib().pfo_endpoint_accept_f32 if data.dtype == np.complex64 else lib().pfo_endpoint_accept_f64
    return bool(fn(nd, ctypes.cast(pp, ctypes.c_void_p), ctypes.cast(shp, ctypes.c_void_p), extra, axis,
                   real_len, _ptr(data), relative, absolute))


def _min_subnormal(dtype) -> float:
    """The smallest positive subnormal of the real type, as f64
    (f32::from_bits(1) / f64::from_bits(1), lib.rs:264-280)."""
    bits = np.array([1], dtype=np.uint32 if _real_dtype(dtype) == np.float32 else np.uint64)
    return float(bits.view(np.float32 if _real_dtype(dtype) == np.float32 else np.float64)[0])


def mixed_axis_model(code: int, n: int):
    """(E_a, factor_a, allowance_a) of one MixedR2cPlan complex-suffix axis
    (mixed.rs:408-417): FFT (n, n, 0); DCT-I 2(n-1), DST-I 2(n+1), types II/III
    4n, type IV 8n with the R2R normalization factor (2(n-1), 2(n+1), 2n) and
    allowance 2; DHT (n, n, 2). None for the identity."""
    if code == AXIS_FFT:
        return n, n, 0
    if code == DCT1:
        return 2 * (n - 1), 2 * (n - 1), 2
    if code == DST1:
        return 2 * (n + 1), 2 * (n + 1), 2
    if code in (DCT2, DCT3, DST2, DST3):
        return 4 * n, 2 * n, 2
    if code in (DCT4, DST4):
        return 8 * n, 2 * n, 2
    if code == DHT:
        return n, n, 2
    return None


def endpoint_thresholds(dtype, shape, selected, reduction_axis, normalize, mixed=None):
    """(relative, absolute) acceptance thresholds of the C2R boundary check
    (normalization_depth / raw_absolute_threshold_for_shape,
    distributed/r2c.rs:4841-4878; for a MixedR2cPlan pass its per-axis codes
    as mixed: mixed_depth_normalization, mixed.rs:4085-4117)."""
    eps = float(np.finfo(_real_dtype(dtype)).eps)
    minsub = _min_subnormal(dtype)
    depth, transverse = 1.0, 1.0
    if mixed is not None:
        for a in range(reduction_axis):
            m = mixed_axis_model(mixed[a], shape[a])
            if m is None:
                continue
            E, factor, allowance = m
            depth += (E - 1).bit_length() + allowance
            transverse *= float(factor)
    else:
        for a, n in enumerate(shape):
            if a != reduction_axis and a in selected:
                if n > 1:
                    depth += (n - 1).bit_length()
                transverse *= float(n)
    relative = 128.0 * eps * depth
    absolute = 128.0 * minsub * depth
    if not normalize:
        absolute *= transverse
    return relative, absolute


def endpoint_thresholds_c(kind, dtype, shape, op, selected=None, r2r=None):
    """The C restatement's thresholds for plan kind and route op (INVERSE or
    BACKWARD): (relative, absolute, depth)."""
    desc = make_desc(kind, shape, 1, selected, r2r, None, True)
    rel, ab, dep = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    st = lib().pfo_endpoint_thresholds(ctypes.byref(desc), 4 if _real_dtype(dtype) == np.float32 else 8, op,
                                       ctypes.byref(rel), ctypes.byref(ab), ctypes.byref(dep))
    if st != OK:
        raise ValueError("oracle thresholds: no real boundary")
    return rel.value, ab.value, dep.value


def mixed_axis_model_c(code: int, n: int):
    E, f, al = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_int()
    if not lib().pfo_mixed_axis_model(code, n, ctypes.byref(E), ctypes.byref(f), ctypes.byref(al)):
        return None
    return E.value, f.value, al.value
