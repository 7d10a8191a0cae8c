"""CPU restatement of the reference's local reductions and broadcast
pointwise operations at one rank (TEST INFRASTRUCTURE ONLY).

Only tests/ import this module, as the checker of the device reductions and
pointwise combinations; the product path never does. Each function follows
the reference's algorithm step by step (file:line in
crates/pencil-array/src/), in the element type the reference folds in, with
plain Python loops: it is meant for the small arrays the tests use.
"""
from __future__ import annotations

import math
from typing import List, Optional, Sequence

import numpy as np


class IntegerOverflow(Exception):
    """CollectiveError::IntegerOverflow."""


class PointwiseError(Exception):
    """PointwiseError / MultiInputError variants (pointwise.rs:8-33, :216-230)."""

    def __init__(self, variant: str, index: Optional[int] = None, **fields):
        self.variant, self.index, self.fields = variant, index, fields
        super().__init__(f"{variant} {index} {fields}")


_INT_LIMITS = {np.dtype(t): (int(np.iinfo(t).min), int(np.iinfo(t).max))
               for t in (np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64)}


def _flags(values: np.ndarray) -> List[int]:
    """values_nonfinite_flags (collectives.rs:300-330, complex :420-470):
    [nan, +inf, -inf] of the real part, then of the imaginary part."""
    v = np.asarray(values)
    out = [0] * 6
    parts = [v.real, v.imag] if np.iscomplexobj(v) else [v]
    for c, p in enumerate(parts):
        if p.dtype.kind == "f":
            out[3 * c] = int(np.isnan(p).any())
            out[3 * c + 1] = int(np.isposinf(p).any())
            out[3 * c + 2] = int(np.isneginf(p).any())
    return out


def _sum_component(value, nan: int, pinf: int, ninf: int, dtype):
    """sum_component (collectives.rs:404-420): the raw input flags decide a
    non-finite result; otherwise the folded partial stands."""
    if nan or (pinf and ninf):
        return dtype.type(np.nan)
    if pinf:
        return dtype.type(np.inf)
    if ninf:
        return dtype.type(-np.inf)
    return value


def global_sum(values: np.ndarray):
    """global_sum (collectives.rs:519-528) at one rank: local_sum is a
    storage-order fold from zero in the element type (integers checked at
    every step, :206-214; floats :300-310, complex :380-395), then the
    collective non-finite rules."""
    v = np.asarray(values).reshape(-1)
    dt = v.dtype
    if dt.kind in "iu":
        lo, hi = _INT_LIMITS[dt]
        s = 0
        for x in v.tolist():
            s += x
            if s < lo or s > hi:
                raise IntegerOverflow()
        return dt.type(s)
    with np.errstate(over="ignore", invalid="ignore"):
        if dt.kind == "c":
            rdt = np.dtype(np.float32) if dt == np.complex64 else np.dtype(np.float64)
            re, im = rdt.type(0.0), rdt.type(0.0)
            for x in v:
                re = rdt.type(re + x.real)
                im = rdt.type(im + x.imag)
            f = _flags(v)
            return dt.type(complex(_sum_component(re, f[0], f[1], f[2], rdt), _sum_component(im, f[3], f[4], f[5], rdt)))
        s = dt.type(0.0)
        for x in v:
            s = dt.type(s + x)
        f = _flags(v)
        return _sum_component(s, f[0], f[1], f[2], dt)


def _extreme(values: np.ndarray, minimum: bool):
    """mapped_extreme_layout (collectives.rs:2066-2094) with local_extreme
    (:315-340): the fold skips NaN and starts at +-inf (integers at MAX /
    MIN); any NaN input makes the result NaN; an empty array gives None."""
    v = np.asarray(values).reshape(-1)
    if v.size == 0:
        return None
    dt = v.dtype
    if dt.kind in "iu":
        lo, hi = _INT_LIMITS[dt]
        r = hi if minimum else lo
        for x in v.tolist():
            r = min(r, x) if minimum else max(r, x)
        return dt.type(r)
    r = dt.type(np.inf if minimum else -np.inf)
    nan = False
    for x in v:
        if np.isnan(x):
            nan = True
            continue
        r = min(r, x) if minimum else max(r, x)
    return dt.type(np.nan) if nan else r


def global_min(values):
    return _extreme(values, True)


def global_max(values):
    return _extreme(values, False)


def _norm_abs(x) -> float:
    """norm_abs (collectives.rs:245, :336, :433, :486): |x| in f64, hypot of
    the parts for complex values."""
    if isinstance(x, (complex, np.complexfloating)):
        return math.hypot(float(x.real), float(x.imag))
    return abs(float(x))


def l2_norm(values: np.ndarray):
    """l2_norm through norm_values (collectives.rs:2210-2250): NaN flags give
    NaN, infinity flags (or an infinite maximum) +inf, a zero maximum or an
    empty array 0, else max * sqrt(sum (|x| / max)^2) in f64, rounded to the
    norm type (f32 for f32 and complex f32, f64 otherwise)."""
    v = np.asarray(values).reshape(-1)
    out_t = np.float32 if v.dtype in (np.float32, np.complex64) else np.float64
    f = _flags(v)
    if f[0] or f[3]:
        return out_t(np.nan)
    mags = [_norm_abs(x) for x in v]
    gmax = max(mags, default=0.0)
    if f[1] or f[2] or f[4] or f[5] or math.isinf(gmax):
        return out_t(np.inf)
    if gmax == 0.0 or v.size == 0:
        return out_t(0.0)
    s = 0.0
    for a in mags:
        r = a / gmax
        s = s + r * r
    with np.errstate(over="ignore"):
        return out_t(gmax * math.sqrt(s))


def norm_state(values: Sequence) -> float:
    """NormState push (collectives.rs:2150-2185): the one-pass scaled norm
    the mapped norms use; returns scale * sqrt(ssq)."""
    scale, ssq = 0.0, 0.0
    for x in values:
        a = _norm_abs(x)
        if math.isinf(a):
            scale = math.inf
            continue
        if a == 0.0 or math.isnan(a) or math.isinf(scale):
            continue
        if a > scale:
            ratio = scale / a
            ssq = 1.0 + (ssq * ratio) * ratio
            scale = a
        else:
            ratio = a / scale
            ssq += ratio * ratio
    return scale * math.sqrt(ssq)


def _truth(x) -> bool:
    """TruthValue::truth: x != 0 (complex: either part), NaN is true."""
    if isinstance(x, (complex, np.complexfloating)):
        return x.real != 0 or x.imag != 0
    return x != 0


def any_(values) -> bool:
    return builtins_any(_truth(x) for x in np.asarray(values).reshape(-1))


def all_(values) -> bool:
    return builtins_all(_truth(x) for x in np.asarray(values).reshape(-1))


builtins_any, builtins_all = any, all


# ---------------------------------------------------------------------------
# Broadcast pointwise operations
# ---------------------------------------------------------------------------

class Pencil:
    """What Pencil::same_layout compares at one rank (pencil.rs:238-240)."""

    def __init__(self, shape, perm=None, decomposition=()):
        self.shape = tuple(shape)
        self.perm = tuple(range(len(shape))) if perm is None else tuple(perm)
        self.decomposition = tuple(decomposition)

    def same_layout(self, other: "Pencil") -> bool:
        return (self.shape, self.perm, self.decomposition) == (other.shape, other.perm, other.decomposition)

    def local_len(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))


def _validate_extents(inp: Sequence[int], out: Sequence[int]):
    """validate_extents (pointwise.rs:178-197)."""
    if len(inp) != len(out):
        raise PointwiseError("ExtraRankMismatch", expected=len(out), actual=len(inp))
    for axis, (e, o) in enumerate(zip(inp, out)):
        if e != o and e != 1:
            raise PointwiseError("ExtraExtentMismatch", axis=axis, output=o, input=e)


def validate_pointwise2(left, right, out):
    """validate_two_inputs (pointwise.rs:155-176): both layouts first, then
    the extents of left and right. Arrays are (pencil, extra_shape) pairs."""
    if not left[0].same_layout(right[0]) or not left[0].same_layout(out[0]):
        raise PointwiseError("IncompatiblePencils")
    for arr in (left, right):
        _validate_extents(arr[1], out[1])


def validate_many(inputs, out):
    """many_plan / pointwise3 (pointwise.rs:331-352, :492-512): EmptyInputs,
    then input by input (layout, extents), errors carrying the input index."""
    if len(inputs) == 0:
        raise PointwiseError("EmptyInputs")
    for i, arr in enumerate(inputs):
        try:
            if not arr[0].same_layout(out[0]):
                raise PointwiseError("IncompatiblePencils")
            _validate_extents(arr[1], out[1])
        except PointwiseError as e:
            raise PointwiseError(e.variant, index=i, **e.fields)


def broadcast_linear(out_index: int, out_dims: Sequence[int], dims: Sequence[int]) -> int:
    """broadcast_linear (pointwise.rs:198-214): the input's extra index for
    output extra index `out_index`, extent-1 axes pinned to 0."""
    idx = []
    rem = out_index
    for d in reversed(out_dims):
        idx.append(rem % d)
        rem //= d
    idx.reverse()
    lin = 0
    for i, d in zip(idx, dims):
        lin = lin * d + (0 if d == 1 else i)
    return lin


def pointwise_many(inputs: Sequence[np.ndarray], extra_shapes: Sequence[Sequence[int]], out_extra: Sequence[int],
                   spatial: int, f):
    """pointwise_many_views (pointwise.rs:354-398): for every output extra
    index, then every spatial element, out = f([inputs at their broadcast
    positions]). `inputs` are flat arrays [extra...][spatial]."""
    count = int(np.prod(out_extra, dtype=np.int64)) if len(out_extra) else 1
    cols = []
    for n in range(count):
        bases = [broadcast_linear(n, out_extra, dims) * spatial for dims in extra_shapes]
        cols.append(f([x[b:b + spatial] for x, b in zip(inputs, bases)]))
    return np.concatenate(cols) if cols else np.zeros(0)
