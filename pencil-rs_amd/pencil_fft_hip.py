"""Python mirror of pencil-rs's FFT plan API over the gfx950 C ABI.

Each class keeps the name, argument meaning and error behaviour of the Rust
type it mirrors (cited per class). All work runs in ``libpencilfft_hip.so``
(hand-written HIP kernels for gfx950) on device-resident torch tensors; torch
is only used for device memory and the current stream. There is no CPU
fallback: constructing a plan without the library or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Iterable, Optional, Sequence

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
# PFH_LIBRARY names another build of the library (A/B measurements).
LIB_PATH = os.environ.get("PFH_LIBRARY") or os.path.join(_HERE, "libpencilfft_hip.so")
MAX_DIMS = 4

# pfh_kind
(LOCAL_C2C, LOCAL_R2C, LOCAL_R2R, LOCAL_DHT, C2C, R2C, R2R, LOCAL_TRANSPOSE, ALLTOALLV_TRANSPOSE, C2R_PLAN, MIXED_C2C,
 MIXED_R2C) = range(1, 13)
# pfh_op
FORWARD, INVERSE, BACKWARD, ADJOINT_FORWARD, ADJOINT_INVERSE, ADJOINT_BACKWARD = range(6)
EXEC_DEFER_STATUS = 1


class R2rKind:
    """R2rKind (crates/pencil-fft/src/r2r.rs:17-84) with the reference's
    descriptor codes (r2r.rs:53-64); DHT is AxisR2rKind::Dht (code 9)."""

    DctI, DctII, DctIII, DctIV, DstI, DstII, DstIII, DstIV = range(1, 9)
    Dht = 9

    _BACKWARD = {1: 1, 2: 3, 3: 2, 4: 4, 5: 5, 6: 7, 7: 6, 8: 8, 9: 9}

    @staticmethod
    def backward_kind(kind: int) -> int:
        return R2rKind._BACKWARD[kind]


class AxisTransform:
    """AxisTransform (crates/pencil-fft/src/distributed/mixed.rs:86-101) of
    the mixed plans: None (identity), FFT, RFFT (the one real-to-half-complex
    axis of a MixedR2cPlan), or an R2rKind code."""

    NONE = 0
    FFT = 10
    RFFT = 11

    @staticmethod
    def r2r(kind: int) -> int:
        return int(kind)


STATUS_NAMES = {
    0: "Ok", 1: "InvalidLength", 2: "LengthOverflow", 3: "NonIntegralBatch", 4: "BufferLengthMismatch",
    5: "BatchCountMismatch", 6: "ScratchTooSmall", 7: "InvalidSpectrumEndpoint", 8: "InvalidDimensions",
    9: "InvalidInputLayout", 10: "InputLayoutMismatch", 11: "OutputLayoutMismatch", 12: "ExtraShapeMismatch",
    13: "WorkspaceMismatch", 14: "WorkspaceTooSmall", 15: "PreparationFailed", 16: "AllocationFailed",
    17: "InvalidSpectrum", 18: "IncompatibleDistribution", 19: "SourceLayoutMismatch",
    20: "DestinationLayoutMismatch", 21: "InvalidPermutation", 22: "ZeroGlobalExtent", 23: "InvalidArgument",
    24: "Unsupported", 25: "Device", 26: "InvalidGraph", 27: "IncompatiblePencils", 28: "ExtraRankMismatch",
    29: "ExtraExtentMismatch", 30: "EmptyInputs", 31: "InvalidInput", 32: "InvalidPath", 33: "SizeLimit",
    34: "MetadataMismatch", 35: "InvalidFile", 36: "IncompleteFile", 37: "Native", 38: "WriteIncomplete",
    39: "CommitUncertain", 40: "InvalidName", 41: "DuplicateName", 42: "NotFound", 43: "InvalidTail",
    44: "Poisoned", 45: "AliasedBuffers",
}


class PfhError(Exception):
    """A non-Ok pfh_status. ``variant`` is the Rust error variant name;
    payload fields mirror the variant's fields."""

    def __init__(self, status: int, info: "_ErrorInfo"):
        self.status = status
        self.variant = STATUS_NAMES.get(status, f"Status{status}")
        self.required = info.required
        self.actual = info.actual
        self.batch = info.batch
        self.index = info.index
        self.hip_error = info.hip_error
        msg = info.message.decode(errors="replace")
        # the variant's string field (IoError reason / field / what / stage / operation)
        self.detail = msg
        super().__init__(f"{self.variant}: {msg}")


class _Desc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("ndim", ctypes.c_int),
        ("topology_dims", ctypes.c_int),
        ("shape", ctypes.c_size_t * MAX_DIMS),
        ("extra_ndim", ctypes.c_int),
        ("extra_shape", ctypes.c_size_t * MAX_DIMS),
        ("selected", ctypes.c_uint),
        ("axis_kind", ctypes.c_int * MAX_DIMS),
        ("direction", ctypes.c_int * MAX_DIMS),
        ("permute_dims", ctypes.c_int),
        ("transpose_method", ctypes.c_int),
        ("src_perm", ctypes.c_int * MAX_DIMS),
        ("dst_perm", ctypes.c_int * MAX_DIMS),
        ("element_bytes", ctypes.c_size_t),
        ("line_len", ctypes.c_size_t),
        ("r2r_kind", ctypes.c_int),
        ("flags", ctypes.c_int),
    ]


class _ErrorInfo(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("required", ctypes.c_size_t),
        ("actual", ctypes.c_size_t),
        ("batch", ctypes.c_size_t),
        ("index", ctypes.c_size_t),
        ("hip_error", ctypes.c_int),
        ("message", ctypes.c_char * 160),
    ]


class _PencilLayout(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_size_t * MAX_DIMS),
        ("perm", ctypes.c_int * MAX_DIMS),
        ("topology_dims", ctypes.c_int),
        ("decomposition", ctypes.c_int * MAX_DIMS),
    ]


class _Array(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("scalar_type", ctypes.c_int),
        ("extra_ndim", ctypes.c_int),
        ("extra_shape", ctypes.c_size_t * MAX_DIMS),
        ("pencil", _PencilLayout),
    ]


class _DatasetInfo(ctypes.Structure):
    _fields_ = [
        ("scalar_type", ctypes.c_int),
        ("ndim", ctypes.c_int),
        ("extra_ndim", ctypes.c_int),
        ("grid_ndim", ctypes.c_int),
        ("shape", ctypes.c_uint64 * MAX_DIMS),
        ("extra_shape", ctypes.c_uint64 * MAX_DIMS),
        ("writer_grid", ctypes.c_uint64 * MAX_DIMS),
        ("writer_perm", ctypes.c_int * MAX_DIMS),
        ("payload_bytes", ctypes.c_uint64),
    ]


class TransformTiming(ctypes.Structure):
    """TransformTiming / StageTiming (distributed.rs:290-325)."""

    _fields_ = [
        ("stages", ctypes.c_int),
        ("stage_ms", ctypes.c_double * 12),
        ("stage_axis", ctypes.c_int * 12),
        ("stage_bytes", ctypes.c_double * 12),
        ("stage_launches", ctypes.c_int * 12),
        ("total_ms", ctypes.c_double),
    ]

    def entries(self):
        """[(axis, ms, bytes)] in execution order; axis -1 is an endpoint check."""
        return [(self.stage_axis[i], self.stage_ms[i], self.stage_bytes[i]) for i in range(self.stages)]

    def launches(self):
        """Kernel launches of each entry (a stage run slab by slab has several)."""
        return [self.stage_launches[i] for i in range(self.stages)]


_lib = None

EXPORTED_SYMBOLS = (
    "pfh_version", "pfh_last_error", "pfh_status_string", "pfh_plan_create", "pfh_plan_destroy",
    "pfh_plan_workspace_bytes", "pfh_plan_layout", "pfh_execute", "pfh_execute_in_place", "pfh_local_execute",
    "pfh_local_execute_in_place", "pfh_execute_with_timing", "pfh_workspace_take_status", "pfh_workspace_init", "pfh_execute_noalloc", "pfh_batch_create",
    "pfh_batch_destroy", "pfh_batch_workspace_bytes", "pfh_batch_execute", "pfh_batch_tuning",
    "pfh_batch_tuning_source", "pfh_batch_tuning_cache_clear",
    "pfh_pointwise_scale_axes", "pfh_pointwise_scale_axes_complex", "pfh_pointwise_quadratic", "pfh_pointwise_quadratic_adjoint", "pfh_sum_squares",
    "pfh_reduce_workspace_bytes", "pfh_reduce", "pfh_pointwise", "pfh_io_workspace_bytes", "pfh_io_write",
    "pfh_io_read", "pfh_io_read_raw", "pfh_io_catalog", "pfh_io_write_named", "pfh_io_append_named",
    "pfh_io_read_named",
)


def load_library(path: str = LIB_PATH):
    """Loads libpencilfft_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    lib.pfh_version.restype = ctypes.c_char_p
    lib.pfh_status_string.restype = ctypes.c_char_p
    lib.pfh_status_string.argtypes = [i]
    lib.pfh_last_error.argtypes = [ctypes.POINTER(_ErrorInfo)]
    lib.pfh_plan_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(vp)]
    lib.pfh_plan_destroy.argtypes = [vp]
    lib.pfh_plan_workspace_bytes.argtypes = [vp, i, ctypes.POINTER(sz)]
    lib.pfh_plan_layout.argtypes = [vp, i, ctypes.POINTER(ctypes.c_int * MAX_DIMS),
                                    ctypes.POINTER(ctypes.c_size_t * MAX_DIMS), ctypes.POINTER(sz),
                                    ctypes.POINTER(sz)]
    lib.pfh_execute.argtypes = [vp, i, vp, sz, vp, sz, vp, sz, vp, i]
    lib.pfh_execute_in_place.argtypes = [vp, i, vp, sz, vp, sz, vp, i]
    lib.pfh_local_execute.argtypes = [vp, i, vp, sz, vp, sz, vp, sz, vp, i]
    lib.pfh_local_execute_in_place.argtypes = [vp, i, vp, sz, sz, vp, sz, vp, i]
    lib.pfh_execute_with_timing.argtypes = [vp, i, vp, sz, vp, sz, vp, sz, vp, ctypes.POINTER(TransformTiming)]
    lib.pfh_workspace_take_status.argtypes = [vp, vp, vp]
    lib.pfh_workspace_init.argtypes = [vp, vp, sz, vp]
    lib.pfh_execute_noalloc.argtypes = [vp, i, vp, sz, vp, sz, vp, i]
    lib.pfh_batch_create.argtypes = [ctypes.POINTER(_Desc), i, i, ctypes.POINTER(vp)]
    lib.pfh_batch_destroy.argtypes = [vp]
    lib.pfh_batch_workspace_bytes.argtypes = [vp, ctypes.POINTER(sz)]
    lib.pfh_batch_execute.argtypes = [vp, i, ctypes.POINTER(vp), ctypes.POINTER(vp), vp, sz, vp, i]
    ip = ctypes.POINTER(i)
    lib.pfh_batch_tuning.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
    lib.pfh_batch_tuning_source.argtypes = [vp, ip]
    lib.pfh_batch_tuning_cache_clear.argtypes = []
    lib.pfh_batch_tuning_cache_clear.restype = None
    d = ctypes.c_double
    lib.pfh_pointwise_scale_axes.argtypes = [vp, i, vp, sz, vp, vp]
    lib.pfh_pointwise_scale_axes_complex.argtypes = [vp, i, vp, sz, vp, vp]
    lib.pfh_pointwise_quadratic.argtypes = [i, vp, vp, sz, d, vp]
    lib.pfh_pointwise_quadratic_adjoint.argtypes = [i, vp, vp, vp, sz, d, vp]
    lib.pfh_sum_squares.argtypes = [i, vp, sz, vp, sz, vp, vp]
    lib.pfh_reduce_workspace_bytes.argtypes = [ctypes.POINTER(sz)]
    lib.pfh_reduce.argtypes = [i, i, vp, sz, vp, sz, vp, vp]
    lib.pfh_pointwise.argtypes = [i, ctypes.POINTER(_Array), i, ctypes.POINTER(_Array), ctypes.POINTER(d), vp]
    ap = ctypes.POINTER(_Array)
    lib.pfh_io_workspace_bytes.argtypes = [ap, ctypes.POINTER(sz)]
    lib.pfh_io_write.argtypes = [ctypes.c_char_p, ap, vp, sz, vp]
    lib.pfh_io_read.argtypes = [ctypes.c_char_p, ap, vp, sz, vp]
    lib.pfh_io_read_raw.argtypes = [ctypes.c_char_p, ap, ctypes.c_uint64, i, vp, sz, vp]
    lib.pfh_io_catalog.argtypes = [ctypes.c_char_p, ctypes.POINTER(_DatasetInfo)]
    for fn in (lib.pfh_io_write_named, lib.pfh_io_append_named, lib.pfh_io_read_named):
        fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ap, vp, sz, vp]
    _lib = lib
    return lib


def _check(status: int):
    if status != 0:
        info = _ErrorInfo()
        _lib.pfh_last_error(ctypes.byref(info))
        raise PfhError(status, info)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pencil_fft_hip needs a ROCm GPU; there is no CPU fallback")


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("pencil_fft_hip buffers must be device tensors")
    if not t.is_contiguous():
        raise ValueError("pencil_fft_hip buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _precision(dtype: torch.dtype) -> int:
    if dtype in (torch.float32, torch.complex64):
        return 1
    if dtype in (torch.float64, torch.complex128):
        return 2
    raise TypeError(f"unsupported dtype {dtype}")


def _real_dtype(precision: int):
    return torch.float32 if precision == 1 else torch.float64


def _complex_dtype(precision: int):
    return torch.complex64 if precision == 1 else torch.complex128


class _Plan:
    _kind = 0

    def __init__(self, desc: _Desc):
        _require_gpu()
        lib = load_library()
        handle = ctypes.c_void_p()
        _check(lib.pfh_plan_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._h = handle
        self._desc = desc
        self.precision = desc.precision
        self._ws = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.pfh_plan_destroy(h)
            self._h = None

    # -- workspace ---------------------------------------------------------
    def workspace_bytes(self, op: int = FORWARD) -> int:
        n = ctypes.c_size_t()
        _check(_lib.pfh_plan_workspace_bytes(self._h, op, ctypes.byref(n)))
        return n.value

    def allocate_workspace(self) -> torch.Tensor:
        """A workspace large enough for every operation of this plan
        (the reference's *Workspace types, bound to one plan)."""
        need = max(self.workspace_bytes(op) for op in range(6))
        return torch.zeros(max(need, 256), dtype=torch.uint8, device="cuda")

    def init_workspace(self, ws: Optional[torch.Tensor] = None) -> torch.Tensor:
        """Binds a workspace to this plan and resets its status area, which
        also clears an in-place poisoned state (pfh_workspace_init)."""
        w = self._workspace(ws)
        _check(_lib.pfh_workspace_init(self._h, _ptr(w), w.numel() * w.element_size(), _stream()))
        return w

    def _workspace(self, ws: Optional[torch.Tensor]) -> torch.Tensor:
        if ws is not None:
            return ws
        if self._ws is None:
            self._ws = self.allocate_workspace()
        return self._ws

    def layout(self, side: int):
        """(memory order, logical shape, elements, element bytes) of the
        input (side 0) or output (side 1) array: input_pencil() /
        output_pencil() (distributed.rs:1316-1329)."""
        perm = (ctypes.c_int * MAX_DIMS)()
        shape = (ctypes.c_size_t * MAX_DIMS)()
        elems, eb = ctypes.c_size_t(), ctypes.c_size_t()
        _check(_lib.pfh_plan_layout(self._h, side, ctypes.byref(perm), ctypes.byref(shape), ctypes.byref(elems),
                                    ctypes.byref(eb)))
        nd = self._desc.ndim if self._desc.ndim else 1
        return tuple(perm[:nd]), tuple(shape[:nd]), elems.value, eb.value

    def take_status(self, ws: Optional[torch.Tensor] = None):
        """Reads the deferred device status of a workspace (synchronizes)."""
        _check(_lib.pfh_workspace_take_status(self._h, _ptr(self._workspace(ws)), _stream()))


def _fill_common(desc: _Desc, kind: int, dtype, shape=(), extra_shape=()):
    desc.kind = kind
    desc.precision = _precision(dtype)
    desc.ndim = len(shape)
    for a, n in enumerate(shape):
        desc.shape[a] = n
    desc.extra_ndim = len(extra_shape)
    for a, n in enumerate(extra_shape):
        desc.extra_shape[a] = n


# ------------------------------------------------------------------ local --

class _LocalPlan(_Plan):
    def _run(self, op, src, dst, ws=None, flags=0):
        w = self._workspace(ws)
        _check(_lib.pfh_local_execute(self._h, op, _ptr(src), src.numel(), _ptr(dst), dst.numel(), _ptr(w),
                                      w.numel(), _stream(), flags))
        return dst

    def _run_in_place(self, op, data, elems, ws=None, flags=0):
        w = self._workspace(ws)
        _check(_lib.pfh_local_execute_in_place(self._h, op, _ptr(data), elems, data.numel() * data.element_size(),
                                               _ptr(w), w.numel(), _stream(), flags))
        return data


class LocalC2cPlan(_LocalPlan):
    """LocalC2cPlan (crates/pencil-fft/src/lib.rs:367-655): batched
    unnormalized C2C lines; inverse = backward * (1/n)."""

    def __init__(self, line_len: int, dtype=torch.complex64):
        d = _Desc()
        _fill_common(d, LOCAL_C2C, dtype)
        d.line_len = line_len
        self.line_len = line_len
        super().__init__(d)

    def forward(self, src, dst=None, workspace=None):
        return self._run(FORWARD, src, torch.empty_like(src) if dst is None else dst, workspace)

    def backward(self, src, dst=None, workspace=None):
        return self._run(BACKWARD, src, torch.empty_like(src) if dst is None else dst, workspace)

    def inverse(self, src, dst=None, workspace=None):
        return self._run(INVERSE, src, torch.empty_like(src) if dst is None else dst, workspace)

    def adjoint(self, op, src, dst=None, workspace=None):
        return self._run(op, src, torch.empty_like(src) if dst is None else dst, workspace)

    def forward_in_place(self, data, workspace=None):
        return self._run_in_place(FORWARD, data, data.numel(), workspace)

    def backward_in_place(self, data, workspace=None):
        return self._run_in_place(BACKWARD, data, data.numel(), workspace)

    def inverse_in_place(self, data, workspace=None):
        return self._run_in_place(INVERSE, data, data.numel(), workspace)


class LocalR2cPlan(_LocalPlan):
    """LocalR2cPlan (crates/pencil-fft/src/r2c.rs:260-813): real n ->
    n/2+1 complex; inverse/backward reject non-zero DC / Nyquist imaginary
    parts before any write (InvalidSpectrumEndpoint {batch, index})."""

    def __init__(self, real_len: int, dtype=torch.float32):
        d = _Desc()
        _fill_common(d, LOCAL_R2C, dtype)
        d.line_len = real_len
        self.real_len = real_len
        self.complex_len = real_len // 2 + 1
        super().__init__(d)

    def _out_real(self, src):
        batch = src.numel() // self.complex_len
        return torch.empty(batch * self.real_len, dtype=_real_dtype(self.precision), device=src.device)

    def _out_complex(self, src):
        batch = src.numel() // self.real_len
        return torch.empty(batch * self.complex_len, dtype=_complex_dtype(self.precision), device=src.device)

    def forward(self, src, dst=None, workspace=None):
        return self._run(FORWARD, src, self._out_complex(src) if dst is None else dst, workspace)

    def inverse(self, src, dst=None, workspace=None):
        return self._run(INVERSE, src, self._out_real(src) if dst is None else dst, workspace)

    def backward(self, src, dst=None, workspace=None):
        return self._run(BACKWARD, src, self._out_real(src) if dst is None else dst, workspace)

    def adjoint_forward(self, src, dst=None, workspace=None):
        return self._run(ADJOINT_FORWARD, src, self._out_real(src) if dst is None else dst, workspace)

    def adjoint_backward(self, src, dst=None, workspace=None):
        return self._run(ADJOINT_BACKWARD, src, self._out_complex(src) if dst is None else dst, workspace)

    def adjoint_inverse(self, src, dst=None, workspace=None):
        return self._run(ADJOINT_INVERSE, src, self._out_complex(src) if dst is None else dst, workspace)


class LocalR2rPlan(_LocalPlan):
    """LocalR2rPlan (crates/pencil-fft/src/r2r.rs:287-778): FFTW
    REDFT/RODFT kinds, backward = paired kind, inverse = paired kind /
    normalization_factor."""

    def __init__(self, line_len: int, kind: int, dtype=torch.float32):
        d = _Desc()
        _fill_common(d, LOCAL_R2R, dtype)
        d.line_len = line_len
        d.r2r_kind = kind
        self.line_len = line_len
        self.kind = kind
        super().__init__(d)

    def normalization_factor(self) -> int:
        n = self.line_len
        return {1: 2 * (n - 1), 5: 2 * (n + 1)}.get(self.kind, 2 * n)

    def forward(self, src, dst=None, workspace=None):
        return self._run(FORWARD, src, torch.empty_like(src) if dst is None else dst, workspace)

    def backward(self, src, dst=None, workspace=None):
        return self._run(BACKWARD, src, torch.empty_like(src) if dst is None else dst, workspace)

    def inverse(self, src, dst=None, workspace=None):
        return self._run(INVERSE, src, torch.empty_like(src) if dst is None else dst, workspace)

    def adjoint(self, op, src, dst=None, workspace=None):
        return self._run(op, src, torch.empty_like(src) if dst is None else dst, workspace)

    def forward_in_place(self, data, workspace=None):
        return self._run_in_place(FORWARD, data, data.numel(), workspace)

    def inverse_in_place(self, data, workspace=None):
        return self._run_in_place(INVERSE, data, data.numel(), workspace)


class LocalDhtPlan(LocalR2rPlan):
    """LocalDhtPlan (crates/pencil-fft/src/dht.rs:32-303): self-paired
    Hartley transform; inverse divides by n."""

    def __init__(self, line_len: int, dtype=torch.float32):
        d = _Desc()
        _fill_common(d, LOCAL_DHT, dtype)
        d.line_len = line_len
        self.line_len = line_len
        self.kind = R2rKind.Dht
        _Plan.__init__(self, d)

    def normalization_factor(self) -> int:
        return self.line_len


# ------------------------------------------------------------ distributed --

class _RoutePlan(_Plan):
    """Common execution of the one-rank distributed plans."""

    def _in_out(self, op):
        fwd_like = op in (FORWARD, ADJOINT_INVERSE, ADJOINT_BACKWARD)
        _, _, in_el, in_b = self.layout(0)
        _, _, out_el, out_b = self.layout(1)
        return ((in_el, in_b), (out_el, out_b)) if fwd_like else ((out_el, out_b), (in_el, in_b))

    def _alloc(self, elems, eb, device):
        real = _real_dtype(self.precision)
        cplx = _complex_dtype(self.precision)
        dt = cplx if eb == 2 * torch.tensor([], dtype=real).element_size() else real
        return torch.empty(elems, dtype=dt, device=device)

    def execute(self, op, src, dst=None, workspace=None, flags=0):
        (_, _), (oe, ob) = self._in_out(op)
        if dst is None:
            dst = self._alloc(oe, ob, src.device)
        w = self._workspace(workspace)
        _check(_lib.pfh_execute(self._h, op, _ptr(src), src.numel(), _ptr(dst), dst.numel(), _ptr(w), w.numel(),
                                _stream(), flags))
        return dst

    def execute_with_timing(self, op, src, dst=None, workspace=None) -> tuple:
        (_, _), (oe, ob) = self._in_out(op)
        if dst is None:
            dst = self._alloc(oe, ob, src.device)
        w = self._workspace(workspace)
        t = TransformTiming()
        _check(_lib.pfh_execute_with_timing(self._h, op, _ptr(src), src.numel(), _ptr(dst), dst.numel(), _ptr(w),
                                            w.numel(), _stream(), ctypes.byref(t)))
        return dst, t

    def execute_in_place(self, op, data, workspace=None, flags=0):
        w = self._workspace(workspace)
        _check(_lib.pfh_execute_in_place(self._h, op, _ptr(data), data.numel() * data.element_size(), _ptr(w),
                                         w.numel(), _stream(), flags))
        return data

    def forward(self, src, dst=None, workspace=None, flags=0):
        return self.execute(FORWARD, src, dst, workspace, flags)

    def inverse(self, src, dst=None, workspace=None, flags=0):
        return self.execute(INVERSE, src, dst, workspace, flags)

    def backward(self, src, dst=None, workspace=None, flags=0):
        return self.execute(BACKWARD, src, dst, workspace, flags)

    def adjoint_forward(self, src, dst=None, workspace=None):
        return self.execute(ADJOINT_FORWARD, src, dst, workspace)

    def adjoint_inverse(self, src, dst=None, workspace=None):
        return self.execute(ADJOINT_INVERSE, src, dst, workspace)

    def adjoint_backward(self, src, dst=None, workspace=None):
        return self.execute(ADJOINT_BACKWARD, src, dst, workspace)

    # In-place, timed and overlapped forms under the reference's method names
    # (C2cPlan::{forward,inverse,backward}_{in_place,with_timing,with_overlap},
    # distributed.rs:1507-1700). At one rank there is no exchange to overlap:
    # the overlapped forms are the plain ones.
    def forward_in_place(self, data, workspace=None, flags=0):
        return self.execute_in_place(FORWARD, data, workspace, flags)

    def inverse_in_place(self, data, workspace=None, flags=0):
        return self.execute_in_place(INVERSE, data, workspace, flags)

    def backward_in_place(self, data, workspace=None, flags=0):
        return self.execute_in_place(BACKWARD, data, workspace, flags)

    def forward_with_timing(self, src, dst=None, workspace=None):
        return self.execute_with_timing(FORWARD, src, dst, workspace)

    def inverse_with_timing(self, src, dst=None, workspace=None):
        return self.execute_with_timing(INVERSE, src, dst, workspace)

    def backward_with_timing(self, src, dst=None, workspace=None):
        return self.execute_with_timing(BACKWARD, src, dst, workspace)

    forward_with_overlap = forward
    inverse_with_overlap = inverse
    backward_with_overlap = backward

    def input_elements(self) -> int:
        return self.layout(0)[2]

    def output_elements(self) -> int:
        return self.layout(1)[2]


def _route_desc(kind, shape, dtype, extra_shape=(), selection=None, permute_dims=True, topology_dims=0):
    d = _Desc()
    _fill_common(d, kind, dtype, shape, extra_shape)
    if selection is not None:
        mask = 0
        for a in selection:
            mask |= 1 << a
        d.selected = mask if mask else 0
        if not mask:
            d.selected = 0x80000000  # explicit empty selection
    d.permute_dims = int(bool(permute_dims))
    d.topology_dims = topology_dims
    return d


class C2cPlan(_RoutePlan):
    """C2cPlan (crates/pencil-fft/src/distributed.rs:1016-1800) on a
    one-rank process grid. Output memory order is the reversed axes when
    permute_dims is set (build_route, distributed.rs:2306-2373)."""

    def __init__(self, shape: Sequence[int], dtype=torch.complex64, extra_shape=(), selection=None,
                 directions=None, permute_dims=True, topology_dims=0):
        d = _route_desc(C2C, shape, dtype, extra_shape, selection, permute_dims, topology_dims)
        if directions is not None:
            for a, v in enumerate(directions):
                d.direction[a] = int(v)
        super().__init__(d)
        self.shape = tuple(shape)


class R2cPlan(_RoutePlan):
    """R2cPlan (crates/pencil-fft/src/distributed/r2c.rs:255-1130): the
    reduction axis is the last selected axis; inverse/backward run the
    endpoint acceptance check (r2c.rs:4709-4878) and raise
    InvalidSpectrum without writing dst."""

    def __init__(self, shape: Sequence[int], dtype=torch.float32, extra_shape=(), selection=None,
                 permute_dims=True, topology_dims=0):
        d = _route_desc(R2C, shape, dtype, extra_shape, selection, permute_dims, topology_dims)
        super().__init__(d)
        self.shape = tuple(shape)


class C2rPlan(_RoutePlan):
    """C2rPlan (crates/pencil-fft/src/distributed/c2r.rs:72-361) on a
    one-rank process grid: the real route in C2R orientation. `shape` is the
    real shape; the reduction axis is the first selected axis. forward = raw
    C2R from the canonical complex pencil (endpoint planes checked, raises
    PfhError("InvalidSpectrum")), backward = raw R2C, inverse = R2C
    normalized by every selected real extent. Out of place only."""

    def __init__(self, shape: Sequence[int], dtype=torch.float32, extra_shape=(), selection=None,
                 permute_dims=True, topology_dims=0):
        d = _route_desc(C2R_PLAN, shape, dtype, extra_shape, selection, permute_dims, topology_dims)
        super().__init__(d)
        self.shape = tuple(shape)

    def execute_in_place(self, op, data, workspace=None, flags=0):
        raise ValueError("C2rPlan is out of place only (distributed/c2r.rs:69-70)")


class MixedC2cPlan(_RoutePlan):
    """MixedC2cPlan (crates/pencil-fft/src/distributed/mixed.rs:392-439,
    1700-1800) on a one-rank process grid: one AxisTransform per axis on
    complex data; an R2R axis transforms the real and imaginary parts
    separately. Output memory order: the reversed axes (permute_dims)."""

    _kind = MIXED_C2C

    def __init__(self, shape: Sequence[int], transforms: Sequence[int], dtype=torch.complex64, extra_shape=(),
                 directions=None, permute_dims=True, topology_dims=0):
        d = _route_desc(self._kind, shape, dtype, extra_shape, None, permute_dims, topology_dims)
        for a, t in enumerate(transforms):
            d.axis_kind[a] = int(t or 0)
        if directions is not None:
            for a, v in enumerate(directions):
                d.direction[a] = int(v)
        super().__init__(d)
        self.shape = tuple(shape)
        self.transforms = tuple(transforms)


class MixedR2cPlan(MixedC2cPlan):
    """MixedR2cPlan (crates/pencil-fft/src/distributed/mixed.rs:441-446,
    2740+): exactly one AxisTransform.RFFT axis (the R2C boundary), real R2R
    or identity stages on the axes after it, complex stages before it; the
    inverse checks the boundary planes (InvalidSpectrum)."""

    _kind = MIXED_R2C

    def __init__(self, shape: Sequence[int], transforms: Sequence[int], dtype=torch.float32, extra_shape=(),
                 directions=None, permute_dims=True, topology_dims=0):
        super().__init__(shape, transforms, dtype, extra_shape, directions, permute_dims, topology_dims)


class R2rPlan(_RoutePlan):
    """R2rPlan (crates/pencil-fft/src/distributed/r2r.rs:71-1460): one
    Option<AxisR2rKind> per axis (None = identity, R2rKind.Dht = Hartley)."""

    def __init__(self, shape: Sequence[int], kinds: Sequence[Optional[int]], dtype=torch.float32, extra_shape=(),
                 permute_dims=True, topology_dims=0):
        d = _route_desc(R2R, shape, dtype, extra_shape, None, permute_dims, topology_dims)
        for a, k in enumerate(kinds):
            d.axis_kind[a] = int(k or 0)
        super().__init__(d)
        self.shape = tuple(shape)
        self.kinds = tuple(kinds)


class DhtPlan(R2rPlan):
    """DhtPlan (crates/pencil-fft/src/distributed/r2r.rs:83-360): Hartley
    transform on the selected axes."""

    def __init__(self, shape: Sequence[int], dtype=torch.float32, extra_shape=(), selection=None,
                 permute_dims=True, topology_dims=0):
        sel = range(len(shape)) if selection is None else selection
        kinds = [R2rKind.Dht if a in sel else None for a in range(len(shape))]
        super().__init__(shape, kinds, dtype, extra_shape, permute_dims, topology_dims)


class LocalTransposePlan(_Plan):
    """LocalTransposePlan (crates/pencil-array/src/local_transpose.rs:46-277):
    bit-exact memory-axis permutation between two layouts of one shape."""

    _kind = LOCAL_TRANSPOSE

    def __init__(self, shape: Sequence[int], src_perm: Sequence[int], dst_perm: Sequence[int], element_bytes: int,
                 extra_shape=()):
        d = _Desc()
        _fill_common(d, self._kind, torch.float32, shape, extra_shape)
        for a in range(len(shape)):
            d.src_perm[a] = src_perm[a]
            d.dst_perm[a] = dst_perm[a]
        d.element_bytes = element_bytes
        super().__init__(d)

    def execute(self, src, dst, backward=False):
        w = self._workspace(None)
        _check(_lib.pfh_execute(self._h, BACKWARD if backward else FORWARD, _ptr(src),
                                src.numel() * src.element_size() // self._desc.element_bytes, _ptr(dst),
                                dst.numel() * dst.element_size() // self._desc.element_bytes, _ptr(w), w.numel(),
                                _stream(), 0))
        return dst

    def execute_in_place(self, data, scratch=None):
        w = scratch if scratch is not None else self._workspace(None)
        _check(_lib.pfh_execute_in_place(self._h, FORWARD, _ptr(data), data.numel() * data.element_size(), _ptr(w),
                                         w.numel(), _stream(), 0))
        return data


class AllToAllvTransposePlan(LocalTransposePlan):
    """AllToAllvTransposePlan (crates/pencil-array/src/alltoallv_transpose.rs
    :50-311) on one rank: pack, self exchange and unpack fused into one
    device permutation."""

    _kind = ALLTOALLV_TRANSPOSE


TUNING_SOURCES = {0: "default", 1: "measured", 2: "process", 3: "file"}


def clear_tuning_cache():
    """Forgets the batch tunings measured or read in this process (the cache
    file, PFH_TUNE_CACHE or ~/.cache/pencil-fft-hip/batch_tuning.txt, is
    kept): the next autotuned BatchC2cPlan reads the file or measures."""
    load_library().pfh_batch_tuning_cache_clear()


class BatchC2cPlan:
    """A batch of independent C2C problems of uneven shapes (one C2cPlan
    each, every axis selected, one precision), executed as one persistent
    launch per route stage, optionally replayed from a captured hipGraph.
    The reference runs such collections one plan after another
    (crates/pencil-fft/src/distributed/collections.rs:229-311); the results
    are those of the per-problem C2cPlan.execute calls."""

    def __init__(self, shapes, dtype=torch.complex64, autotune: bool = True, extra_shapes=None):
        _require_gpu()
        lib = load_library()
        n = len(shapes)
        if n == 0:
            raise ValueError("empty batch")
        descs = (_Desc * n)()
        for k, shp in enumerate(shapes):
            extra = extra_shapes[k] if extra_shapes is not None else ()
            descs[k] = _route_desc(C2C, shp, dtype, extra)
        handle = ctypes.c_void_p()
        _check(lib.pfh_batch_create(descs, n, int(bool(autotune)), ctypes.byref(handle)))
        self._h = handle
        self.shapes = [tuple(s) for s in shapes]
        self.precision = _precision(dtype)
        self._ws = None
        nb = ctypes.c_size_t()
        _check(lib.pfh_batch_workspace_bytes(self._h, ctypes.byref(nb)))
        self.workspace_bytes = nb.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.pfh_batch_destroy(h)
            self._h = None

    def tuning(self):
        """(threads per block target, largest radix, lines per column tile,
        one launch per stage over all compile-time plans, persistent
        workgroups rather than one workgroup per tile, fused plane stages,
        streams the launches run on)."""
        b, r, c, m, p, q, k = (ctypes.c_int() for _ in range(7))
        _check(_lib.pfh_batch_tuning(self._h, ctypes.byref(b), ctypes.byref(r), ctypes.byref(c), ctypes.byref(m),
                                     ctypes.byref(p), ctypes.byref(q), ctypes.byref(k)))
        return b.value, r.value, c.value, bool(m.value), bool(p.value), bool(q.value), k.value

    @property
    def tuning_source(self) -> str:
        """Where the tuning came from: "default" (autotune off), "measured"
        (timed by this creation), "process" (measured earlier in this
        process) or "file" (read from the tuning cache file)."""
        v = ctypes.c_int()
        _check(_lib.pfh_batch_tuning_source(self._h, ctypes.byref(v)))
        return TUNING_SOURCES[v.value]

    def allocate_workspace(self) -> torch.Tensor:
        return torch.zeros(max(self.workspace_bytes, 256), dtype=torch.uint8, device="cuda")

    def execute(self, op, srcs, dsts=None, workspace=None, use_graph=True):
        if len(srcs) != len(self.shapes):
            raise ValueError("one source per problem")
        if dsts is None:
            dsts = [torch.empty_like(s) for s in srcs]
        if workspace is None:
            if self._ws is None:
                self._ws = self.allocate_workspace()
            workspace = self._ws
        n = len(srcs)
        sp = (ctypes.c_void_p * n)(*[_ptr(t).value for t in srcs])
        dp = (ctypes.c_void_p * n)(*[_ptr(t).value for t in dsts])
        _check(_lib.pfh_batch_execute(self._h, op, sp, dp, _ptr(workspace), workspace.numel(), _stream(),
                                      int(bool(use_graph))))
        return dsts

    def forward(self, srcs, dsts=None, workspace=None, use_graph=True):
        return self.execute(FORWARD, srcs, dsts, workspace, use_graph)

    def inverse(self, srcs, dsts=None, workspace=None, use_graph=True):
        return self.execute(INVERSE, srcs, dsts, workspace, use_graph)

    def backward(self, srcs, dsts=None, workspace=None, use_graph=True):
        return self.execute(BACKWARD, srcs, dsts, workspace, use_graph)


# ------------------------------------------------------------- pointwise --
def scale_axes(plan, side: int, data: torch.Tensor, factors: torch.Tensor) -> torch.Tensor:
    """data *= prod_a factors_a[i_a] in place over the plan's input (side 0)
    or output (side 1) array; `factors` is the per-axis arrays concatenated
    in logical axis order (LocalGrid coordinates,
    crates/pencil-array/src/grid.rs:20-120), in the plan's real dtype, or in
    its complex dtype for complex factors (i k_a: a spectral derivative)."""
    fn = _lib.pfh_pointwise_scale_axes_complex if factors.is_complex() else _lib.pfh_pointwise_scale_axes
    _check(fn(plan._h, side, _ptr(data), data.numel(), _ptr(factors), _stream()))
    return data


def quadratic(w: torch.Tensor, a: float, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """out = w + a w^2 (real tensors)."""
    out = torch.empty_like(w) if out is None else out
    _check(_lib.pfh_pointwise_quadratic(_precision(w.dtype), _ptr(w), _ptr(out), w.numel(), float(a), _stream()))
    return out


def quadratic_adjoint(w: torch.Tensor, g: torch.Tensor, a: float, out: Optional[torch.Tensor] = None):
    """out = g (1 + 2 a w)."""
    out = torch.empty_like(g) if out is None else out
    _check(_lib.pfh_pointwise_quadratic_adjoint(_precision(w.dtype), _ptr(w), _ptr(g), _ptr(out), w.numel(),
                                                float(a), _stream()))
    return out


def sum_squares(x: torch.Tensor, workspace: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Sum of squares of every real component (complex: both parts), f64,
    fixed summation order; returns a one-element device tensor."""
    ws = torch.empty(8192, dtype=torch.uint8, device=x.device) if workspace is None else workspace
    res = torch.empty(1, dtype=torch.float64, device=x.device)
    n = x.numel() * (2 if x.is_complex() else 1)
    _check(_lib.pfh_sum_squares(_precision(x.dtype), _ptr(x), n, _ptr(ws), ws.numel(), _ptr(res), _stream()))
    return res


# ------------------------------------------------------------ reductions --
# Scalar type codes: IoElement::CODE (crates/pencil-io/src/format.rs:50-60, :95-126).
SCALAR_TYPES = {
    torch.int8: 1, torch.uint8: 2, torch.int16: 3, torch.uint16: 4, torch.int32: 5, torch.uint32: 6,
    torch.int64: 7, torch.uint64: 8, torch.float32: 9, torch.float64: 10, torch.complex64: 11, torch.complex128: 12,
}
REDUCE_SUM, REDUCE_MIN, REDUCE_MAX, REDUCE_L2_NORM, REDUCE_ANY, REDUCE_ALL = range(1, 7)
_REDUCE_VALUE, _REDUCE_NONE, _REDUCE_INTEGER_OVERFLOW = range(3)
_NORM_DTYPE = {torch.float32: torch.float32, torch.complex64: torch.float32}


class CollectiveError(Exception):
    """CollectiveError (crates/pencil-array/src/error.rs); ``variant`` names
    the Rust variant (``IntegerOverflow``)."""

    def __init__(self, variant: str):
        self.variant = variant
        super().__init__(variant)


def reduce_workspace_bytes() -> int:
    load_library()
    n = ctypes.c_size_t()
    _check(_lib.pfh_reduce_workspace_bytes(ctypes.byref(n)))
    return n.value


def reduce_async(op: int, x: torch.Tensor, result: Optional[torch.Tensor] = None,
                 workspace: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Runs one reduction in stream order; returns the 32-byte device result
    (pfh_reduce_result: value bytes, then the outcome word)."""
    load_library()
    if x.dtype not in SCALAR_TYPES:
        raise TypeError(f"unsupported dtype {x.dtype}")
    ws = torch.empty(reduce_workspace_bytes(), dtype=torch.uint8, device=x.device) if workspace is None else workspace
    res = torch.empty(32, dtype=torch.uint8, device=x.device) if result is None else result
    xp = _ptr(x) if x.numel() else None
    _check(_lib.pfh_reduce(op, SCALAR_TYPES[x.dtype], xp, x.numel(), _ptr(ws), ws.numel(), _ptr(res), _stream()))
    return res


def _reduce_value(op: int, x: torch.Tensor):
    raw = reduce_async(op, x).cpu()
    outcome = int(raw[16:20].view(torch.int32).item())
    if outcome == _REDUCE_INTEGER_OVERFLOW:
        raise CollectiveError("IntegerOverflow")
    if outcome == _REDUCE_NONE:
        return None
    if op in (REDUCE_ANY, REDUCE_ALL):
        return bool(raw[0].item())
    dt = _NORM_DTYPE.get(x.dtype, torch.float64) if op == REDUCE_L2_NORM else x.dtype
    width = torch.empty(0, dtype=dt).element_size()
    return raw[:width].view(dt).item()


def global_sum(x: torch.Tensor):
    """global_sum (collectives.rs:519-528) of a device array at one rank:
    integer sums are checked at every storage-order step (raises
    CollectiveError("IntegerOverflow")); NaN wins, opposite infinities give
    NaN, a lone infinity sign is kept (complex parts independently)."""
    return _reduce_value(REDUCE_SUM, x)


def global_min(x: torch.Tensor):
    """global_min (collectives.rs:530-539): None for an empty array, NaN if
    any value is NaN."""
    return _reduce_value(REDUCE_MIN, x)


def global_max(x: torch.Tensor):
    """global_max (collectives.rs:541-550)."""
    return _reduce_value(REDUCE_MAX, x)


def l2_norm(x: torch.Tensor):
    """l2_norm (collectives.rs:552-565): scaled, NaN -> NaN, infinity -> +inf,
    empty -> 0; f32 for f32 / complex64 arrays, f64 otherwise."""
    return _reduce_value(REDUCE_L2_NORM, x)


def any(x: torch.Tensor) -> bool:  # noqa: A001 - the reference's name
    """any (collectives.rs:567-575): some value is non-zero."""
    return _reduce_value(REDUCE_ANY, x)


def all(x: torch.Tensor) -> bool:  # noqa: A001 - the reference's name
    """all (collectives.rs:577-588): every value is non-zero (true when empty)."""
    return _reduce_value(REDUCE_ALL, x)


# ----------------------------------------------------- broadcast pointwise --
POINTWISE_LINEAR, POINTWISE_MUL, POINTWISE_MUL_ADD = 1, 2, 3


class PencilLayout:
    """The spatial layout of a pencil at one rank (Pencil, pencil.rs:69-112):
    global shape, memory permutation, ordered decomposition."""

    def __init__(self, shape: Sequence[int], perm: Optional[Sequence[int]] = None,
                 decomposition: Sequence[int] = ()):
        self.shape = tuple(int(v) for v in shape)
        self.perm = tuple(range(len(self.shape))) if perm is None else tuple(int(v) for v in perm)
        self.decomposition = tuple(int(v) for v in decomposition)

    def local_len(self) -> int:
        n = 1
        for v in self.shape:
            n *= v
        return n

    def _c(self) -> _PencilLayout:
        c = _PencilLayout()
        c.ndim = len(self.shape)
        for a in range(len(self.shape)):
            c.shape[a] = self.shape[a]
            c.perm[a] = self.perm[a] if a < len(self.perm) else -1
        c.topology_dims = len(self.decomposition)
        for a, v in enumerate(self.decomposition):
            c.decomposition[a] = v
        return c


class PencilArray:
    """A device pencil array (PencilArray, array.rs:17-56): ``data`` holds
    [extra...][memory-ordered spatial] elements of a float or complex dtype."""

    def __init__(self, data: torch.Tensor, layout: PencilLayout, extra_shape: Sequence[int] = ()):
        self.data, self.layout, self.extra_shape = data, layout, tuple(int(v) for v in extra_shape)
        count = layout.local_len()
        for v in self.extra_shape:
            count *= v
        if data.numel() != count:
            raise ValueError(f"array holds {data.numel()} elements, its shapes need {count}")

    def _c(self) -> _Array:
        c = _Array()
        c.data = _ptr(self.data).value if self.data.numel() else None
        c.scalar_type = SCALAR_TYPES.get(self.data.dtype, 0)
        c.extra_ndim = len(self.extra_shape)
        for a, v in enumerate(self.extra_shape[:MAX_DIMS]):
            c.extra_shape[a] = v
        c.pencil = self.layout._c()
        return c


def _pointwise(op: int, inputs: Sequence[PencilArray], out: PencilArray, coeffs: Sequence[complex]):
    load_library()
    arr = (_Array * max(1, len(inputs)))(*[a._c() for a in inputs])
    co = (ctypes.c_double * max(2, 2 * len(coeffs)))()
    for k, c in enumerate(coeffs):
        c = complex(c)
        co[2 * k], co[2 * k + 1] = c.real, c.imag
    oc = out._c()
    _check(_lib.pfh_pointwise(op, arr, len(inputs), ctypes.byref(oc), co, _stream()))
    return out


def pointwise_linear(inputs: Sequence[PencilArray], coeffs: Sequence[complex], out: PencilArray) -> PencilArray:
    """out = sum_i coeffs[i] * inputs[i] with broadcast extras
    (pointwise_many, pointwise.rs:354-420; out may be inputs[0], the in-place
    form :421-490)."""
    return _pointwise(POINTWISE_LINEAR, inputs, out, coeffs)


def pointwise_mul(a: PencilArray, b: PencilArray, out: PencilArray, coeff: complex = 1.0) -> PencilArray:
    """out = coeff * a * b (pointwise2, pointwise.rs:40-133)."""
    return _pointwise(POINTWISE_MUL, [a, b], out, [coeff])


def pointwise_mul_add(a: PencilArray, b: PencilArray, c: PencilArray, out: PencilArray, coeff_ab: complex = 1.0,
                      coeff_c: complex = 1.0) -> PencilArray:
    """out = coeff_ab * a * b + coeff_c * c (pointwise3, pointwise.rs:492-598)."""
    return _pointwise(POINTWISE_MUL_ADD, [a, b, c], out, [coeff_ab, coeff_c])


# ------------------------------------------------------------ file formats --
BYTE_ORDER = {"native": 0, "little": 1, "big": 2}


def _path(path) -> bytes:
    return os.fsencode(path)


def _io_workspace(array: PencilArray, workspace: Optional[torch.Tensor]) -> torch.Tensor:
    c = array._c()
    n = ctypes.c_size_t()
    _check(_lib.pfh_io_workspace_bytes(ctypes.byref(c), ctypes.byref(n)))
    if workspace is not None:
        return workspace
    return torch.empty(max(1, n.value), dtype=torch.uint8, device=array.data.device)


def write_mpi(path, array: PencilArray, workspace: Optional[torch.Tensor] = None):
    """write_mpi (crates/pencil-io/src/mpi_io.rs:198-208) of a device array at
    one rank: exclusive creation, versioned header, canonical little-endian
    payload, commit marker after the flushed payload."""
    load_library()
    ws = _io_workspace(array, workspace)
    c = array._c()
    _check(_lib.pfh_io_write(_path(path), ctypes.byref(c), _ptr(ws), ws.numel(), _stream()))


def read_mpi(path, array: PencilArray, workspace: Optional[torch.Tensor] = None) -> PencilArray:
    """read_mpi (mpi_io.rs:454-472): validates the header against the
    destination's type, shapes and payload size, stages the payload, then
    writes the destination in its own memory order; a rejected read leaves
    it unchanged."""
    load_library()
    ws = _io_workspace(array, workspace)
    c = array._c()
    _check(_lib.pfh_io_read(_path(path), ctypes.byref(c), _ptr(ws), ws.numel(), _stream()))
    return array


def read_mpi_raw(path, array: PencilArray, offset: int = 0, byte_order: str = "native",
                 workspace: Optional[torch.Tensor] = None) -> PencilArray:
    """read_mpi_raw (crates/pencil-io/src/raw.rs:24-238): headerless canonical
    values at a byte offset, each real component in the given byte order."""
    load_library()
    ws = _io_workspace(array, workspace)
    c = array._c()
    _check(_lib.pfh_io_read_raw(_path(path), ctypes.byref(c), int(offset), BYTE_ORDER[byte_order], _ptr(ws),
                                ws.numel(), _stream()))
    return array


class DatasetInfo:
    """DatasetInfo (crates/pencil-io/src/catalog.rs:113-160) of one file."""

    def __init__(self, c: _DatasetInfo):
        inv = {v: k for k, v in SCALAR_TYPES.items()}
        self.scalar_type = inv.get(c.scalar_type, c.scalar_type)
        self.global_shape = tuple(int(c.shape[i]) for i in range(c.ndim))
        self.extra_shape = tuple(int(c.extra_shape[i]) for i in range(c.extra_ndim))
        self.writer_grid = tuple(int(c.writer_grid[i]) for i in range(c.grid_ndim))
        self.writer_permutation = tuple(int(c.writer_perm[i]) for i in range(c.ndim))
        self.payload_bytes = int(c.payload_bytes)


def read_mpi_catalog(path) -> DatasetInfo:
    """read_mpi_catalog (catalog.rs:272-300): the validated metadata of a
    committed file without reading its payload."""
    load_library()
    info = _DatasetInfo()
    _check(_lib.pfh_io_catalog(_path(path), ctypes.byref(info)))
    return DatasetInfo(info)


def _named(fn, path, name: str, array: PencilArray, workspace: Optional[torch.Tensor]):
    load_library()
    ws = _io_workspace(array, workspace)
    c = array._c()
    raw = name.encode("utf-8") if isinstance(name, str) else bytes(name)
    if b"\0" in raw:
        # a C string cannot carry NUL: the name is invalid (valid_name_bytes, lib.rs:343-348)
        raw = b""
    _check(fn(_path(path), raw, ctypes.byref(c), _ptr(ws), ws.numel(), _stream()))
    return array


def write_mpi_named(path, name: str, array: PencilArray, workspace: Optional[torch.Tensor] = None):
    """write_mpi_named (crates/pencil-io/src/named_mpi.rs:996-1019): a new
    named (v2) container holding one committed record."""
    _named(_lib_fn("pfh_io_write_named"), path, name, array, workspace)


def append_mpi_named(path, name: str, array: PencilArray, workspace: Optional[torch.Tensor] = None):
    """append_mpi_named: one more committed record (InvalidTail if the
    committed prefix does not reach the end of the file, DuplicateName if the
    name is taken)."""
    _named(_lib_fn("pfh_io_append_named"), path, name, array, workspace)


def read_mpi_named(path, name: str, array: PencilArray, workspace: Optional[torch.Tensor] = None) -> PencilArray:
    """read_mpi_named: the committed record `name` into the destination."""
    return _named(_lib_fn("pfh_io_read_named"), path, name, array, workspace)


def _lib_fn(name: str):
    return getattr(load_library(), name)
