"""Device layer: the counterpart of src/device.jl and the device half of
ext/KPMCUDAExt.jl.

- ``ACTIVE_DEVICE`` / ``whichcore`` mirror src/device.jl:17,25.  The HIP
  device becomes active when libkpm_hip.so loads and sees a GPU (the
  counterpart of the extension's ``__init__``, ext/KPMCUDAExt.jl:15-20).
- ``to_device(H)`` converts a host sparse matrix into a device CSR operator
  (ext/KPMCUDAExt.jl:26-27); device objects pass through unchanged
  (ext/KPMCUDAExt.jl:29-34).
- ``DeviceBlock`` is a device NH x NR recurrence block
  (``maybe_on_device_zeros(dt_cplx, NH, NR)``; ext/KPMCUDAExt.jl:43).

There is no CPU execution path for the moment engine: without a GPU the
engines raise :class:`KPMHipMissing`.
"""
from __future__ import annotations

import atexit
import ctypes as C
import threading

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import KPMHipMissing, check, dptr, i64ptr


class AbstractDevice:
    pass


class CPUDevice(AbstractDevice):
    def __repr__(self):
        return "CPUDevice()"


class HIPDevice(AbstractDevice):
    def __init__(self, index: int = 0):
        self.index = index

    def __repr__(self):
        return f"HIPDevice({self.index})"


ACTIVE_DEVICE = [CPUDevice()]
_ctx = {}
_ctx_lock = threading.Lock()
_shutdown = [False]


@atexit.register
def _mark_shutdown():
    # handles freed after this point would race the context teardown order of
    # the interpreter; the process exit releases device memory instead
    _shutdown[0] = True


def device_count() -> int:
    lib = _lib.load_library()
    n = C.c_int(0)
    check(lib.kpm_device_count(C.byref(n)), "kpm_device_count")
    return int(n.value)


def activate(index: int = 0) -> bool:
    """Activate the HIP device if one is visible; returns whichcore()."""
    try:
        if device_count() > index:
            ACTIVE_DEVICE[0] = HIPDevice(index)
    except KPMHipMissing:
        pass
    return whichcore()


def whichcore() -> bool:
    """True when the HIP device is active (src/device.jl:25)."""
    return not isinstance(ACTIVE_DEVICE[0], CPUDevice)


class Context:
    """One device + one HIP stream + cached workspaces (a kpm_ctx)."""

    def __init__(self, index: int = 0):
        self.lib = _lib.load_library()
        h = C.c_void_p()
        check(self.lib.kpm_ctx_create(index, C.byref(h)), "kpm_ctx_create")
        self.handle = h
        self.index = index

    def synchronize(self):
        check(self.lib.kpm_ctx_synchronize(self.handle), "kpm_ctx_synchronize")

    def set_dual_pass(self, enable: bool):
        """Fuse the kpm_2d / kpm_3d right-chain step with the Jb product of the
        previous order (default) or run them as two passes; same tables."""
        check(self.lib.kpm_ctx_set_dual_pass(self.handle, 1 if enable else 0), "kpm_ctx_set_dual_pass")

    def set_bf16_planes(self, enable: bool):
        """In the split-BF16 contraction mode, write kpm_2d's slots as BF16
        hi/lo planes from the recurrence passes (default) or as FP64 / FP32
        slots split on every read; bitwise the same tables."""
        check(self.lib.kpm_ctx_set_bf16_planes(self.handle, 1 if enable else 0), "kpm_ctx_set_bf16_planes")

    def set_merged_limit(self, nbytes: int):
        """Cap the merged operator of the dual pass (bytes; negative: no cap).
        Over the cap kpm_2d / kpm_3d run the two passes separately."""
        check(self.lib.kpm_ctx_set_merged_limit(self.handle, int(nbytes)), "merged limit")

    def counters(self, reset: bool = False) -> dict:
        """Launch counters of this context: recurrence passes, dual passes,
        dot-only passes, separate moment reductions, host synchronisations."""
        buf = (C.c_int64 * _lib.COUNT_N)()
        check(self.lib.kpm_ctx_counters(self.handle, buf, _lib.COUNT_N, 1 if reset else 0),
              "counters")
        return {"step": int(buf[_lib.COUNT_STEP]), "dual": int(buf[_lib.COUNT_DUAL]),
                "dot": int(buf[_lib.COUNT_DOT]), "combine": int(buf[_lib.COUNT_COMBINE]),
                "sync": int(buf[_lib.COUNT_SYNC])}

    def free_memory(self) -> int:
        f = C.c_size_t(0)
        t = C.c_size_t(0)
        check(self.lib.kpm_device_free_memory(self.handle, C.byref(f), C.byref(t)), "free memory")
        return int(f.value)

    def total_memory(self) -> int:
        f = C.c_size_t(0)
        t = C.c_size_t(0)
        check(self.lib.kpm_device_free_memory(self.handle, C.byref(f), C.byref(t)), "free memory")
        return int(t.value)

    def name(self) -> str:
        buf = C.create_string_buffer(256)
        check(self.lib.kpm_device_name(self.handle, buf, 256), "device name")
        return buf.value.decode()

    def event_record(self, slot: int):
        check(self.lib.kpm_event_record(self.handle, slot), "event record")

    def event_elapsed(self, a: int, b: int) -> float:
        ms = C.c_float(0.0)
        check(self.lib.kpm_event_elapsed(self.handle, a, b, C.byref(ms)), "event elapsed")
        return float(ms.value)

    def destroy(self):
        if getattr(self, "handle", None) and not _shutdown[0]:
            self.lib.kpm_ctx_destroy(self.handle)
        self.handle = None


def get_context(index: int | None = None) -> Context:
    """The per-device context of the active HIP device (created on first use)."""
    if index is None:
        dev = ACTIVE_DEVICE[0]
        if not isinstance(dev, HIPDevice):
            if not activate(0):
                raise KPMHipMissing(
                    "no HIP device is visible: the KPM moment engine runs only on the GPU "
                    "(libkpm_hip.so); there is no CPU fallback")
            dev = ACTIVE_DEVICE[0]
        index = dev.index
    with _ctx_lock:
        if index not in _ctx:
            _ctx[index] = Context(index)
        return _ctx[index]


def device_free_memory(device=None) -> int:
    """device_free_memory (src/device.jl:66-67; ext/KPMCUDAExt.jl:45)."""
    return get_context().free_memory()


# --------------------------------------------------------------------------
# operators


def _as_csc_complex(A):
    if isinstance(A, np.ndarray):
        A = sp.csc_matrix(A)
    if not sp.issparse(A):
        raise TypeError(f"expected a scipy sparse matrix or ndarray, got {type(A)}")
    A = sp.csc_matrix(A, dtype=np.complex128)
    A.sort_indices()
    n, m = A.shape
    if n != m:
        raise _lib.ArgumentError(f"operator must be square (got {A.shape})")
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
    nz = np.ascontiguousarray(A.data, dtype=np.complex128)
    return n, colptr, rowval, nz


class DeviceCSR:
    """A device CSR operator (kpm_csr)."""

    def __init__(self, A, ctx: Context | None = None, adjoint: bool = False):
        self.ctx = ctx or get_context()
        n, colptr, rowval, nz = _as_csc_complex(A)
        h = C.c_void_p()
        fn = self.ctx.lib.kpm_csr_adjoint_from_csc if adjoint else self.ctx.lib.kpm_csr_from_csc
        check(fn(self.ctx.handle, n, i64ptr(colptr), i64ptr(rowval),
                 nz.ctypes.data_as(C.POINTER(C.c_double)), 0, C.byref(h)),
              "kpm_csr_from_csc")
        self.handle = h
        self.shape = (n, n)
        nn = C.c_int64(0)
        nnz = C.c_int64(0)
        isr = C.c_int(0)
        check(self.ctx.lib.kpm_csr_info(h, C.byref(nn), C.byref(nnz), C.byref(isr)), "csr info")
        self.nnz = int(nnz.value)
        self.is_real = bool(isr.value)
        self.adjoint = adjoint

    def __del__(self):
        try:
            if getattr(self, "handle", None) and not _shutdown[0]:
                self.ctx.lib.kpm_csr_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def to_device(A, ctx: Context | None = None) -> DeviceCSR:
    """maybe_to_device for operators: host sparse/dense -> DeviceCSR;
    a DeviceCSR passes through (ext/KPMCUDAExt.jl:26-34)."""
    if isinstance(A, DeviceCSR):
        return A
    if hasattr(A, "device_operator"):  # BdG and scaled operators assemble themselves
        return A.device_operator(ctx)
    return DeviceCSR(A, ctx)


def to_device_adjoint(A, ctx: Context | None = None) -> DeviceCSR:
    """Device operator of adjoint(A) (the Jα' of src/moment.jl:680)."""
    if isinstance(A, DeviceCSR):
        raise TypeError("pass the host matrix; the adjoint is built from its CSC arrays")
    return DeviceCSR(A, ctx, adjoint=True)


PREC = {"f64": _lib.VEC_F64, "f32": _lib.VEC_F32, "fp64": _lib.VEC_F64, "fp32": _lib.VEC_F32}


def prec_code(precision) -> int:
    if isinstance(precision, int):
        return precision
    try:
        return PREC[str(precision).lower()]
    except KeyError:
        raise _lib.ArgumentError(f"precision must be 'f64' or 'f32' (got {precision!r})") from None


class DeviceBlock:
    """A device NH x NR recurrence block (kpm_block), ComplexF64 or ComplexF32."""

    def __init__(self, NH: int, NR: int, precision="f64", ctx: Context | None = None):
        self.ctx = ctx or get_context()
        self.NH = int(NH)
        self.NR = int(NR)
        self.prec = prec_code(precision)
        h = C.c_void_p()
        check(self.ctx.lib.kpm_block_create(self.ctx.handle, self.NH, self.NR, self.prec,
                                            C.byref(h)), "kpm_block_create")
        self.handle = h

    @classmethod
    def from_host(cls, psi: np.ndarray, precision="f64", ctx: Context | None = None):
        psi = np.asarray(psi)
        if psi.ndim == 1:
            psi = psi.reshape(-1, 1)
        b = cls(psi.shape[0], psi.shape[1], precision, ctx)
        b.upload(psi)
        return b

    def upload(self, psi: np.ndarray):
        a = np.asfortranarray(psi, dtype=np.complex128)
        if a.shape != (self.NH, self.NR):
            raise _lib.ArgumentError(f"block shape {a.shape} != {(self.NH, self.NR)}")
        check(self.ctx.lib.kpm_block_upload(self.handle, dptr(a)), "upload")

    def download(self) -> np.ndarray:
        out = np.zeros((self.NH, self.NR), dtype=np.complex128, order="F")
        check(self.ctx.lib.kpm_block_download(self.handle, dptr(out)), "download")
        return out

    def copy_from(self, other: "DeviceBlock"):
        check(self.ctx.lib.kpm_block_copy(self.handle, other.handle), "block copy")

    def zero(self):
        check(self.ctx.lib.kpm_block_zero(self.handle), "block zero")

    def random_phase(self, seed: int, lanes: int = 1024):
        """Fill with exp(2 pi i U), unit-norm columns, drawn on the device from
        the same lane-parallel xoshiro256++ stream as
        ``random_phase_vectors(Xoshiro256pp(seed, lanes), NH, NR)``."""
        check(self.ctx.lib.kpm_block_random_phase(self.handle, int(seed) & 0xFFFFFFFFFFFFFFFF,
                                                  int(lanes)), "random_phase")
        return self

    def __del__(self):
        try:
            if getattr(self, "handle", None) and not _shutdown[0]:
                self.ctx.lib.kpm_block_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class DeviceAccumulator:
    """The device moment accumulator (kpm_accum): a zeroed NR x NC
    ComplexF64 matrix in device memory, the value of
    ``moment_accumulator(::CuArray, target)`` (ext/KPMCUDAExt.jl:125-126)."""

    def __init__(self, NR: int, NC: int, ctx: Context | None = None):
        self.ctx = ctx or get_context()
        self.NR = int(NR)
        self.NC = int(NC)
        h = C.c_void_p()
        check(self.ctx.lib.kpm_accum_create(self.ctx.handle, self.NR, self.NC, C.byref(h)),
              "moment_accumulator")
        self.handle = h

    @property
    def shape(self):
        return (self.NR, self.NC)

    def dot(self, n: int, A: DeviceBlock, B: DeviceBlock, alpha=1.0, beta_col: int = 0,
            beta_scale=0.0):
        """columnwise_dot!(self, n, A, B; alpha, beta_col, beta_scale), 1-based
        n and beta_col (src/utils/vectors.jl:103-117; ext/KPMCUDAExt.jl:128-145)."""
        check(self.ctx.lib.kpm_accum_dot(self.ctx.handle, self.handle, int(n), A.handle, B.handle,
                                         float(alpha), int(beta_col), float(beta_scale)),
              "columnwise_dot!")

    def download(self, out: np.ndarray | None = None) -> np.ndarray:
        """copy_moment_accumulator!(out, self) (ext/KPMCUDAExt.jl:147-150)."""
        if out is None:
            out = np.zeros((self.NR, self.NC), dtype=np.complex128, order="F")
        if out.shape != (self.NR, self.NC) or out.dtype != np.complex128 or not out.flags.f_contiguous:
            raise _lib.ArgumentError("target must be a Fortran-ordered complex128 NR x NC matrix")
        check(self.ctx.lib.kpm_accum_download(self.ctx.handle, self.handle, dptr(out)),
              "copy_moment_accumulator!")
        return out

    def __del__(self):
        try:
            if getattr(self, "handle", None) and not _shutdown[0]:
                self.ctx.lib.kpm_accum_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def maybe_on_device_zeros(NH: int, NR: int, nslots: int | None = None, precision="f64"):
    """maybe_on_device_zeros(dt_cplx, NH, NR[, nslots]) (src/device.jl:54)."""
    if nslots is None:
        return DeviceBlock(NH, NR, precision)
    return [DeviceBlock(NH, NR, precision) for _ in range(nslots)]


def maybe_to_host(x):
    """maybe_to_host (src/device.jl:46; ext/KPMCUDAExt.jl:36-41)."""
    if isinstance(x, DeviceBlock):
        return x.download()
    return x
