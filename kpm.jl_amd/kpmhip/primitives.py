"""Recurrence and reduction primitives on device blocks -- the method table of
ext/KPMCUDAExt.jl:84-150 (chebyshev_iter_single, chebyshev_iter,
chebyshev_iter_wrap, columnwise_dot!, moment accumulator), and the vector
utilities of src/utils/vectors.jl:14-20,125-132.
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, check, dptr
from .device import DeviceAccumulator, DeviceBlock, DeviceCSR, get_context, to_device


def _H(H, ctx) -> DeviceCSR:
    return to_device(H, ctx)


def chebyshev_iter_single(H, V_pp: DeviceBlock, V_p: DeviceBlock, V_out: DeviceBlock | None = None):
    """V_pp <- 2 H V_p - V_pp (in place), or the 3-address form
    V_out <- 2 H V_p - V_pp leaving V_pp intact
    (src/utils/chebyshev_iteration.jl:59-75,157-167; ext/KPMCUDAExt.jl:84-93)."""
    ctx = V_p.ctx
    Hd = _H(H, ctx)
    out = V_pp if V_out is None else V_out
    check(ctx.lib.kpm_cheb_step(ctx.handle, Hd.handle, V_pp.handle, V_p.handle, out.handle),
          "chebyshev_iter_single")


def chebyshev_iter(H, slots: list, n: int):
    """Fill slots 3..n: slots[i] = 2 H slots[i-1] - slots[i-2]
    (src/utils/chebyshev_iteration.jl:16-20, 1-based slot numbering)."""
    Hd = _H(H, slots[0].ctx)
    for i in range(3, n + 1):
        chebyshev_iter_single(Hd, slots[i - 3], slots[i - 2], slots[i - 1])


def chebyshev_iter_wrap(H, slots: list, n: int | None = None):
    """Continue past the end of an n-slot ring: slots 1, 2 receive the next two
    vectors (src/utils/chebyshev_iteration.jl:31-34)."""
    n = len(slots) if n is None else n
    Hd = _H(H, slots[0].ctx)
    chebyshev_iter_single(Hd, slots[n - 2], slots[n - 1], slots[0])
    chebyshev_iter_single(Hd, slots[n - 1], slots[0], slots[1])


def spmm(A, X: DeviceBlock, Y: DeviceBlock, alpha: float = 1.0, beta: float = 0.0):
    """mul!(Y, A, X, alpha, beta) on device blocks."""
    ctx = X.ctx
    Ad = _H(A, ctx)
    check(ctx.lib.kpm_spmm(ctx.handle, Ad.handle, float(alpha), X.handle, float(beta), Y.handle),
          "mul!")


def moment_accumulator(alpha_all, target):
    """moment_accumulator(α_all, target) (src/utils/vectors.jl:101,
    ext/KPMCUDAExt.jl:125-126): the host target itself when the workspace is
    on the host, a zeroed device accumulator of the target's shape when it is
    a device block (or a list of device blocks)."""
    blk = alpha_all[0] if isinstance(alpha_all, (list, tuple)) else alpha_all
    if isinstance(blk, DeviceBlock):
        NR, NC = np.shape(target)
        return DeviceAccumulator(NR, NC, blk.ctx)
    return target


def copy_moment_accumulator(target: np.ndarray, source):
    """copy_moment_accumulator!(target, source) (src/utils/vectors.jl:119-122,
    ext/KPMCUDAExt.jl:147-150)."""
    if isinstance(source, DeviceAccumulator):
        source.download(target)
    elif source is not target:
        np.copyto(target, source)


def columnwise_dot(target, n: int, A: DeviceBlock, B: DeviceBlock, *, alpha=1.0,
                   beta_col: int = 0, beta_scale=0.0):
    """columnwise_dot!(target, n, A, B; alpha, beta_col, beta_scale)
    (src/utils/vectors.jl:103-117): target[i,n] = alpha <A_i|B_i> +
    beta_scale target[i,beta_col], 1-based n and beta_col.  A device
    accumulator takes the dots on the device without a host round trip;
    a host target is written through one synchronising call.  Complex alpha
    or beta_scale raise ArgumentError (the device reduction takes reals)."""
    for name, v in (("alpha", alpha), ("beta_scale", beta_scale)):
        if np.iscomplexobj(v) and np.imag(v) != 0:
            raise ArgumentError(f"{name} must be real (got {v!r})")
    if isinstance(target, DeviceAccumulator):
        if target.NR != A.NR:
            raise ArgumentError("target rows must equal the number of probe columns")
        target.dot(n, A, B, float(np.real(alpha)), int(beta_col), float(np.real(beta_scale)))
        return
    if target.dtype != np.complex128 or not target.flags.f_contiguous:
        raise ArgumentError("target must be a Fortran-ordered complex128 NR x NC matrix")
    NR, NC = target.shape
    if NR != A.NR:
        raise ArgumentError("target rows must equal the number of probe columns")
    ctx = A.ctx
    check(ctx.lib.kpm_colwise_dot(ctx.handle, dptr(target), NC, int(n), A.handle,
                                  B.handle, float(np.real(alpha)), int(beta_col),
                                  float(np.real(beta_scale))), "columnwise_dot!")


def normalize_by_col(psi_in: np.ndarray, NR: int | None = None, centering: bool = False):
    """In-place column normalisation without centering by default
    (src/utils/vectors.jl:14-20)."""
    NR = psi_in.shape[1] if NR is None else NR
    for i in range(NR):
        col = psi_in[:, i]
        if centering:
            col -= col.mean()
        col /= np.linalg.norm(col)
    return psi_in


# ring helpers, src/utils/vectors.jl:125-132 (1-based slot numbers)
def r_i(n: int) -> int:
    return (n - 1) % 3 + 1


def r_ip(n: int) -> int:
    return (n - 2) % 3 + 1


def r_ipp(n: int) -> int:
    return (n - 3) % 3 + 1


def r2_i(n: int) -> int:
    return (n - 1) % 2 + 1


def r2_ip(n: int) -> int:
    return (n - 2) % 2 + 1


def r2_ipp(n: int) -> int:
    return (n - 1) % 2 + 1


__all__ = [
    "chebyshev_iter_single", "chebyshev_iter", "chebyshev_iter_wrap", "spmm", "moment_accumulator",
    "copy_moment_accumulator", "columnwise_dot", "normalize_by_col", "r_i", "r_ip", "r_ipp", "r2_i", "r2_ip", "r2_ipp",
    "get_context",
]
