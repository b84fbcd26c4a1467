"""Damping kernels and the h_n factors (src/kernels/kernels.jl:2,14-117,
src/kernels/jackson_kernel.jl:7-8, src/kernels/lorentz_kernel.jl:10-18).
Host FP64; these act on the NC-long (or NC^d) moment tables only."""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError


def hn(n):
    """h_0 = 1, h_n = 2 otherwise (src/kernels/kernels.jl:2)."""
    n = np.asarray(n)
    return (n != 0).astype(np.int64) + 1 if n.ndim else int(n != 0) + 1


def JacksonKernel(n, N: int):
    """Jackson kernel g_n for N moments (src/kernels/jackson_kernel.jl:7-8);
    zero for n >= N."""
    n = np.asarray(n, dtype=np.float64)
    Np1 = N + 1.0
    g = (1.0 / Np1) * ((Np1 - n) * np.cos(np.pi * n / Np1) +
                       np.sin(np.pi * n / Np1) / np.tan(np.pi / Np1))
    return g * (n < N)


def LorentzKernels(lam: float):
    """Return LorentzKernel(n, N) = sinh(λ(1 - n/N))/sinh(λ)
    (src/kernels/lorentz_kernel.jl:10-18)."""
    if not lam > 0:
        raise ArgumentError(
            f"LorentzKernels: λ must be positive (got {lam}); λ ≲ 2 also degrades the "
            "Lorentzian-broadening correspondence")

    def LorentzKernel(n, N: int):
        n = np.asarray(n, dtype=np.float64)
        return np.sinh(lam * (1.0 - n / N)) / np.sinh(lam)

    return LorentzKernel


def _kernel_vec(kernel, NC: int) -> np.ndarray:
    ns = np.arange(NC)
    return np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)


def mu2D_apply_kernel_and_h(mu, NC: int, kernel, dims=(1, 2)):
    """μ̃ = μ ⋅ (g h) along each listed axis (1-based dims), complex copy
    (src/kernels/kernels.jl:14-43)."""
    kernels = kernel if isinstance(kernel, (list, tuple)) else [kernel, kernel]
    mt = np.array(mu, dtype=np.complex128, copy=True)
    if len(dims) == 0:
        return mt
    if 1 in dims:
        mt *= _kernel_vec(kernels[0], NC)[:, None]
    if 2 in dims:
        mt *= _kernel_vec(kernels[1], NC)[None, :]
    return mt


mu2D_apply_kernel_and_h_no_mutate = mu2D_apply_kernel_and_h


def muND_apply_kernel_and_h(mu, NC: int, kernel, dims=(1, 2, 3)):
    """Kernel and h_n along each listed axis of an N-d table
    (src/kernels/kernels.jl:92-103)."""
    mt = np.array(mu, dtype=np.complex128, copy=True)
    kv = _kernel_vec(kernel, NC)
    for d in dims:
        shape = [1] * max(mt.ndim, d)
        shape[d - 1] = NC
        mt = mt * kv.reshape(shape)
    return mt


mu3D_apply_kernel_and_h = muND_apply_kernel_and_h
