"""Spectral rescaling: normalizeH (src/utils/external.jl:41-76).

The reference estimates the spectral edges with Arpack ``eigs`` (tol 1e-3,
maxiter 300) on the host.  ``method="lanczos"`` (the default) runs an FP64
Lanczos recurrence on the device (``kpm_spectral_edges``) and reads the
extreme Ritz values; ``method="arpack"`` uses the reference's own ARPACK
routines through scipy (``eigsh``), with exact dense eigenvalues for small
operators where ARPACK cannot run (n <= 64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ._lib import ArgumentError, check

_DENSE_CUTOFF = 64


def _hermitian_check(H):
    Hs = sp.csr_matrix(H) if not sp.issparse(H) else H.tocsr()
    D = (Hs - Hs.conj().T).tocsr()
    if D.nnz == 0 or np.max(np.abs(D.data)) == 0:
        return
    dev = np.max(np.abs(D.data))
    scale = np.max(np.abs(Hs.data)) if Hs.nnz else 0.0
    eps = np.finfo(np.float64).eps
    if not dev <= np.sqrt(eps) * scale:
        raise ArgumentError(f"normalizeH: H must be Hermitian (max |H - H'| = {dev}).")


def _edges(H, which):
    n = H.shape[0]
    if n <= _DENSE_CUTOFF:
        ev = np.linalg.eigvalsh(H.toarray() if sp.issparse(H) else np.asarray(H))
        return ev
    Hs = sp.csr_matrix(H, dtype=np.complex128)
    if which == "LM":
        vals = spla.eigsh(Hs, k=min(6, n - 1), which="LM", tol=1e-3, maxiter=300,
                          return_eigenvectors=False)
    else:
        vals = spla.eigsh(Hs, k=1, which=which, tol=1e-3, maxiter=300,
                          return_eigenvectors=False)
    return np.real(vals)


def spectral_edges(H, *, max_iter: int = 1000, tol: float = 1e-4, seed: int = 42):
    """(Emin, Emax) of Hermitian H from the device Lanczos recurrence."""
    from .device import get_context, to_device

    ctx = get_context()
    Hd = to_device(H, ctx)
    lo, hi, it = C.c_double(), C.c_double(), C.c_int64()
    check(ctx.lib.kpm_spectral_edges(ctx.handle, Hd.handle, int(max_iter), float(tol), int(seed),
                                     C.byref(lo), C.byref(hi), C.byref(it)), "spectral_edges")
    return lo.value, hi.value


def normalizeH(H, *, eps: float = 0.1, fixed_a: float = 0.0, center: bool = False,
               method: str = "lanczos"):
    """Returns ``(a, H_norm)`` with H_norm = H/a, a = 2 Emax/(2-eps); or with
    ``center=True`` ``(a, b, H_norm)``, H_norm = (H - bI)/a,
    a = (Emax-Emin)/(2-eps), b = (Emax+Emin)/2; ``fixed_a`` skips the edge
    estimate (src/utils/external.jl:41-76).  ``method``: "lanczos" (device
    FP64 Lanczos) or "arpack" (host ARPACK, the reference's routine)."""
    if method not in ("lanczos", "arpack"):
        raise ArgumentError(f"normalizeH: method must be 'lanczos' or 'arpack' (got {method!r})")
    _hermitian_check(H)
    Hs = sp.csc_matrix(H, dtype=np.complex128) if not sp.issparse(H) else H.astype(np.complex128)
    if fixed_a != 0:
        if center:
            raise ArgumentError(
                "normalizeH: center=true is incompatible with fixed_a — the center shift is "
                "derived from the spectral edges, which fixed_a skips.")
        a = fixed_a
        return a, (Hs / a).tocsc()
    n = Hs.shape[0]
    if method == "lanczos":
        Emin, Emax = spectral_edges(Hs)
        if center:
            a = (Emax - Emin) / (2 - eps)
            b = (Emax + Emin) / 2
            Hn = ((Hs - b * sp.identity(n, dtype=np.complex128, format="csc")) / a).tocsc()
            return a, b, Hn
        a = 2 * max(abs(Emin), abs(Emax)) / (2 - eps)
        return a, (Hs / a).tocsc()
    if center:
        if n <= _DENSE_CUTOFF:
            ev = _edges(Hs, "LA")
            Emax, Emin = float(np.max(ev)), float(np.min(ev))
        else:
            Emax = float(np.max(_edges(Hs, "LA")))
            Emin = float(np.min(_edges(Hs, "SA")))
        a = (Emax - Emin) / (2 - eps)
        b = (Emax + Emin) / 2
        Hn = ((Hs - b * sp.identity(n, dtype=np.complex128, format="csc")) / a).tocsc()
        return a, b, Hn
    es = _edges(Hs, "LM")
    Emax = float(np.max(np.abs(es)))
    a = 2 * Emax / (2 - eps)
    return a, (Hs / a).tocsc()
