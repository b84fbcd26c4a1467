"""Spectral rescaling: normalizeH (src/utils/external.jl:41-76).

The reference estimates the spectral edges with Arpack ``eigs`` (tol 1e-3,
maxiter 300) on the host; this mirror uses the same ARPACK routines through
scipy (``eigsh``, Lanczos for Hermitian H), and exact dense eigenvalues for
small operators where ARPACK cannot run (n <= 64).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ._lib import ArgumentError

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


def normalizeH(H, *, eps: float = 0.1, fixed_a: float = 0.0, center: bool = False):
    """Returns ``(a, H_norm)`` with H_norm = H/a, a = 2 Emax/(2-eps); or with
    ``center=True`` ``(a, b, H_norm)``, H_norm = (H - bI)/a,
    a = (Emax-Emin)/(2-eps), b = (Emax+Emin)/2; ``fixed_a`` skips the edge
    estimate (src/utils/external.jl:41-76)."""
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
