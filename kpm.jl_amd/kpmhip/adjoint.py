"""Reverse-mode differentiation of the moments (no reference counterpart).

``kpm_1d_adjoint(H, NC, psi, g)`` returns dL/dH on H's stored entries for a
loss L(mu) of the probe-averaged equal-vector moments
mu_n = (1/NR) sum_r <psi_r|T_n(H)|psi_r> (the moments of kpm_1d), given
g = dL/dmu.  It runs libkpm_hip's reverse Chebyshev (Clenshaw) recurrence
with sqrt(NC) checkpointing on the GPU.

``moments_rrule`` packages it the way ChainRulesCore.rrule does for a Julia
function: it returns the primal moments and a pullback closure g -> dL/dH.
``dos_loss`` gives the L2 distance between the Jackson-smoothed DOS on an
energy grid and a target, with its gradient in the moments.
``site_gradient`` and ``bond_gradient`` map dL/dH_norm to per-site
potentials and per-bond hoppings of the physical Hamiltonian
(H_norm = (H - b I)/a, with a and b held fixed).
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, check, dptr
from .device import get_context, prec_code, to_device
from .kernels import JacksonKernel, hn
from .moment import kpm_1d


def _csc(H):
    A = sp.csc_matrix(H, dtype=np.complex128)
    A.sort_indices()
    return A


def kpm_1d_adjoint(H, NC: int, psi, g, *, precision="f64", ckpt_stride: int = 0) -> sp.csc_matrix:
    """dL/dH_ab on H's sparsity pattern (a CSC matrix with H's pattern):
    dL = sum_ab dH_ab * grad[a, b] for any perturbation dH on the pattern."""
    A = _csc(H)
    NC = int(NC)
    psi = np.asfortranarray(np.asarray(psi).reshape(A.shape[0], -1), dtype=np.complex128)
    NH, NR = psi.shape
    g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1))
    if g.shape[0] != NC:
        raise ArgumentError(f"g must have NC={NC} entries (got {g.shape[0]})")
    ctx = get_context()
    Hd = to_device(A, ctx)
    out = np.zeros(A.nnz, dtype=np.complex128)
    check(ctx.lib.kpm_1d_adjoint(ctx.handle, Hd.handle, NC, NR, NH, dptr(psi), dptr(g), dptr(out),
                                 prec_code(precision), int(ckpt_stride)), "kpm_1d_adjoint")
    return sp.csc_matrix((out, A.indices.copy(), A.indptr.copy()), shape=A.shape)


def moments_rrule(H, NC: int, psi, *, precision="f64"):
    """(mu, pullback): mu = probe-averaged kpm_1d moments (real, NC), and
    pullback(g) = dL/dH on H's pattern for g = dL/dmu."""
    mu = kpm_1d(H, NC, np.asarray(psi).shape[1], psi_in=psi, precision=precision)

    def pullback(g):
        return kpm_1d_adjoint(H, NC, psi, g, precision=precision)

    return mu, pullback


def dos_loss(mu, a: float, b: float, E, rho_target, kernel=JacksonKernel):
    """L = sum_k (rho(E_k) - rho*(E_k))^2 with
    rho(E) = sum_n h_n g_n mu_n T_n(x) / (a pi sqrt(1-x^2)), x = (E-b)/a
    (the reconstruction of src/applications/dos.jl:76-149); returns (L, dL/dmu)."""
    mu = np.asarray(mu, dtype=np.float64)
    NC = len(mu)
    E = np.asarray(E, dtype=np.float64)
    x = (E - b) / a
    if np.any(np.abs(x) >= 1):
        raise ArgumentError("energies must lie strictly inside the rescaled band")
    ns = np.arange(NC)
    w = np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)
    T = np.cos(ns[None, :] * np.arccos(x)[:, None])           # (nE, NC)
    D = T * w[None, :] / (a * math.pi * np.sqrt(1 - x * x))[:, None]
    rho = D @ mu
    r = rho - np.asarray(rho_target, dtype=np.float64)
    return float(np.sum(r * r)), 2.0 * (D.T @ r)


def site_gradient(grad_Hnorm, a: float) -> np.ndarray:
    """dL/deps_i for on-site potentials eps_i of H = ... + diag(eps), from
    dL/dH_norm (H_norm = (H - bI)/a): Re(grad_ii)/a.  The diagonal must be
    stored in H's pattern (explicit zeros are fine)."""
    G = sp.csc_matrix(grad_Hnorm)
    d = G.diagonal()
    return np.real(d) / a


def bond_gradient(H, grad_Hnorm, a: float) -> sp.coo_matrix:
    """dL/dt_ij for real bond amplitudes t_ij, H_ij = t_ij u_ij and
    H_ji = t_ij conj(u_ij) (u_ij the fixed unit phase of the stored entry),
    for every stored off-diagonal pair i < j: Re(u M_ij + conj(u) M_ji)/a
    (one O(nnz) pass in libkpm_hip over the shared CSC pattern)."""
    import ctypes as C

    from ._lib import i64ptr, load_library

    A = _csc(H)
    G = sp.csc_matrix(grad_Hnorm, dtype=np.complex128)
    G.sort_indices()
    if not (G.shape == A.shape and np.array_equal(G.indptr, A.indptr) and
            np.array_equal(G.indices, A.indices)):
        raise ArgumentError("grad_Hnorm must carry H's sparsity pattern")
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
    hz = np.ascontiguousarray(A.data)
    gz = np.ascontiguousarray(G.data)
    bi = np.empty(A.nnz, dtype=np.int64)
    bj = np.empty(A.nnz, dtype=np.int64)
    val = np.empty(A.nnz, dtype=np.float64)
    nb = C.c_int64(0)
    check(load_library().kpm_bond_gradient(A.shape[0], i64ptr(colptr), i64ptr(rowval), dptr(hz),
                                           dptr(gz), float(a), i64ptr(bi), i64ptr(bj), dptr(val),
                                           C.byref(nb)), "bond_gradient")
    m = nb.value
    return sp.coo_matrix((val[:m], (bi[:m], bj[:m])), shape=A.shape)
