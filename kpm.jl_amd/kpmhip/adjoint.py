"""Reverse-mode differentiation of the moments (no reference counterpart).

``kpm_1d_adjoint(H, NC, psi, g)`` returns dL/dH on H's stored entries for a
loss L(mu) of the probe-averaged equal-vector moments
mu_n = (1/NR) sum_r <psi_r|T_n(H)|psi_r> (the moments of kpm_1d), given
g = dL/dmu.  It runs libkpm_hip's reverse Chebyshev (Clenshaw) recurrence
with sqrt(NC) checkpointing on the GPU.

``kpm_1d_adjoint_forward`` / ``AdjointTape.pullback`` are its two halves;
``moments_rrule`` packages them the way ChainRulesCore.rrule does for a
Julia function: the primal moments and a pullback closure g -> dL/dH from
one forward pass.
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
from .device import DeviceCSR, get_context, prec_code, to_device
from .kernels import JacksonKernel, hn


def _csc(H):
    A = sp.csc_matrix(H, dtype=np.complex128)
    A.sort_indices()
    return A


def kpm_1d_adjoint(H, NC: int, psi, g, *, precision="f64", ckpt_stride: int = 0,
                   operator=None) -> sp.csc_matrix:
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
    Hd = _operator(A, operator, ctx)
    out = np.zeros(A.nnz, dtype=np.complex128)
    ctx._adjoint_tape = None
    check(ctx.lib.kpm_1d_adjoint(ctx.handle, Hd.handle, NC, NR, NH, dptr(psi), dptr(g), dptr(out),
                                 prec_code(precision), int(ckpt_stride)), "kpm_1d_adjoint")
    return sp.csc_matrix((out, A.indices.copy(), A.indptr.copy()), shape=A.shape)


class AdjointTape:
    """The forward half of kpm_1d_adjoint, kept on the device.

    ``kpm_1d_adjoint_forward`` runs the forward recurrence once, keeps its
    vectors (or sqrt(NC) checkpoints) in the context and returns this tape
    with the moments of the same passes; ``pullback(g)`` runs the reverse
    pass on them.  The context keeps one forward pass: a later
    ``kpm_1d_adjoint_forward`` or ``kpm_1d_adjoint`` on it retires this tape,
    and its pullback then raises instead of differentiating the wrong pass.
    """

    def __init__(self, ctx, Hd, A, NC, NR, mu_all):
        self.ctx, self.Hd, self.A = ctx, Hd, A
        self.NC, self.NR = NC, NR
        self.mu_all = mu_all              # NR x NC, <psi_r|T_n(H)|psi_r>
        self.mu = mu_all.real.mean(axis=0)  # the probe-averaged moments of kpm_1d
        self.token = object()
        ctx._adjoint_tape = self.token

    def pullback(self, g) -> sp.csc_matrix:
        """dL/dH on H's pattern for g = dL/dmu (NC reals)."""
        g = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(-1))
        if g.shape[0] != self.NC:
            raise ArgumentError(f"g must have NC={self.NC} entries (got {g.shape[0]})")
        if getattr(self.ctx, "_adjoint_tape", None) is not self.token:
            raise ArgumentError("this forward pass was replaced by a later one on the same context")
        out = np.zeros(self.A.nnz, dtype=np.complex128)
        check(self.ctx.lib.kpm_1d_adjoint_reverse(self.ctx.handle, self.Hd.handle, dptr(g), dptr(out)),
              "kpm_1d_adjoint_reverse")
        return sp.csc_matrix((out, self.A.indices.copy(), self.A.indptr.copy()), shape=self.A.shape)


def _operator(A, operator, ctx):
    """The device operator of A: a caller's DeviceCSR built from the same
    matrix (reused across calls), or a fresh upload."""
    if operator is None:
        return to_device(A, ctx)
    if not isinstance(operator, DeviceCSR) or operator.adjoint:
        raise ArgumentError("operator must be the DeviceCSR of H (to_device(H))")
    if operator.shape != A.shape or operator.nnz != A.nnz or operator.ctx is not ctx:
        raise ArgumentError("operator does not match H (shape, stored entries or context)")
    return operator


def kpm_1d_adjoint_forward(H, NC: int, psi, *, precision="f64", ckpt_stride: int = 0,
                           operator=None) -> AdjointTape:
    """Forward half of kpm_1d_adjoint: one recurrence pass that keeps what the
    reverse pass needs and returns the moments of the same passes (tape.mu,
    tape.mu_all).  ``operator`` may pass H's DeviceCSR (to_device(H)) to skip
    the per-call upload; it must come from the same matrix."""
    A = _csc(H)
    NC = int(NC)
    if NC < 2:
        raise ArgumentError(f"NC must be >= 2 (got {NC})")
    psi = np.asfortranarray(np.asarray(psi).reshape(A.shape[0], -1), dtype=np.complex128)
    NH, NR = psi.shape
    ctx = get_context()
    Hd = _operator(A, operator, ctx)
    mu_all = np.zeros((NC, NR), dtype=np.complex128)  # column-major NR x NC
    ctx._adjoint_tape = None
    check(ctx.lib.kpm_1d_adjoint_forward(ctx.handle, Hd.handle, NC, NR, NH, dptr(psi), dptr(mu_all),
                                         prec_code(precision), int(ckpt_stride)),
          "kpm_1d_adjoint_forward")
    return AdjointTape(ctx, Hd, A, NC, NR, np.ascontiguousarray(mu_all.T))


def moments_rrule(H, NC: int, psi, *, precision="f64", ckpt_stride: int = 0, operator=None):
    """(mu, pullback): mu = probe-averaged equal-vector moments (real, NC,
    the moments of kpm_1d to rounding), and pullback(g) = dL/dH on H's
    pattern for g = dL/dmu.  The primal and the pullback share one forward
    pass."""
    tape = kpm_1d_adjoint_forward(H, NC, psi, precision=precision, ckpt_stride=ckpt_stride,
                                  operator=operator)
    return tape.mu, tape.pullback


_DOS_CACHE: dict = {}


def _dos_matrix(NC, a, b, E, kernel):
    """D[k, n] = h_n g_n T_n(x_k) / (a pi sqrt(1 - x_k^2)), x = (E - b)/a; the
    last few (NC, a, b, E, kernel) are kept, since a training loop asks for the
    same grid every step."""
    key = (NC, a, b, E.tobytes(), kernel)
    D = _DOS_CACHE.get(key)
    if D is None:
        x = (E - b) / a
        if np.any(np.abs(x) >= 1):
            raise ArgumentError("energies must lie strictly inside the rescaled band")
        ns = np.arange(NC)
        w = np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)
        T = np.cos(ns[None, :] * np.arccos(x)[:, None])           # (nE, NC)
        D = T * w[None, :] / (a * math.pi * np.sqrt(1 - x * x))[:, None]
        D.setflags(write=False)
        if len(_DOS_CACHE) >= 4:
            _DOS_CACHE.pop(next(iter(_DOS_CACHE)))
        _DOS_CACHE[key] = D
    return D


def dos_loss(mu, a: float, b: float, E, rho_target, kernel=JacksonKernel):
    """L = sum_k (rho(E_k) - rho*(E_k))^2 with
    rho(E) = sum_n h_n g_n mu_n T_n(x) / (a pi sqrt(1-x^2)), x = (E-b)/a
    (the reconstruction of src/applications/dos.jl:76-149); returns (L, dL/dmu)."""
    mu = np.asarray(mu, dtype=np.float64)
    E = np.asarray(E, dtype=np.float64)
    D = _dos_matrix(len(mu), float(a), float(b), E, kernel)
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


# ---- reverse mode through kpm_2d and the Kubo-Bastin head ---------------------


def kpm_2d_adjoint(H, Ja, Jb, NC: int, psi_l, psi_r, gbar):
    """Derivatives of a real loss through the kpm_2d table
    mu[n,m] = (1/NR) sum_r <T_m(H) psi_l_r | Jb T_n(H) Ja psi_r_r>
    (src/moment.jl:987-1088) for the cotangent gbar (NC x NC complex,
    dL = Re sum conj(gbar) dmu).  Returns (gH, gJa, gJb): CSC matrices on the
    patterns of H, Ja and Jb with dL = Re sum_k g[k] dA[k] (FP64, GPU)."""
    A, Aa, Ab = _csc(H), _csc(Ja), _csc(Jb)
    NC = int(NC)
    n = A.shape[0]
    pl = np.asfortranarray(np.asarray(psi_l).reshape(n, -1), dtype=np.complex128)
    pr = np.asfortranarray(np.asarray(psi_r).reshape(n, -1), dtype=np.complex128)
    if pl.shape != pr.shape:
        raise ArgumentError(f"psi_l and psi_r must have equal size (got {pl.shape} and {pr.shape})")
    gb = np.asfortranarray(np.asarray(gbar, dtype=np.complex128))
    if gb.shape != (NC, NC):
        raise ArgumentError(f"gbar must be {NC} x {NC} (got {gb.shape})")
    from .device import to_device_adjoint

    ctx = get_context()
    Hd, Jad, Jbd = to_device(A, ctx), to_device(Aa, ctx), to_device(Ab, ctx)
    Jbh = to_device_adjoint(Ab, ctx)
    outs = [np.zeros(X.nnz, dtype=np.complex128) for X in (A, Aa, Ab)]
    check(ctx.lib.kpm_2d_adjoint(ctx.handle, Hd.handle, Jad.handle, Jbd.handle, Jbh.handle, NC,
                                 pl.shape[1], n, dptr(pl), dptr(pr), dptr(gb), dptr(outs[0]),
                                 dptr(outs[1]), dptr(outs[2])), "kpm_2d_adjoint")
    return tuple(sp.csc_matrix((o, X.indices.copy(), X.indptr.copy()), shape=X.shape)
                 for o, X in zip(outs, (A, Aa, Ab)))


def kubo_bastin_weights(NC: int, a: float, Ef: float, *, b: float = 0.0, NH: int, area: float,
                        beta: float = math.inf, kernel=JacksonKernel, quad_N: int | None = None,
                        edge_cutoff: float = 1e-3) -> np.ndarray:
    """The NC x NC complex K with kubo_bastin_cond(mu, ...) = Re sum K * mu
    (src/applications/conductivity.jl:101-157 is linear in the moments):
    K[n,m] = -2 NH/(area a^2) w_n w_m sum_k phi_k Gamma_nm(x_k), with the same
    Gauss-Chebyshev nodes, Fermi weights and edge cutoff as the head and the
    kernel weights w_n = g_n h_n.  Its conjugate is the cotangent of mu for
    the loss sigma."""
    from .applications import fermiFunctions, gausschebyshevt

    NC = int(NC)
    quad_N = 8 * NC if quad_N is None else int(quad_N)
    ns = np.arange(NC)
    w = np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)
    nodes, weights = gausschebyshevt(quad_N)
    fv = fermiFunctions(Ef, beta)(a * nodes + b)
    keep = (np.abs(nodes) < 1 - edge_cutoff) & (fv != 0)
    if not np.any(keep):
        return np.zeros((NC, NC), dtype=np.complex128)
    x = nodes[keep]
    phi = weights[keep] * fv[keep] / (1 - x ** 2) ** 1.5
    th = np.arccos(x)[None, :]
    s = np.sqrt(1 - x ** 2)[None, :]
    Cn = np.cos(ns[:, None] * th)
    Em = np.exp(1j * ns[:, None] * th)
    WM = (x[None, :] - 1j * ns[:, None] * s) * Em        # Gamma's m-factor of the T_n term
    WN = (x[None, :] + 1j * ns[:, None] * s) * np.conj(Em)  # n-factor of the T_m term
    G = (Cn * phi[None, :]) @ WM.T + (WN * phi[None, :]) @ Cn.T
    return (-2.0 * NH / (area * a * a)) * (w[:, None] * w[None, :]) * G


def sigma_rrule(H, Ja, Jb, NC: int, psi, a: float, Ef: float, *, b: float = 0.0, area: float,
                psi_r=None, **head_kwargs):
    """(sigma, pullback): sigma = kubo_bastin_cond of the FP64 kpm_2d table of
    the rescaled H (src/frontend.jl:270-327 + src/applications/
    conductivity.jl:101-157), and pullback(sbar) = (gH, gJa, gJb), the
    derivatives on the stored entries of H, Ja and Jb (dsigma = Re sum g dA)."""
    from .moment import kpm_2d

    psi = np.asarray(psi)
    psi_r = psi if psi_r is None else np.asarray(psi_r)
    NH = H.shape[0]
    mu = kpm_2d(H, Ja, Jb, int(NC), psi.shape[1], psi_in_l=psi, psi_in_r=psi_r, gemm="f64")
    K = kubo_bastin_weights(NC, a, Ef, b=b, NH=NH, area=area, **head_kwargs)
    sigma = float(np.real(np.sum(K * mu)))

    def pullback(sbar=1.0):
        return kpm_2d_adjoint(H, Ja, Jb, NC, psi, psi_r, float(sbar) * np.conj(K))

    return sigma, pullback


def hopping_gradient(H, Ja, Jb, gH, gJa, gJb, a: float):
    """Per-site and per-bond derivatives from the entry derivatives of
    kpm_2d_adjoint, for the parametrisation of a tight-binding model whose
    current operators are built from its hoppings: H = diag(eps) + sum_b t_b
    (h_b + h_b^H) and J_x = sum_b t_b (j_b + j_b^H), so H_ij / t_b and
    J_ij / t_b are fixed (t_b = |H_ij|, J on H's pattern or a subset).  The
    table is of H_norm = (H - b I)/a with a, b fixed, so gH is per H_norm
    entry.  Returns (d/deps_i for every site, COO of d/dt_b for every stored
    pair i < j)."""
    A = _csc(H)
    inv = _inv_magnitude(A)
    unit = sp.csc_matrix(A.multiply(inv))               # H_ij / |H_ij|
    contrib = sp.csc_matrix(sp.csc_matrix(gH).multiply(unit)).real / a
    for J, gJ in ((Ja, gJa), (Jb, gJb)):
        Jhat = sp.csc_matrix(_csc(J).multiply(inv))     # J_ij / |H_ij|
        contrib = contrib + sp.csc_matrix(sp.csc_matrix(gJ).multiply(Jhat)).real
    site = np.real(sp.csc_matrix(gH).diagonal()) / a
    C = sp.csc_matrix(contrib)
    bonds = sp.triu(C + C.T, k=1).tocoo()
    return site, bonds


def _inv_magnitude(A):
    """1/|A_ij| on A's pattern (0 where A_ij = 0)."""
    M = _csc(A)
    m = np.abs(M.data)
    M.data = np.where(m > 0, 1.0 / np.where(m > 0, m, 1.0), 0.0).astype(np.complex128)
    return M
