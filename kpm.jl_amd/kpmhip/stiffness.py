"""Superfluid stiffness and the generic two-energy response
(src/applications/stiffness.jl).

The moment tables come from the device engine: superfluid_stiffness makes
two kpm_2d calls, on the superconducting and the normal-state BdG operators
(stiffness.jl:606,621). diamagnetic_term makes one kpm_1d_current call
(:393). Both run on the assembled, rescaled Nambu matrices (ScaledOperator
of a BdGOperator), uploaded once per call.

The vertices (nambu_current_q, nambu_diamagnetic) are O(nnz) host builders
of sparse matrices. The two-energy reconstruction is host FP64 on the
NC x NC table, as in the reference.
"""
from __future__ import annotations

import math
import warnings
from types import SimpleNamespace

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, KPMUnstableError
from .applications import fermiFunctions, gausschebyshevt
from .bdg import BdGOperator, ScaledOperator, _canonical_hole_convention, chebyshev_stability_probe
from .bdg import rescale_bdg
from .device import get_context
from .kernels import JacksonKernel, hn, mu2D_apply_kernel_and_h
from .moment import kpm_1d_current, kpm_2d_


def two_energy_response(mu2D, a: float | None = None, *, b: float | None = None, beta: float,
                        eta: float, omega: float = 0.0, Ef: float = 0.0, NH: int | None = None,
                        volume: float, g_J: float = 1.0, kernel=JacksonKernel, NC: int | None = None,
                        Np: int | None = None) -> complex:
    """two_energy_response (stiffness.jl:37-150):

        Pi(omega) = (g_J/volume) sum_{p,q} (Ja)_pq (Jb)_qp (f_p - f_q) /
                    (omega + E_p - E_q + i eta)

    from mu2D[n,m] = Tr[Ja T_{m-1} Jb T_{n-1}]/D. Gauss-Chebyshev nodes
    absorb both spectral weights. With eta = 0 only the static
    finite-temperature case is defined (the Fermi divided difference, with
    -beta f (1 - f) at coincident energies). The typed form takes a
    ConductivityMoments as the first argument (src/frontend.jl:1083-1105)."""
    from .frontend import ConductivityMoments

    if isinstance(mu2D, ConductivityMoments):
        m = mu2D
        if a is not None:
            raise ArgumentError("a is stored in the moments object; do not pass it separately")
        if b is not None:
            raise ArgumentError("b is stored in the moments object; do not pass it separately")
        if NH is not None:
            raise ArgumentError("NH is stored in the moments object; do not pass it separately")
        return two_energy_response(m.mu, m.a, b=m.b, beta=beta, eta=eta, omega=omega, Ef=Ef, NH=m.NH,
                                   volume=volume, g_J=g_J, kernel=kernel, NC=NC, Np=Np)
    b = 0.0 if b is None else float(b)
    mu2D = np.asarray(mu2D)
    if mu2D.ndim != 2 or mu2D.shape[0] != mu2D.shape[1]:
        raise ArgumentError(f"two_energy_response: mu2D must be square (got {mu2D.shape})")
    NC = mu2D.shape[0] if NC is None else int(NC)
    Np = 2 * NC if Np is None else int(Np)
    if NC < 2:
        raise ArgumentError(f"two_energy_response: NC must be at least 2 (got {NC})")
    if mu2D.shape[0] < NC:
        raise ArgumentError(f"two_energy_response: mu2D size {mu2D.shape} is smaller than NC={NC}")
    if a is None or not a > 0:
        raise ArgumentError(f"two_energy_response: a must be positive (got {a})")
    if not eta >= 0:
        raise ArgumentError(f"two_energy_response: eta must be nonnegative (got {eta})")
    if not beta > 0:
        raise ArgumentError(f"two_energy_response: beta must be positive (got {beta})")
    if not Np > 0:
        raise ArgumentError(f"two_energy_response: Np must be positive (got {Np})")
    if NH is None or not NH > 0:
        raise ArgumentError(f"two_energy_response: NH must be positive (got {NH})")
    if not volume > 0:
        raise ArgumentError(f"two_energy_response: volume must be positive (got {volume})")
    if eta == 0 and not (omega == 0 and math.isfinite(beta)):
        raise ArgumentError("eta=0 requires omega=0 and finite beta")
    mt = mu2D_apply_kernel_and_h(mu2D[:NC, :NC], NC, kernel)
    nodes, _ = gausschebyshevt(Np)
    Cm = np.cos(np.arange(NC)[:, None] * np.arccos(nodes)[None, :])  # NC x Np
    Gamma = Cm.T @ mt @ Cm
    E = a * nodes + b
    fermi = fermiFunctions(float(Ef), float(beta))
    f = fermi(E)
    dE = E[:, None] - E[None, :]
    if eta == 0:
        thr = 1e-8 * max(a, 1.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            F = (f[:, None] - f[None, :]) / dE
        close = np.abs(dE) < thr
        if np.any(close):
            fm = fermi((E[:, None] + E[None, :]) / 2)
            F = np.where(close, -beta * fm * (1 - fm), F)
    else:
        F = (f[:, None] - f[None, :]) / (omega + dE + 1j * eta)
    return complex(g_J * NH / volume / Np ** 2 * np.sum(Gamma * F))


def _bond_lists(h, pos, disp, ndim):
    Hc = sp.coo_matrix(h)
    I, J, V = Hc.row, Hc.col, Hc.data.astype(np.complex128)
    if disp is None:
        d = pos[I, :] - pos[J, :]
    else:
        d = np.array([np.asarray(disp(int(i), int(j)), dtype=np.float64) for i, j in zip(I, J)])
        if d.size and d.shape[1] != ndim:
            raise ArgumentError(f"disp has length {d.shape[1]}; expected {ndim}")
        d = d.reshape(len(I), ndim)
    keep = ~np.all(d == 0, axis=1)
    return I[keep], J[keep], V[keep], d[keep]


def _check_vertex_inputs(h, pos, caller, dir, q=None):
    N = h.shape[0]
    if h.shape[1] != N:
        raise ArgumentError(f"{caller}: h must be square (got {h.shape})")
    pos = np.asarray(pos, dtype=np.float64)
    if pos.shape[0] != N:
        raise ArgumentError(f"{caller}: pos has {pos.shape[0]} rows; expected {N}")
    ndim = pos.shape[1]
    if q is not None and len(q) != ndim:
        raise ArgumentError(f"{caller}: q has length {len(q)}; expected {ndim}")
    if not (1 <= dir <= ndim):
        raise ArgumentError(f"{caller}: dir must satisfy 1 <= dir <= {ndim} (got {dir})")
    return N, pos, ndim


def nambu_current_q(h, pos, q, *, dir: int = 1, disp=None, hole_convention: str = "conjugate"):
    """nambu_current_q (stiffness.jl:152-228): J(q)_ij = h_ij d_dir
    e^{-i q.m_ij}, J(q)_(i+N,j+N) = j_hole e^{-i q.m_ij} with m_ij = r_j +
    d/2, d = r_i - r_j (minimum image through disp), j_hole = conj(h_ij d_dir)
    ('conjugate') or h_ij d_dir ('intervalley'); zero-displacement entries
    are skipped. 2N x 2N CSC complex; dir is 1-based."""
    N, pos, ndim = _check_vertex_inputs(h, pos, "nambu_current_q", dir, q)
    hc = _canonical_hole_convention(hole_convention, caller="nambu_current_q")
    q = np.asarray(q, dtype=np.float64)
    I, J, V, d = _bond_lists(h, pos, disp, ndim)
    phase = np.exp(-1j * ((pos[J, :] + d / 2) @ q))
    bond = V * d[:, dir - 1]
    hole = np.conj(bond) if hc == "conjugate" else bond
    rows = np.concatenate([I, I + N])
    cols = np.concatenate([J, J + N])
    vals = np.concatenate([bond * phase, hole * phase])
    return sp.csc_matrix((vals, (rows, cols)), shape=(2 * N, 2 * N))


def nambu_diamagnetic(h, pos, *, dir: int = 1, disp=None, hole_convention: str = "conjugate"):
    """nambu_diamagnetic (stiffness.jl:230-308): Dhat_ij = -h_ij d_dir^2,
    Dhat_(i+N,j+N) = +conj(h_ij) d_dir^2 ('conjugate') or +h_ij d_dir^2
    ('intervalley'); zero-displacement entries skipped."""
    N, pos, ndim = _check_vertex_inputs(h, pos, "nambu_diamagnetic", dir)
    hc = _canonical_hole_convention(hole_convention, caller="nambu_diamagnetic")
    I, J, V, d = _bond_lists(h, pos, disp, ndim)
    d2 = d[:, dir - 1] ** 2
    part = -V * d2
    hole = (np.conj(V) if hc == "conjugate" else V) * d2
    rows = np.concatenate([I, I + N])
    cols = np.concatenate([J, J + N])
    return sp.csc_matrix((np.concatenate([part, hole]), (rows, cols)), shape=(2 * N, 2 * N))


def _unstable(stability):
    raise KPMUnstableError(
        f"Chebyshev recurrence is unstable (maximum probe norm {stability} > 1.5); use "
        "rescale(...; bound=:gershgorin) or a larger eps.")


def _probes(psi_in, rng, NH, NR, caller):
    from .frontend import random_phase_vectors
    from .rng import Xoshiro256pp

    if psi_in is None:
        psi_in = random_phase_vectors(Xoshiro256pp(0) if rng is None else rng, NH, NR)
    else:
        psi_in = np.asarray(psi_in)
        NR = psi_in.shape[1]
    if psi_in.shape != (NH, NR):
        raise ArgumentError(f"{caller}: psi_in has size {psi_in.shape}; expected ({NH}, {NR})")
    return np.asfortranarray(psi_in, dtype=np.complex128), NR


def diamagnetic_term(op: BdGOperator, pos, *, dir: int = 1, disp=None, beta: float, NC: int = 256,
                     NR: int = 8, rng=None, psi_in=None, volume: float, g_J: float = 1.0,
                     kernel=JacksonKernel, Np: int | None = None, rescale_eps: float = 0.2,
                     a: float | None = None, Hs=None) -> float:
    """diamagnetic_term (stiffness.jl:310-404): (g_J/volume) sum_n f(E_n)
    <n|Dhat|n> at BdG Fermi level zero, from the device moments
    Gamma_m = Tr[Dhat T_m(Hs)]/(2N) (kpm_1d_current) and a Gauss-Chebyshev
    reconstruction."""
    if not sp.issparse(op.h) and not isinstance(op.h, np.ndarray):
        raise ArgumentError("diamagnetic_term needs an assembled sparse h to build the vertex")
    if np.asarray(pos).shape[0] != op.N:
        raise ArgumentError(f"diamagnetic_term: pos has {np.asarray(pos).shape[0]} rows; expected {op.N}")
    if not beta > 0:
        raise ArgumentError(f"diamagnetic_term: beta must be positive (got {beta})")
    if not volume > 0:
        raise ArgumentError(f"diamagnetic_term: volume must be positive (got {volume})")
    if not NC >= 2:
        raise ArgumentError(f"diamagnetic_term: NC must be at least 2 (got {NC})")
    Np = 2 * NC if Np is None else int(Np)
    if not Np > 0:
        raise ArgumentError(f"diamagnetic_term: Np must be positive (got {Np})")
    if not NR > 0:
        raise ArgumentError(f"diamagnetic_term: NR must be positive (got {NR})")
    if not (0 < rescale_eps < 2):
        raise ArgumentError(f"diamagnetic_term: rescale_eps must satisfy 0 < rescale_eps < 2 "
                            f"(got {rescale_eps})")
    Dhat = nambu_diamagnetic(sp.csc_matrix(op.h), pos, dir=dir, disp=disp,
                             hole_convention=op.hole_convention)
    scale = rescale_bdg(op, eps=float(rescale_eps)).a if a is None else float(a)
    if not (math.isfinite(scale) and scale > 0):
        raise ArgumentError(f"diamagnetic_term: a must be finite and positive (got {scale})")
    if Hs is None:
        Hs = ScaledOperator(op, scale, 0.0)
    elif Hs.a != scale:
        raise ArgumentError(f"diamagnetic_term: supplied Hs has a={Hs.a} but the requested scale is {scale}")
    NH = 2 * op.N
    NC = int(NC)
    Hd = Hs.device_operator(get_context())
    stability = chebyshev_stability_probe(Hd, NH, NC)
    if not stability <= 1.5:
        _unstable(stability)
    psi_in, NR = _probes(psi_in, rng, NH, int(NR), "diamagnetic_term")
    Gamma = np.real(kpm_1d_current(Hd, Dhat, NC, NR, NH, psi_in=psi_in))
    gh = np.asarray(kernel(np.arange(NC), NC), dtype=np.float64) * hn(np.arange(NC))
    nodes, _ = gausschebyshevt(Np)
    wf = fermiFunctions(0.0, float(beta))(scale * nodes) / Np
    Cm = np.cos(np.arange(NC)[:, None] * np.arccos(nodes)[None, :])
    Dia_raw = float(np.dot(gh * Gamma, Cm @ wf))
    return float(g_J * NH / volume * Dia_raw)


def superfluid_stiffness(op: BdGOperator, pos, q, *, beta: float, eta: float, omega: float = 0.0,
                         dir: int = 1, disp=None, NC: int = 256, NR: int = 8, rng=None, psi_in=None,
                         volume: float, g_J: float = 1.0, kernel=JacksonKernel, Np: int | None = None,
                         moment_parity="NONE", arr_size="auto", right_block="auto",
                         rescale_eps: float = 0.2, verbose: int = 0,
                         include_diamagnetic: bool = False):
    """superfluid_stiffness (stiffness.jl:406-728): D_s/pi = Re Pi_N -
    Re Pi_SC with Ja = J(q), Jb = J(-q), common probes and the common scale
    a = max(a_SC, a_N); the normal reference keeps h, mu, U and n and zeroes
    the pairing. Both kpm_2d tables run on the device. Returns a namespace
    with the reference's NamedTuple fields (plus Dia_SC, Dia_N and
    Ds_over_pi_complete with include_diamagnetic).

    arr_size / right_block default to the engine's slot plan (every slot
    resident when the workspace allows); the reference's defaults are 3 and
    16, and any blocking gives the same tables to rounding."""
    if not (sp.issparse(op.h) or isinstance(op.h, np.ndarray)):
        raise ArgumentError(
            "superfluid_stiffness needs an assembled sparse h to build vertices; matrix-free users "
            "can call nambu_current_q-equivalent vertices + kpm_2d + two_energy_response directly")
    pos = np.asarray(pos, dtype=np.float64)
    if pos.shape[0] != op.N:
        raise ArgumentError(f"superfluid_stiffness: pos has {pos.shape[0]} rows; expected {op.N}")
    if not (0 < rescale_eps < 2):
        raise ArgumentError(f"superfluid_stiffness: rescale_eps must satisfy 0 < rescale_eps < 2 "
                            f"(got {rescale_eps})")
    op_n = BdGOperator(op.h, mu=op.mu, U=op.U.copy(), n=op.n.copy(),
                       Delta=np.zeros(op.N, dtype=np.complex128), hole_convention=op.hole_convention,
                       h_hole=op.h_hole if op.hole_convention == "conjugate" else None,
                       assume_intervalley=op.hole_convention == "intervalley")
    h_sparse = sp.csc_matrix(op.h)
    q = np.asarray(q, dtype=np.float64)
    Jq = nambu_current_q(h_sparse, pos, q, dir=dir, disp=disp, hole_convention=op.hole_convention)
    Jmq = nambu_current_q(h_sparse, pos, -q, dir=dir, disp=disp, hole_convention=op.hole_convention)
    qn = float(np.linalg.norm(q))
    if qn > 0 and abs(q[dir - 1]) > 1e-12 * qn:
        warnings.warn("superfluid_stiffness: q has a longitudinal component; transverse stiffness "
                      "requires q perpendicular to dir")
    rh_sc = rescale_bdg(op, eps=float(rescale_eps))
    rh_n = rescale_bdg(op_n, eps=float(rescale_eps))
    a_common = max(rh_sc.a, rh_n.a)
    ctx = get_context()
    Hs_sc = ScaledOperator(op, a_common, 0.0)
    Hs_n = ScaledOperator(op_n, a_common, 0.0)
    Hd_sc, Hd_n = Hs_sc.device_operator(ctx), Hs_n.device_operator(ctx)
    NH = 2 * op.N
    psi_in, NR = _probes(psi_in, rng, NH, int(NR), "superfluid_stiffness")
    NC = int(NC)
    stab = max(chebyshev_stability_probe(Hd_sc, NH, NC), chebyshev_stability_probe(Hd_n, NH, NC))
    if not stab <= 1.5:
        _unstable(stab)
    mu_sc = np.zeros((NC, NC), dtype=np.complex128, order="F")
    mu_n = np.zeros((NC, NC), dtype=np.complex128, order="F")
    kw = dict(arr_size=arr_size, right_block=right_block, verbose=verbose, moment_parity=moment_parity)
    kpm_2d_(Hd_sc, Jq, Jmq, NC, NR, NH, mu_sc, psi_in, psi_in, **kw)
    kpm_2d_(Hd_n, Jq, Jmq, NC, NR, NH, mu_n, psi_in, psi_in, **kw)
    Np = 2 * NC if Np is None else int(Np)
    resp = dict(b=0.0, beta=beta, eta=eta, omega=omega, Ef=0.0, NH=NH, volume=volume, g_J=g_J,
                kernel=kernel, Np=Np)
    Pi_SC = two_energy_response(mu_sc, a_common, **resp)
    Pi_N = two_energy_response(mu_n, a_common, **resp)
    out = dict(Ds_over_pi=Pi_N.real - Pi_SC.real, Pi_SC=Pi_SC, Pi_N=Pi_N, a_SC=rh_sc.a, a_N=rh_n.a,
               a_common=a_common, q=q.copy(), dir=int(dir), eta=float(eta), beta=float(beta),
               omega=float(omega), NC=NC, NR=NR, Np=Np)
    if include_diamagnetic:
        dkw = dict(dir=dir, disp=disp, beta=beta, NC=NC, NR=NR, psi_in=psi_in, volume=volume, g_J=g_J,
                   kernel=kernel, Np=Np, rescale_eps=rescale_eps, a=a_common)
        Dia_SC = diamagnetic_term(op, pos, Hs=Hs_sc, **dkw)
        Dia_N = diamagnetic_term(op_n, pos, Hs=Hs_n, **dkw)
        out.update(Dia_SC=Dia_SC, Dia_N=Dia_N, Ds_over_pi_complete=out["Ds_over_pi"] + Dia_SC - Dia_N)
    return SimpleNamespace(**out)
