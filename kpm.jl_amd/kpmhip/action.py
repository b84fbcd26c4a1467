"""Coefficient-accumulating matrix-function action and its consumers.

``chebyshev_action`` / ``chebyshev_action_`` mirror chebyshev_action /
chebyshev_action! (src/utils/chebyshev_action.jl:47-159):

    out[:, :, k] = sum_{n=1..NC} C[n, k] T_{n-1}(Hn) V

computed by libkpm_hip's ``kpm_cheb_action`` (windowed accumulation over a
ring of recurrence slots, relative stability guard on the device).  On top of
it, as in the reference:

- time evolution: evolution_tail / evolution_order / evolution_coefficients /
  evolve_ / evolve (src/applications/evolution.jl:1-170, src/frontend.jl:725-760);
- the Fermi operator: fermi_coefficients / fermi_projector
  (src/applications/chern_marker.jl:9-82, src/frontend.jl:783-800);
- the Bianco-Resta local Chern marker: chern_marker / chern_marker_region /
  chern_marker_average (src/applications/chern_marker.jl:84-152,
  src/frontend.jl:835-981), whose two projector actions, diagonal position
  scalings and column contraction all stay on the device.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
from scipy.special import jv

from ._lib import ArgumentError, check, dptr
from .applications import fermiFunctions, gausschebyshevt
from .device import DeviceBlock, DeviceCSR, get_context, prec_code, to_device
from .kernels import JacksonKernel, hn

_EPS = np.finfo(np.float64).eps
_TINY = np.finfo(np.float64).tiny


def _shares(a, b) -> bool:
    return isinstance(a, np.ndarray) and isinstance(b, np.ndarray) and np.shares_memory(a, b)


def _as_matrix(V) -> np.ndarray:
    a = np.asarray(V)
    return a.reshape(-1, 1) if a.ndim == 1 else a


def chebyshev_action_(out, Hn, V, C, *, slots=None, check_every: int = 16, verbose: int = 0,
                      precision="f64"):
    """chebyshev_action! (src/utils/chebyshev_action.jl:47-125): overwrite
    ``out`` (NH x NR x K complex; NH x NR when C is a vector) with
    sum_n C[n, k] T_{n-1}(Hn) V.  ``slots`` is validated like the
    reference's two preallocated workspaces; the recurrence itself runs in
    the library's own slot ring.  ``precision`` selects the recurrence
    vectors (FP64 or FP32); the accumulation is FP64."""
    Vm = _as_matrix(V)
    Cm = np.asarray(C)
    vec_c = Cm.ndim == 1
    if vec_c:
        Cm = Cm.reshape(-1, 1)
    out3 = out.reshape(out.shape[0], out.shape[1], 1) if (vec_c and out.ndim == 2) else out
    NH, NR = Vm.shape
    NC, K = Cm.shape
    if NC < 1:
        raise ArgumentError("C must have at least one coefficient row")
    if NR < 1:
        raise ArgumentError("V must have at least one column")
    if K < 1:
        raise ArgumentError("C must have at least one coefficient column")
    if Hn.shape[1] != NH:
        raise ArgumentError(f"size(Hn, 2) = {Hn.shape[1]} must equal size(V, 1) = {NH}")
    if not (out3.ndim == 3 and out3.shape[0] == NH and out3.shape[1] == NR):
        raise ArgumentError(f"out must be {NH} × {NR} × K, got {out.shape}")
    if out3.shape[2] != K:
        raise ArgumentError(f"size(out, 3) = {out3.shape[2]} must equal size(C, 2) = {K}")
    if check_every < 0:
        raise ArgumentError("check_every must be >= 0")
    if _shares(out, V):
        raise ArgumentError("out must not alias V")
    if _shares(out, Hn):
        raise ArgumentError("out must not alias Hn")
    if slots is not None:
        if len(slots) != 2:
            raise ArgumentError("slots must hold two workspaces")
        for s in slots:
            if not (np.shape(s) == (NH, NR) and np.iscomplexobj(s)):
                raise ArgumentError(f"slots must be two complex {NH} × {NR} workspaces")
        if (_shares(slots[0], slots[1]) or any(_shares(s, out) or _shares(s, V) for s in slots)
                or any(_shares(s, Hn) for s in slots)):
            raise ArgumentError("slots must not alias each other, out, V, or Hn")
    if not np.iscomplexobj(out):
        raise ArgumentError("out must be complex")
    ctx = get_context()
    Hd = to_device(Hn, ctx)
    Vh = np.asfortranarray(Vm, dtype=np.complex128)
    Ch = np.asfortranarray(Cm, dtype=np.complex128)
    res = np.zeros((NH, NR, K), dtype=np.complex128, order="F")
    if verbose >= 1:
        print(f"chebyshev_action: NC = {NC}, K = {K}")
    check(ctx.lib.kpm_cheb_action(ctx.handle, Hd.handle, NC, K, NR, NH, dptr(Vh), dptr(Ch),
                                  dptr(res), prec_code(precision), int(check_every)),
          "chebyshev_action!")
    out3[...] = res
    return out


def chebyshev_action(Hn, V, C, **kwargs) -> np.ndarray:
    """Allocating form (src/utils/chebyshev_action.jl:148-159): matrix V and
    matrix C give NH x NR x K; a vector C drops the last axis, a vector V the
    probe axis."""
    Vm = _as_matrix(V)
    Cm = np.asarray(C)
    K = 1 if Cm.ndim == 1 else Cm.shape[1]
    out = np.zeros((Vm.shape[0], Vm.shape[1], K), dtype=np.complex128, order="F")
    chebyshev_action_(out, Hn, Vm, Cm.reshape(Cm.shape[0], -1) if Cm.ndim == 1 else Cm, **kwargs)
    if Cm.ndim == 1:
        out = out[:, :, 0]
    if np.ndim(V) == 1:
        out = out[:, 0] if out.ndim == 2 else out[:, 0, :]
    return out


def action_block(Hd: DeviceCSR, V: DeviceBlock, C, out: DeviceBlock, *, check_every: int = 16):
    """Device-resident action: ``out`` (FP64 block of (K*NH) x NR) receives
    sum_n C[n, k] T_{n-1}(Hd) V in rows [k*NH, (k+1)*NH)."""
    Cm = np.asarray(C, dtype=np.complex128)
    Cm = np.asfortranarray(Cm.reshape(Cm.shape[0], -1))
    NC, K = Cm.shape
    ctx = V.ctx
    check(ctx.lib.kpm_cheb_action_block(ctx.handle, Hd.handle, NC, K, V.handle, dptr(Cm),
                                        out.handle, int(check_every)), "chebyshev_action!")
    return out


# ---- time evolution (src/applications/evolution.jl) -------------------------


def evolution_tail(z: float, NC: int) -> float:
    """2 sum_{m >= NC} |J_m(z)|, summed until three consecutive terms fall
    below the running sum's ulp (src/applications/evolution.jl:20-37)."""
    z = abs(float(z))
    if not math.isfinite(z):
        raise ArgumentError(f"z must be finite, got {z}")
    if NC < 1:
        raise ArgumentError(f"NC must be >= 1, got {NC}")
    s = 0.0
    m = int(NC)
    negligible = 0
    while negligible < 3:
        term = 2.0 * abs(float(jv(m, z)))
        s += term
        negligible = negligible + 1 if term <= _EPS * max(s, _TINY) else 0
        m += 1
    return s


def _capcheck(x, NC_cap):
    if not (math.isfinite(x) and x < NC_cap):
        raise ArgumentError(
            f"evolution order would exceed NC_cap = {NC_cap} (reached {x}); pass NC explicitly "
            "or split the propagation into shorter times")


def evolution_order(a: float, t: float, *, tol: float = 1e-12, NC_min: int = 8,
                    NC_cap: int = 10_000_000) -> int:
    """Smallest NC >= NC_min whose tail bound is below tol
    (src/applications/evolution.jl:49-86)."""
    if not (math.isfinite(a) and a > 0):
        raise ArgumentError(f"a must be finite and positive, got {a}")
    if not math.isfinite(t):
        raise ArgumentError(f"t must be finite, got {t}")
    if not 0 < tol < 1:
        raise ArgumentError(f"tol must be in (0, 1), got {tol}")
    if NC_min < 2:
        raise ArgumentError(f"NC_min must be >= 2, got {NC_min}")
    if NC_min > NC_cap:
        raise ArgumentError(f"NC_min = {NC_min} must not exceed NC_cap = {NC_cap}")
    z = abs(a * t)
    _capcheck(z, NC_cap)
    n = math.ceil(z)
    while not (abs(jv(n, z)) < tol / 2 and abs(jv(n + 1, z)) < tol / 2):
        n += 1
        _capcheck(n, NC_cap)
    NC = n + 1
    while evolution_tail(z, NC) >= tol:
        NC += 1
        _capcheck(NC, NC_cap)
    return max(NC, int(NC_min))


def evolution_coefficients(a: float, b: float, ts, *, NC: int = 0, tol: float = 1e-12):
    """(C, NC, tail): C[n, k] = e^{-i b t_k} (2 - δ_{n1}) (-i)^{n-1} J_{n-1}(a t_k)
    (src/applications/evolution.jl:103-144)."""
    ts = np.atleast_1d(np.asarray(ts, dtype=np.float64))
    if not (math.isfinite(a) and a > 0):
        raise ArgumentError(f"a must be finite and positive, got {a}")
    if not math.isfinite(b):
        raise ArgumentError(f"b must be finite, got {b}")
    if ts.size == 0:
        raise ArgumentError("ts must not be empty")
    if not np.all(np.isfinite(ts)):
        raise ArgumentError("all times must be finite")
    if not 0 < tol < 1:
        raise ArgumentError(f"tol must be in (0, 1), got {tol}")
    adaptive = NC == 0
    if adaptive:
        NC = max(evolution_order(a, float(t), tol=tol) for t in ts)
    elif NC < 2:
        raise ArgumentError(f"NC must be >= 2 (or 0 for adaptive selection), got {NC}")
    NT = ts.size
    C = np.zeros((NC, NT), dtype=np.complex128, order="F")
    tail = np.zeros(NT)
    n = np.arange(NC)
    minus_i_pow = (-1j) ** (n % 4)
    for k, t in enumerate(ts):
        z = a * t
        phase = np.exp(-1j * b * t)
        Jn = jv(n, z)
        C[:, k] = phase * np.where(n == 0, 1.0, 2.0) * minus_i_pow * Jn
        tail[k] = evolution_tail(z, NC)
    if not adaptive and tail.max() > tol:
        import warnings
        warnings.warn(f"evolution series tail estimate {tail.max()} exceeds tol = {tol}; "
                      "increase NC or tol")
    return C, NC, tail


def evolve_(out, H_norm, a: float, b: float, psi0, ts, *, NC: int = 0, tol: float = 1e-12,
            check_every: int = 16, verbose: int = 0, precision="f64"):
    """evolve! (src/applications/evolution.jl:154-170)."""
    C, NC_used, tail = evolution_coefficients(a, b, ts, NC=NC, tol=tol)
    if verbose >= 1:
        print(f"evolve: NC = {NC_used}, max tail estimate = {tail.max()}")
    chebyshev_action_(out, H_norm, psi0, C, check_every=check_every, verbose=verbose,
                      precision=precision)
    return out


def evolve(h, psi0, ts, *, NC: int = 0, tol: float = 1e-12, check_every: int = 16,
           verbose: int = 0, precision="f64") -> np.ndarray:
    """e^{-iHt} psi0 for every t (src/frontend.jl:725-760): NH, NH x NR,
    NH x NT or NH x NR x NT with singleton axes of vector psi0 / scalar t
    dropped."""
    NH = h.H.shape[0]
    if np.shape(psi0)[0] != NH:
        raise ArgumentError(f"psi0 has {np.shape(psi0)[0]} rows; expected NH = {NH}")
    scalar_t = np.ndim(ts) == 0
    tv = np.atleast_1d(np.asarray(ts, dtype=np.float64))
    if tv.size == 0:
        raise ArgumentError("ts must not be empty")
    psi_mat = _as_matrix(psi0)
    out = np.zeros((NH, psi_mat.shape[1], tv.size), dtype=np.complex128, order="F")
    evolve_(out, h.H, h.a, h.b, psi_mat, tv, NC=NC, tol=tol, check_every=check_every,
            verbose=verbose, precision=precision)
    if np.ndim(psi0) == 1:
        out = out[:, 0, :]
    if scalar_t:
        out = out[..., 0]
    return out


# ---- Fermi operator and Chern marker (src/applications/chern_marker.jl) -----


def fermi_coefficients(a: float, b: float, Ef: float, *, beta: float = math.inf, NC: int,
                       kernel=JacksonKernel, Np: int | None = None) -> np.ndarray:
    """Kernel-damped Chebyshev coefficients of f_beta(H - Ef) in
    H_norm = (H - b)/a, (2 - δ_{n0}) folded in
    (src/applications/chern_marker.jl:44-82)."""
    Np = 2 * NC if Np is None else Np
    if not (math.isfinite(a) and a > 0):
        raise ArgumentError(f"a must be finite and positive, got {a}")
    if not math.isfinite(b):
        raise ArgumentError(f"b must be finite, got {b}")
    if not math.isfinite(Ef):
        raise ArgumentError(f"Ef must be finite, got {Ef}")
    if not beta > 0:
        raise ArgumentError(f"beta must be positive, got {beta}")
    if NC < 1:
        raise ArgumentError(f"NC must be >= 1, got {NC}")
    if Np < 1:
        raise ArgumentError(f"Np must be >= 1, got {Np}")
    xF = (Ef - b) / a
    if not -1 < xF < 1:
        raise ArgumentError(
            f"Ef = {Ef} maps to x̃_F = {xF} outside the rescaled spectral window (-1, 1)")
    n = np.arange(NC)
    if math.isinf(beta):
        thF = math.acos(xF)
        c = np.empty(NC)
        c[0] = 1 - thF / math.pi
        c[1:] = -(2 / math.pi) * np.sin(n[1:] * thF) / n[1:]
    else:
        nodes, _ = gausschebyshevt(Np)
        wf = fermiFunctions(float(Ef), float(beta))(a * nodes + b) / Np
        c = hn(n) * (np.cos(n[:, None] * np.arccos(nodes)[None, :]) @ wf)
    return c * np.asarray(kernel(n, NC), dtype=np.float64)


def fermi_projector(h, V, *, Ef: float, beta: float = math.inf, NC: int, kernel=JacksonKernel,
                    check_every: int = 16, verbose: int = 0, precision="f64") -> np.ndarray:
    """P V with P = f_beta(H - Ef) (src/frontend.jl:783-800)."""
    NH = h.H.shape[0]
    if np.shape(V)[0] != NH:
        raise ArgumentError(f"V has {np.shape(V)[0]} rows; expected NH = {NH}")
    C = fermi_coefficients(h.a, h.b, Ef, beta=beta, NC=NC, kernel=kernel)
    Vm = _as_matrix(V)
    out = np.zeros(Vm.shape, dtype=np.complex128, order="F")
    chebyshev_action_(out, h.H, Vm, C, check_every=check_every, verbose=verbose,
                      precision=precision)
    return out[:, 0] if np.ndim(V) == 1 else out


def _diag_operator(d, ctx) -> DeviceCSR:
    n = len(d)
    return DeviceCSR(sp.csc_matrix((np.asarray(d, dtype=np.complex128), (np.arange(n), np.arange(n))),
                                   shape=(n, n)), ctx)


def _pxqyp_imdiag(Hd: DeviceCSR, C, Xd: DeviceCSR, Yd: DeviceCSR, V: np.ndarray, ctx, *,
                  check_every: int = 16) -> np.ndarray:
    """r[c] = sum_j x_j Im(conj(U[j,c]) W[j,c]) with U = P V and
    W = P (y .* U) (src/applications/chern_marker.jl:84-129), all on the
    device: two actions, two diagonal SpMMs and one column contraction."""
    NH, nb = V.shape
    Vb = DeviceBlock.from_host(V, "f64", ctx)
    U = DeviceBlock(NH, nb, "f64", ctx)
    YU = DeviceBlock(NH, nb, "f64", ctx)
    W = DeviceBlock(NH, nb, "f64", ctx)
    action_block(Hd, Vb, C, U, check_every=check_every)
    check(ctx.lib.kpm_spmm(ctx.handle, Yd.handle, 1.0, U.handle, 0.0, YU.handle), "y .* U")
    action_block(Hd, YU, C, W, check_every=check_every)
    XW = YU  # y .* U is no longer needed
    check(ctx.lib.kpm_spmm(ctx.handle, Xd.handle, 1.0, W.handle, 0.0, XW.handle), "x .* W")
    tgt = np.zeros((nb, 1), dtype=np.complex128, order="F")
    check(ctx.lib.kpm_colwise_dot(ctx.handle, dptr(tgt), 1, 1, U.handle, XW.handle, 1.0, 0, 0.0),
          "column contraction")
    return tgt[:, 0].imag.copy()


def _coords(x, y, NH):
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if len(x) != NH:
        raise ArgumentError(f"x has length {len(x)}; expected NH = {NH}")
    if len(y) != NH:
        raise ArgumentError(f"y has length {len(y)}; expected NH = {NH}")
    if not (np.all(np.isfinite(x)) and np.all(np.isfinite(y))):
        raise ArgumentError("coordinate vectors must be finite")
    return x, y


def chern_marker(h, x, y, *, Ef: float, sites, beta: float = math.inf, NC: int,
                 kernel=JacksonKernel, batch_size: int = 32, check_every: int = 16,
                 verbose: int = 0) -> np.ndarray:
    """Bianco-Resta local marker m_i = -4π Im<i|P X Q Y P|i> at the 1-based
    basis indices ``sites`` (src/frontend.jl:835-891)."""
    NH = h.H.shape[0]
    x, y = _coords(x, y, NH)
    sites = np.asarray(sites, dtype=np.int64).reshape(-1)
    if sites.size == 0:
        raise ArgumentError("sites must not be empty")
    if not np.all((sites >= 1) & (sites <= NH)):
        raise ArgumentError(f"sites must lie in 1:{NH}")
    if batch_size < 1:
        raise ArgumentError("batch_size must be >= 1")
    C = fermi_coefficients(h.a, h.b, Ef, beta=beta, NC=NC, kernel=kernel)
    ctx = get_context()
    Hd = to_device(h.H, ctx)
    Xd, Yd = _diag_operator(x, ctx), _diag_operator(y, ctx)
    markers = np.empty(sites.size)
    for lo in range(0, sites.size, batch_size):
        hi = min(lo + batch_size, sites.size)
        V = np.zeros((NH, hi - lo), dtype=np.complex128, order="F")
        V[sites[lo:hi] - 1, np.arange(hi - lo)] = 1.0
        markers[lo:hi] = 4 * math.pi * _pxqyp_imdiag(Hd, C, Xd, Yd, V, ctx,
                                                     check_every=check_every)
    return markers


def chern_marker_region(h, x, y, *, Ef: float, region, rng, NR: int = 16,
                        beta: float = math.inf, NC: int, kernel=JacksonKernel,
                        batch_size: int = 32, check_every: int = 16,
                        verbose: int = 0) -> np.ndarray:
    """NR stochastic estimates of the region-summed marker from random-phase
    probes supported on ``region`` (src/frontend.jl:917-981)."""
    from .frontend import random_phase_vectors

    NH = h.H.shape[0]
    x, y = _coords(x, y, NH)
    region = np.asarray(region, dtype=np.int64).reshape(-1)
    if region.size == 0:
        raise ArgumentError("region must not be empty")
    if not np.all((region >= 1) & (region <= NH)):
        raise ArgumentError(f"region must lie in 1:{NH}")
    if np.unique(region).size != region.size:
        raise ArgumentError("region must not contain duplicates")
    if NR < 1:
        raise ArgumentError("NR must be >= 1")
    if batch_size < 1:
        raise ArgumentError("batch_size must be >= 1")
    C = fermi_coefficients(h.a, h.b, Ef, beta=beta, NC=NC, kernel=kernel)
    ctx = get_context()
    Hd = to_device(h.H, ctx)
    Xd, Yd = _diag_operator(x, ctx), _diag_operator(y, ctx)
    phases = random_phase_vectors(rng, region.size, int(NR))
    est = np.empty(NR)
    for lo in range(0, NR, batch_size):
        hi = min(lo + batch_size, NR)
        V = np.zeros((NH, hi - lo), dtype=np.complex128, order="F")
        V[region - 1, :] = phases[:, lo:hi]
        est[lo:hi] = 4 * math.pi * region.size * _pxqyp_imdiag(Hd, C, Xd, Yd, V, ctx,
                                                                check_every=check_every)
    return est


def chern_marker_average(markers, *, area: float) -> float:
    """sum(markers) / area (src/applications/chern_marker.jl:146-152)."""
    if not (math.isfinite(area) and area > 0):
        raise ArgumentError(f"area must be finite and positive, got {area}")
    return float(np.sum(np.atleast_1d(markers))) / area
