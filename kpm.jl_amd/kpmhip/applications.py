"""FP64 reconstruction heads that consume the moments: the DOS
(src/applications/dos.jl:76-199 with src/utils/chebyshev.jl:7-39 and
src/utils/chebyshev_lintrans.jl:6-27) and the Kubo–Bastin conductivity
(src/applications/conductivity.jl:101-157), plus the Fermi functions
(src/utils/physics.jl:15-35).

The heavy parts run on the device in FP64, as the reference's device methods
do: ``chebyshev_lin_trans`` (ext/KPMCUDAExt.jl:207-236) through
``kpm_cheb_lin_trans`` and the Kubo–Bastin node contraction through
``kpm_kubo_bastin_nodes``.  Node selection, weights and the energy-derivative
path (which the reference also keeps on the host, dos.jl:134-145) stay on the
host.
"""
from __future__ import annotations

import math
import warnings

import numpy as np

from ._lib import ArgumentError, check, dptr, i64ptr
from .device import get_context
from .kernels import JacksonKernel, hn, mu2D_apply_kernel_and_h


def chebyshevT(n, x):
    """T_n(x) = cos(n acos x) (src/utils/chebyshev.jl:7)."""
    return np.cos(np.multiply(n, np.arccos(x)))


def chebyshevT_0(n):
    """T_n(0): 0 odd, +1 n ≡ 0 (mod 4), -1 n ≡ 2 (mod 4) (src/utils/chebyshev.jl:10)."""
    n = np.asarray(n)
    return ((n + 1) % 2) * (2 - (n + 1) % 4)


def chebyshevT_xn(x_grid, n_grid):
    """T[x, n] matrix (src/utils/chebyshev.jl:33-39)."""
    x = np.atleast_1d(np.asarray(x_grid, dtype=np.float64))
    n = np.asarray(n_grid)
    return np.cos(n[None, :] * np.arccos(x)[:, None])


def chebyshev_lin_trans(x_grid, n_grid, mu_tilde):
    """Σ_n μ̃_n T_n(x) on a point or grid (src/utils/chebyshev_lintrans.jl:6-27),
    evaluated on the device (one wave per point, fixed-order sums)."""
    scalar = np.ndim(x_grid) == 0
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x_grid, dtype=np.float64)))
    n = np.ascontiguousarray(np.asarray(n_grid, dtype=np.int64).reshape(-1))
    m = np.ascontiguousarray(np.asarray(mu_tilde, dtype=np.complex128).reshape(-1))
    if m.shape[0] != n.shape[0]:
        raise ArgumentError("n_grid and mu_tilde must have the same length")
    y = np.zeros(x.shape[0], dtype=np.complex128)
    ctx = get_context()
    check(ctx.lib.kpm_cheb_lin_trans(ctx.handle, x.shape[0], dptr(x), n.shape[0], i64ptr(n),
                                     dptr(m), dptr(y)), "chebyshev_lin_trans")
    if not np.iscomplexobj(mu_tilde):
        y = y.real.copy()
    return y[0] if scalar else y


# ---- truncated Taylor arithmetic for dE_order derivatives ---------------------
# dos.jl:134-145 takes nested ForwardDiff derivatives of _dos_single; the same
# derivatives come exactly from Taylor coefficients in E (x = (E-b)/a is
# linear in E), with T_n built by the three-term recurrence on series.


def _tmul(a, b):
    k = len(a)
    out = np.zeros(k, dtype=np.result_type(a, b))
    for i in range(k):
        out[i] = np.dot(a[: i + 1], b[i::-1])
    return out


def _tsqrt(a):
    k = len(a)
    out = np.zeros(k, dtype=np.result_type(a, np.float64))
    out[0] = math.sqrt(a[0])
    for i in range(1, k):
        s = a[i] - np.dot(out[1:i], out[i - 1:0:-1])
        out[i] = s / (2 * out[0])
    return out


def _tdiv(a, b):
    k = len(a)
    out = np.zeros(k, dtype=np.result_type(a, b))
    for i in range(k):
        s = a[i] - np.dot(out[:i], b[i:0:-1])
        out[i] = s / b[0]
    return out


def _dos_taylor(mu_tilde, a, b, E, order):
    k = order + 1
    x = np.zeros(k)
    x[0] = (E - b) / a
    if k > 1:
        x[1] = 1.0 / a
    t_prev = np.zeros(k)
    t_prev[0] = 1.0
    acc = mu_tilde[0] * t_prev.astype(np.result_type(mu_tilde, np.float64))
    if len(mu_tilde) > 1:
        t_cur = x.copy()
        acc = acc + mu_tilde[1] * t_cur
        for n in range(2, len(mu_tilde)):
            t_next = 2.0 * _tmul(x, t_cur) - t_prev
            acc = acc + mu_tilde[n] * t_next
            t_prev, t_cur = t_cur, t_next
    one_minus = -_tmul(x, x)
    one_minus[0] += 1.0
    denom = a * math.pi * _tsqrt(one_minus)
    series = _tdiv(acc, denom)
    return series[order] * math.factorial(order)


def _dos_single(mu_tilde, a, b, E, NC):
    """ρ at one energy, no derivative (src/applications/dos.jl:151-161)."""
    x = (E - b) / a
    return chebyshev_lin_trans(x, np.arange(NC), mu_tilde) / (a * math.pi * math.sqrt(1 - x * x))


def dos(mu, H_rescale_factor, *, b=0.0, E_grid=None, N_tilde: int = 0, E_range=None, NC: int = 0,
        kernel=JacksonKernel, dE_order: int = 0):
    """ρ(E) = Σ_n h_n g_n μ_n T_n(x) / (aπ√(1-x²)), x = (E-b)/a; zero outside
    the band (src/applications/dos.jl:76-149).  Returns (E_grid, rho)."""
    mu = np.asarray(mu)
    if mu.ndim != 1:
        raise AssertionError("The input need to be 1D array")
    if not H_rescale_factor > 0:
        raise AssertionError("H_rescale_factor > 0")
    a = float(H_rescale_factor)
    if NC == 0:
        NC = len(mu)
    else:
        if NC > len(mu):
            warnings.warn(f"NC={NC} exceeds the number of moments; decreased to {len(mu)}.")
        NC = min(NC, len(mu))
    if E_grid is None:
        if N_tilde == 0:
            N_tilde = NC * 2
        if E_range is None:
            E_range = [b - 0.99 * a, b + 0.99 * a]
        E_grid = (np.arange(N_tilde + 1) * (E_range[1] - E_range[0])) / N_tilde + E_range[0]
    else:
        if E_range is not None:
            raise AssertionError("Should not set `E_grid` and `E_range` simultaneously.")
        E_grid = np.asarray(E_grid, dtype=np.float64)
        if N_tilde != 0 and len(E_grid) != N_tilde:
            raise AssertionError("`N_tilde` does not match with `E_grid`.")
    E_grid = np.asarray(E_grid, dtype=np.float64)
    rho_full = np.zeros(len(E_grid), dtype=np.float64)
    idx = np.abs(E_grid - b) < a
    x_in = (E_grid[idx] - b) / a
    ns = np.arange(NC)
    hgn = np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)
    mu_tilde = mu[:NC] * hgn
    if dE_order == 0:
        rho = chebyshev_lin_trans(x_in, ns, mu_tilde)
        rho = rho / (a * np.pi * np.sqrt(1 - x_in ** 2))
    else:
        rho = np.array([_dos_taylor(mu_tilde, a, b, E, dE_order) for E in E_grid[idx]])
    rho_full[idx] = np.real(rho)
    return E_grid, rho_full


def dos0(mu, H_rescale_factor, *, NC: int = 0, kernel=JacksonKernel, dE_order: int = 0):
    """DOS (or its second energy derivative) at x = 0 from the closed-form
    T_n(0) (src/applications/dos.jl:163-199)."""
    if not H_rescale_factor > 0:
        raise AssertionError("H_rescale_factor > 0")
    a = float(H_rescale_factor)
    mu = np.asarray(mu)
    if NC == 0:
        NC = len(mu)
    else:
        if NC > len(mu):
            warnings.warn(f"NC={NC} exceeds the number of moments; decreased to {len(mu)}.")
        NC = min(NC, len(mu))
    n0 = np.arange(0, NC, 4)
    n2 = np.arange(2, NC, 4)
    m0, m2 = mu[n0], mu[n2]
    k0 = np.asarray(kernel(n0, NC)) * hn(n0)
    k2 = np.asarray(kernel(n2, NC)) * hn(n2)
    if dE_order == 0:
        return (np.sum(m0 * k0) - np.sum(m2 * k2)) / (a * np.pi)
    if dE_order == 2:
        res = np.sum(m0 * k0 * (1 - n0 ** 2)) + np.sum(m2 * k2 * (n2 ** 2 - 1))
        return res / (a ** 3 * np.pi)
    raise NotImplementedError(f"unimplemented for derivative of order {dE_order}.")


# ---- Fermi functions (src/utils/physics.jl:15-35) ------------------------------


def fermiFunction(E, E_f, beta):
    with np.errstate(over="ignore"):  # exp overflow -> occupation 0, as intended
        return 1.0 / (np.exp((E - E_f) * beta) + 1.0)


def fermiFunctions(E_f: float, beta: float):
    E_f = float(E_f)
    beta = float(beta)
    if math.isinf(beta):
        return lambda x: ((np.asarray(x) < E_f).astype(np.float64) +
                          (np.asarray(x) <= E_f).astype(np.float64)) / 2.0
    return lambda x: fermiFunction(np.asarray(x, dtype=np.float64), E_f, beta)


def gausschebyshevt(n: int):
    """Gauss–Chebyshev (first kind) nodes and weights: x_k = cos((2k-1)π/(2n)),
    w_k = π/n (the rule FastGaussQuadrature.gausschebyshevt returns; nodes via
    sin for accuracy near ±1)."""
    k = np.arange(1, n + 1)
    x = np.sin(np.pi * (n - 2 * k + 1) / (2.0 * n))
    return x[::-1].copy(), np.full(n, np.pi / n)


def kubo_bastin_cond(mu2D, a: float, Ef: float, *, b: float = 0.0, NH: int, area: float,
                     beta: float = math.inf, kernel=JacksonKernel, NC: int | None = None,
                     quad_N: int | None = None, edge_cutoff: float = 1e-3) -> float:
    """σ_αβ(Ef) in e²/h from the NC x NC moments
    (src/applications/conductivity.jl:101-157)."""
    if not a > 0:
        raise ArgumentError("kubo_bastin_cond: a must be positive.")
    if not area > 0:
        raise ArgumentError("kubo_bastin_cond: area must be positive.")
    if not NH > 0:
        raise ArgumentError("kubo_bastin_cond: NH must be positive.")
    mu2D = np.asarray(mu2D)
    NC = mu2D.shape[0] if NC is None else NC
    quad_N = 8 * NC if quad_N is None else quad_N
    if not quad_N > 0:
        raise ArgumentError("kubo_bastin_cond: quad_N must be positive.")
    if not (0 <= edge_cutoff < 1):
        raise ArgumentError("kubo_bastin_cond: edge_cutoff must be in [0, 1).")
    NC = min(NC, mu2D.shape[0], mu2D.shape[1])
    mt = mu2D_apply_kernel_and_h(mu2D[:NC, :NC], NC, kernel)
    ff = fermiFunctions(Ef, beta)
    nodes, weights = gausschebyshevt(quad_N)
    fv = ff(a * nodes + b)
    keep = (np.abs(nodes) < 1 - edge_cutoff) & (fv != 0)
    if not np.any(keep):
        return 0.0
    x = nodes[keep]
    phi = weights[keep] * fv[keep] / (1 - x ** 2) ** 1.5
    mtf = np.asfortranarray(mt, dtype=np.complex128)
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ps = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros(2)
    ctx = get_context()
    check(ctx.lib.kpm_kubo_bastin_nodes(ctx.handle, NC, dptr(mtf), xs.shape[0], dptr(xs), dptr(ps),
                                        dptr(out)), "kubo_bastin_cond")
    acc = complex(out[0], out[1])
    return float(-2 * NH / (area * a ** 2) * acc.real)


# ---- Kubo–Bastin helpers (src/applications/conductivity.jl, dc_cond_util.jl) ---


def chebyshevT_accurate(n, x):
    """T_n(x) by the three-term recurrence, exact at x = 0 (src/utils/chebyshev.jl:23-29)."""
    n = int(n)
    if x == 0:
        return float(chebyshevT_0(n))
    t0, t1 = 1.0, float(x)
    if n == 0:
        return t0
    for _ in range(n - 1):
        t0, t1 = t1, 2.0 * x * t1 - t0
    return t1


def Gamma_nm(n, m, eps):
    """The Kubo–Bastin kernel Γnm(ε) of Garcia et al., PRL 114, 116602 (2015),
    elementwise over broadcast n, m (src/applications/dc_cond_util.jl:9-12)."""
    n = np.asarray(n)
    m = np.asarray(m)
    s = math.sqrt(1 - eps * eps)
    th = math.acos(eps)
    return ((eps - 1j * m * s) * np.exp(1j * m * th) * chebyshevT(n, eps) +
            (eps + 1j * n * s) * np.exp(-1j * n * th) * chebyshevT(m, eps))


def Gamma_nm_mu(mu_tilde, eps: float, NC: int):
    """Σ_nm Γnm(ε) μ̃nm by the rank-one factorisation of Γ: two matrix-vector
    products, no NC x NC Γ (src/applications/dc_cond_util.jl:21-32)."""
    mt = np.asarray(mu_tilde)
    if mt.shape != (NC, NC):
        raise AssertionError("size(μtilde) == (NC, NC)")
    th = math.acos(eps)
    s = math.sqrt(1 - eps * eps)
    ns = np.arange(NC)
    c = np.cos(ns * th)
    e = np.exp(1j * ns * th)
    w_m = (eps - 1j * ns * s) * e
    w_n = (eps + 1j * ns * s) * np.conj(e)
    return c @ (mt @ w_m) + w_n @ (mt @ c)


def _tmul_rows(A, B):
    """Row-wise truncated series products of (N, k) coefficient arrays."""
    k = A.shape[1]
    out = np.zeros((A.shape[0], k), dtype=np.result_type(A, B))
    for i in range(k):
        for j in range(i + 1):
            out[:, i] += A[:, j] * B[:, i - j]
    return out


def _dsigma_taylor(mu_tilde, a, b, E, order):
    """Taylor coefficients in E of Re[Σ Γnm μ̃nm] / ((1-x²)² a²), x = (E-b)/a:
    T_n and U_{n-1} by their recurrences on series, s = √(1-x²) as a series,
    e^{±inθ} = T_n ± i s U_{n-1}; returns the order-th E derivative."""
    k = order + 1
    NC = mu_tilde.shape[0]
    x = np.zeros(k)
    x[0] = (E - b) / a
    if k > 1:
        x[1] = 1.0 / a
    one_minus = -_tmul(x, x)
    one_minus[0] += 1.0
    s = _tsqrt(one_minus)
    T = np.zeros((NC, k))
    U = np.zeros((NC, k))  # U[n] = U_{n-1}(x); U_{-1} = 0
    T[0, 0] = 1.0
    if NC > 1:
        T[1] = x
        U[1, 0] = 1.0
    for n in range(2, NC):
        T[n] = 2.0 * _tmul(x, T[n - 1]) - T[n - 2]
        U[n] = 2.0 * _tmul(x, U[n - 1]) - U[n - 2]
    sU = _tmul_rows(np.broadcast_to(s, (NC, k)), U)
    ns = np.arange(NC)[:, None]
    e_plus = T + 1j * sU
    e_minus = T - 1j * sU
    xs = np.broadcast_to(x, (NC, k))
    w_m = _tmul_rows(xs - 1j * ns * np.broadcast_to(s, (NC, k)), e_plus)
    w_n = _tmul_rows(xs + 1j * ns * np.broadcast_to(s, (NC, k)), e_minus)
    mw = mu_tilde @ w_m
    mc = mu_tilde @ T
    acc = _tmul_rows(T.astype(np.complex128), mw).sum(axis=0) + _tmul_rows(w_n, mc).sum(axis=0)
    denom = _tmul(one_minus, one_minus) * a * a
    series = _tdiv(acc, denom.astype(np.complex128))
    return float(np.real(series[order])) * math.factorial(order)


def d_dc_cond(mu, a: float, E, *, b: float = 0.0, NC: int = 0, kernel=JacksonKernel,
              dE_order: int = 0) -> np.ndarray:
    """Kubo–Bastin integrand dσ(E) = Re[Σ_nm Γnm(x) μ̃nm] / ((1-x²)² a²),
    x = (E-b)/a, at the energies E (0 outside the band), or its dE_order-th
    energy derivative (src/applications/conductivity.jl:199-245; the
    reference's nested ForwardDiff derivatives are Taylor coefficients here)."""
    mu = np.asarray(mu)
    E_arr = np.atleast_1d(np.asarray(E, dtype=np.float64))
    if NC == 0:
        NC = mu.shape[0]
    else:
        if NC > mu.shape[0]:
            warnings.warn(f"NC={NC} exceeds the maximum size of μ; decreased to {mu.shape[0]}.")
        NC = min(mu.shape[0], NC)
    mt = mu2D_apply_kernel_and_h(mu[:NC, :NC], NC, kernel)
    out = np.zeros(E_arr.shape, dtype=np.float64)
    idx = np.abs(E_arr - b) < abs(a)
    for i in np.nonzero(idx)[0]:
        if dE_order == 0:
            eps = (E_arr[i] - b) / a
            out[i] = float(np.real(Gamma_nm_mu(mt, eps, NC) / ((1 - eps * eps) ** 2) / (a * a)))
        else:
            out[i] = _dsigma_taylor(mt, a, b, E_arr[i], int(dE_order))
    return out


def dc_cond0(mu, H_rescale_factor, *, kernel=JacksonKernel, NC: int | None = None):
    """Σ_nm μ̃nm T_n(0) T_m(0) / a by mod-4 index slices (not in physical
    units; src/applications/conductivity.jl:247-262)."""
    mu = np.asarray(mu)
    NC = mu.shape[0] if NC is None else int(NC)
    mt = mu2D_apply_kernel_and_h(mu[:NC, :NC], NC, kernel)
    oneone = mt[0::4, 0::4].sum()
    onethree = mt[0::4, 2::4].sum()
    threeone = mt[2::4, 0::4].sum()
    threethree = mt[2::4, 2::4].sum()
    return (oneone + threethree - onethree - threeone) / H_rescale_factor


def dc_cond_single(mu, H_rescale_factor, Ef, *, b=0.0, kernel=JacksonKernel,
                   NC: int | None = None):
    """Σ_nm μ̃nm T_n(x_F) T_m(x_F) / a with x_F = (Ef - b)/a (not in physical
    units; src/applications/conductivity.jl:264-285)."""
    mu = np.asarray(mu)
    NC = mu.shape[0] if NC is None else int(NC)
    ef = (Ef - b) / H_rescale_factor
    mt = mu2D_apply_kernel_and_h(mu[:NC, :NC], NC, kernel)
    t = np.array([chebyshevT_accurate(n, ef) for n in range(NC)])
    return (mt * t[:, None] * t[None, :]).sum() / H_rescale_factor
