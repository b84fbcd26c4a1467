"""Circular photogalvanic (three-current) response from the kpm_3d table
(src/applications/cpge.jl).

The hot part is the node function: at each quadrature node it evaluates
three trilinear forms of the kernel-weighted NC^3 table. The raw table is
uploaded once per call to a device-resident kpm_table3, which applies the
kernel weights on the device. Every refinement round of the adaptive
integral then evaluates all its nodes in one kpm_trilinear_forms_rec call:
the device expands each coefficient row from its recurrence seed (cos theta,
or the resolvent's g0 and q), then runs one FP64-MFMA pass over the table.
The host keeps the per-node seeds, the Fermi weight and the quadrature
bookkeeping (FP64).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import ArgumentError, DomainError, check, dptr
from .device import get_context
from .kernels import JacksonKernel, hn
from .spectral import (chebyshev_delta_theta, deduplicate_sorted, green_coefficients, green_seed,
                       spectral_integral)

# Paper coefficients (u, v, w) multiply package tensor slots (2, 1, 3):
# Gamma[n,m,p] = mu[m,n,p] in Joao-Lopes / Wu et al. (cpge.jl:1-4).
CPGE_SLOTS = (2, 1, 3)


class Table3:
    """A kernel-weighted NC^3 table resident on the device (kpm_table3),
    stored n fastest (Julia column-major mu[n, m, p]). With ``kvec`` the
    table is the raw moments and the device applies the per-axis weights
    k[n] = kernel(n, NC) h_n; without, it is uploaded as given."""

    def __init__(self, mu: np.ndarray, NC: int, kvec=None):
        self.ctx = get_context()
        self.NC = int(NC)
        host = np.asfortranarray(np.asarray(mu)[:NC, :NC, :NC], dtype=np.complex128)
        h = C.c_void_p()
        if kvec is None:
            check(self.ctx.lib.kpm_table3_create(self.ctx.handle, self.NC, dptr(host), C.byref(h)),
                  "kpm_table3_create")
        else:
            kv = np.ascontiguousarray(kvec, dtype=np.float64)
            if kv.shape != (self.NC,):
                raise ArgumentError("kernel vector must have NC entries")
            check(self.ctx.lib.kpm_table3_create_weighted(self.ctx.handle, self.NC, dptr(host), dptr(kv),
                                                          C.byref(h)), "kpm_table3_create_weighted")
        self.handle = h

    def forms_rec(self, da: np.ndarray, db: np.ndarray, dc: np.ndarray) -> np.ndarray:
        """Forms with device-generated coefficient rows from (Q, 5)
        descriptors (kind, p0r, p0i, p1r, p1i): kind 0 delta row (p0r =
        cos theta), 1 resolvent row (g0 = p0, q = p1), 2 zero row."""
        da, db, dc = (np.ascontiguousarray(x, dtype=np.float64) for x in (da, db, dc))
        Q = da.shape[0]
        if da.shape != (Q, 5) or db.shape != da.shape or dc.shape != da.shape:
            raise ArgumentError("descriptors must be Q x 5")
        out = np.zeros(Q, dtype=np.complex128)
        step = 1 << 20
        for q0 in range(0, Q, step):
            q1 = min(Q, q0 + step)
            part = np.zeros(q1 - q0, dtype=np.complex128)
            check(self.ctx.lib.kpm_trilinear_forms_rec(self.handle, q1 - q0, dptr(da[q0:q1]),
                                                       dptr(db[q0:q1]), dptr(dc[q0:q1]), dptr(part)),
                  "kpm_trilinear_forms_rec")
            out[q0:q1] = part
        return out

    def forms(self, a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
        """out[q] = sum mu~[n,m,p] a[q,n] b[q,m] c[q,p] (plain products)."""
        a = np.ascontiguousarray(a, dtype=np.complex128)
        b = np.ascontiguousarray(b, dtype=np.complex128)
        c = np.ascontiguousarray(c, dtype=np.complex128)
        Q = a.shape[0]
        if a.shape != (Q, self.NC) or b.shape != a.shape or c.shape != a.shape:
            raise ArgumentError("coefficient rows must be Q x NC")
        out = np.zeros(Q, dtype=np.complex128)
        step = 1 << 20
        for q0 in range(0, Q, step):
            q1 = min(Q, q0 + step)
            part = np.zeros(q1 - q0, dtype=np.complex128)
            check(self.ctx.lib.kpm_trilinear_forms(self.handle, q1 - q0, dptr(a[q0:q1]), dptr(b[q0:q1]),
                                                   dptr(c[q0:q1]), dptr(part)), "kpm_trilinear_forms")
            out[q0:q1] = part
        return out

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            check(self.ctx.lib.kpm_table3_destroy(self.handle), "kpm_table3_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check_cpge_inputs(mu3, NC, omega1, omega2, lam, delta):
    """_check_cpge_inputs (cpge.jl:6-28)."""
    if not NC > 0:
        raise ArgumentError("NC must be positive")
    mu3 = np.asarray(mu3)
    if mu3.ndim != 3:
        raise ArgumentError("mu3 must be a three-dimensional moment table")
    if not all(mu3.shape[d] >= NC for d in range(3)):
        raise ArgumentError("NC exceeds the supplied moment dimensions")
    if not lam >= 0:
        raise ArgumentError(f"lambda must be nonnegative (got {lam})")
    if not delta >= 0:
        raise ArgumentError(f"delta must be nonnegative (got {delta})")
    Omega = omega1 + omega2
    if lam == 0 and delta == 0 and any(x in (-2, 0, 2) for x in (omega1, omega2, Omega)):
        raise ArgumentError(
            "coincident shifted Green-function edges require lambda > 0 or delta > 0 when "
            "omega1, omega2, or Omega = omega1 + omega2 is 0 or +/-2; see the cpge docstring "
            "for the regularization recipe.")


def _cpge_green(z, NC, lam, delta, branch):
    """_cpge_green_coefficients! (cpge.jl:30-36): zero on the exclusion
    annulus 1 - delta < |z| < 1 + delta."""
    z = np.asarray(z, dtype=np.float64)
    out = np.zeros((z.size, NC), dtype=np.complex128)
    keep = ~((delta > 0) & (1 - delta < np.abs(z)) & (np.abs(z) < 1 + delta))
    if np.any(keep):
        out[keep] = green_coefficients(z[keep], NC, lam, branch)
    return out


def cpge_exclusion_breakpoints(shifts, delta):
    """_cpge_exclusion_breakpoints (cpge.jl:38-49)."""
    if delta == 0:
        return []
    pts = []
    for shift in shifts:
        for radius in (1 - delta, 1 + delta):
            if radius < 0:
                continue
            for edge in (-radius, radius):
                xstar = edge - shift
                if -1 < xstar < 1:
                    pts.append(math.acos(xstar))
    return deduplicate_sorted(pts)


def cpge_node_coefficients(theta, NC, omega1, omega2, lam, delta):
    """Slot coefficient rows of the three forms RR, RA, AA at each node
    (cpge.jl:66-120 with _cpge_slot_coefficients, :51-57): returns (a, b,
    c), each (3 N, NC), rows ordered node-major (RR, RA, AA) so that form
    q = 3 k + t belongs to node k."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    N = theta.size
    Omega = omega1 + omega2
    D = chebyshev_delta_theta(theta, NC).astype(np.complex128)
    x = np.cos(theta)
    gRO = _cpge_green(x + Omega, NC, lam, delta, "R")
    gR2 = _cpge_green(x + omega2, NC, lam, delta, "R")
    gR1 = _cpge_green(x + omega1, NC, lam, delta, "R")
    gAm2 = _cpge_green(x - omega2, NC, lam, delta, "A")
    gAm1 = _cpge_green(x - omega1, NC, lam, delta, "A")
    gAmO = _cpge_green(x - Omega, NC, lam, delta, "A")
    # (u, v, w) per form; slot1 = v (index n), slot2 = u (m), slot3 = w (p)
    u = np.stack([gRO, gR1, D], axis=1)
    v = np.stack([gR2, D, gAm1], axis=1)
    w = np.stack([D, gAm2, gAmO], axis=1)
    return (v.reshape(3 * N, NC), u.reshape(3 * N, NC), w.reshape(3 * N, NC))


def _green_desc(z, lam, delta, branch):
    """(N, 5) resolvent-row descriptors: (1, g0, q), or (2, ...) on the
    delta annulus."""
    z = np.asarray(z, dtype=np.float64)
    d = np.zeros((z.size, 5))
    d[:, 0] = 2.0
    keep = ~((delta > 0) & (1 - delta < np.abs(z)) & (np.abs(z) < 1 + delta))
    if np.any(keep):
        g0, q = green_seed(z[keep], lam, branch)
        d[keep, 0] = 1.0
        d[keep, 1], d[keep, 2] = g0.real, g0.imag
        d[keep, 3], d[keep, 4] = q.real, q.imag
    return d


def cpge_node_descriptors(theta, omega1, omega2, lam, delta):
    """Row descriptors of the forms RR, RA, AA at each node (the recurrence
    seeds of cpge_node_coefficients): (a, b, c), each (3 N, 5), node-major."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    N = theta.size
    Omega = omega1 + omega2
    x = np.cos(theta)
    D = np.zeros((N, 5))
    D[:, 1] = x
    gRO = _green_desc(x + Omega, lam, delta, "R")
    gR2 = _green_desc(x + omega2, lam, delta, "R")
    gR1 = _green_desc(x + omega1, lam, delta, "R")
    gAm2 = _green_desc(x - omega2, lam, delta, "A")
    gAm1 = _green_desc(x - omega1, lam, delta, "A")
    gAmO = _green_desc(x - Omega, lam, delta, "A")
    u = np.stack([gRO, gR1, D], axis=1)
    v = np.stack([gR2, D, gAm1], axis=1)
    w = np.stack([D, gAm2, gAmO], axis=1)
    return (v.reshape(3 * N, 5), u.reshape(3 * N, 5), w.reshape(3 * N, 5))


def _node_function(table: Table3, NC, omega1, omega2, lam, delta):
    def F(theta):
        theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
        a, b, c = cpge_node_descriptors(theta, omega1, omega2, lam, delta)
        forms = table.forms_rec(a, b, c).reshape(theta.size, 3)
        return (forms[:, 0] + forms[:, 1] + forms[:, 2])[:, None]  # rr + ra + aa

    return F


def _kernel_vector(kernel, NC):
    """k[n] = kernel(n, NC) h_n, the per-axis weight of
    mu3D_apply_kernel_and_h (cpge.jl:122-125)."""
    ns = np.arange(NC)
    return np.asarray(kernel(ns, NC), dtype=np.float64) * hn(ns)


def _weighted_table(mu3, NC, kernel):
    return Table3(np.asarray(mu3), NC, kvec=_kernel_vector(kernel, NC))


def cpge(mu3, NC: int, omega1_tilde: float, omega2_tilde: float, *, E_f: float = 0.0,
         beta: float = math.inf, lambda_: float = 0.0, delta: float = 0.0, kernel=JacksonKernel,
         quad_rtol: float = 1e-8, quad_atol: float = 0.0, maxevals: int = 10 ** 6) -> complex:
    """cpge (cpge.jl:127-220): i * integral over theta of
    f(cos theta) [RR + RA + AA](theta), in rescaled units. (``lambda_`` is
    the reference's ``lambda``: the Green broadening.) No Omega prefactor and
    no 1/(omega1 omega2); chi = D/(V a^2) * cpge(...)."""
    NC = int(NC)
    omega1_tilde, omega2_tilde = float(omega1_tilde), float(omega2_tilde)
    _check_cpge_inputs(mu3, NC, omega1_tilde, omega2_tilde, lambda_, delta)
    Omega = omega1_tilde + omega2_tilde
    shifts = (Omega, omega2_tilde, omega1_tilde, -omega2_tilde, -omega1_tilde, -Omega)
    kv = _kernel_vector(kernel, NC)
    # structural zero of mu~ (all moments zero, or every weight zero)
    exact_zero = not (np.any(np.asarray(mu3)[:NC, :NC, :NC]) and np.any(kv))
    with _weighted_table(mu3, NC, kernel) as table:
        F = _node_function(table, NC, omega1_tilde, omega2_tilde, lambda_, delta)
        integral, _ = spectral_integral(F, 1, NC, shifts, float(E_f), float(beta), float(lambda_),
                                        rtol=quad_rtol, atol=quad_atol, maxevals=maxevals,
                                        breakpoints=cpge_exclusion_breakpoints(shifts, delta),
                                        exact_zero=[exact_zero])
    return complex(1j * integral[0])


def d_cpge(mu3, NC: int, omega1_tilde: float, omega2_tilde: float, x=None, *, lambda_: float = 0.0,
           delta: float = 0.0, kernel=JacksonKernel, N_int: int | None = None, e_range=(-1.0, 1.0)):
    """d_cpge (cpge.jl:222-293): the energy-resolved bracket RR + RA + AA
    per unit dx, F(acos x)/sqrt(1 - x^2), without the Fermi weight or the
    factor i. With x: one complex value (-1 < x < 1). Without x: the grid
    form, returning (grid, values) on N_int (default 2 NC) midpoints of
    e_range."""
    NC = int(NC)
    omega1_tilde, omega2_tilde = float(omega1_tilde), float(omega2_tilde)
    _check_cpge_inputs(mu3, NC, omega1_tilde, omega2_tilde, lambda_, delta)
    if x is not None:
        if not (-1 < x < 1):
            raise DomainError(f"{x}: the resolved integrand requires -1 < x < 1")
        grid = np.array([float(x)])
    else:
        N_int = 2 * NC if N_int is None else int(N_int)
        if not N_int > 0:
            raise ArgumentError("N_int must be positive")
        if len(e_range) != 2:
            raise ArgumentError("e_range must have two endpoints")
        if not e_range[0] < e_range[1]:
            raise ArgumentError("e_range endpoints must be increasing")
        grid = ((np.arange(N_int) + 0.5) / N_int) * (e_range[1] - e_range[0]) + e_range[0]
        if not np.all((grid > -1) & (grid < 1)):
            raise DomainError(f"{list(e_range)}: resolved grid nodes must lie strictly inside (-1,1)")
    with _weighted_table(mu3, NC, kernel) as table:
        F = _node_function(table, NC, omega1_tilde, omega2_tilde, lambda_, delta)
        vals = F(np.arccos(grid))[:, 0] / np.sqrt(1 - grid ** 2)
    if x is not None:
        return complex(vals[0])
    return grid, vals
