"""Spectral-integral machinery shared by the Green-function reconstructions
(src/applications/spectral_coefficients.jl): Chebyshev delta and resolvent
coefficients, the Fermi weight, and the adaptive integral over theta in
(0, pi) with shifted-edge breakpoints and endpoint maps.

The adaptive rule restates QuadGK's h-adaptive scheme, which the reference
calls with order 7 (spectral_coefficients.jl:135-344):
- a (7-point Gauss, 15-point Kronrod) pair on every segment;
- the segment error is norm(I_K - I_G);
- the segment with the largest error is bisected until E <= atol or
  E <= rtol * norm(I), or the evaluation budget is spent.

In the reference the pieces of one pass are separate quadgk calls, run one
after another. Here they advance in lockstep: every piece still follows the
serial rule (same segments, same bisection order), but the nodes of one
refinement round of all pieces are handed to the node function as one batch,
which is what lets a device contraction evaluate them in a single launch.

Node functions take an array of theta values (N,) and return (N, ncomp)
complex values.
"""
from __future__ import annotations

import heapq
import math

import numpy as np

from ._lib import ArgumentError, DomainError, QuadratureError

# 15-point Gauss-Kronrod pair (QUADPACK qk15, Piessens et al. 1983): Kronrod
# abscissae in (0, 1] descending, then the centre; Gauss weights belong to
# abscissae 1, 3, 5 (0-based) and the centre.
_XGK = np.array([
    0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
    0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
    0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
    0.207784955007898467600689403773245, 0.0])
_WGK = np.array([
    0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
    0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
    0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
    0.204432940075298892414161999234649, 0.209482141084727828012999174891714])
_WG = np.array([
    0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
    0.381830050505118944950369775488975, 0.417959183673469387755102040816327])
# abscissa offsets of one segment, in the order the values are stored:
# 7 pairs (a + (1 - x) s, a + (1 + x) s), then the centre
_OFFS = np.concatenate([np.stack([1.0 - _XGK[:7], 1.0 + _XGK[:7]], axis=1).ravel(), [1.0]])
NODES_PER_SEGMENT = 15


def gauss_kronrod_15():
    """(abscissae, Kronrod weights, Gauss weights) of the 15-point rule on
    [-1, 1]: nodes in increasing order; Gauss weights are zero on the
    Kronrod-only nodes."""
    x = np.concatenate([-_XGK[:7], [0.0], _XGK[6::-1]])
    wk = np.concatenate([_WGK[:7], [_WGK[7]], _WGK[6::-1]])
    wg = np.zeros(15)
    for i, k in enumerate((1, 3, 5)):
        wg[k] = wg[14 - k] = _WG[i]
    wg[7] = _WG[3]
    return x, wk, wg


def chebyshev_delta_theta(theta, NC: int) -> np.ndarray:
    """delta[:, n] = cos(n theta)/pi by the recurrence of
    chebyshev_delta_theta! (spectral_coefficients.jl:22-33); theta (N,) ->
    (N, NC) real."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    out = np.empty((theta.size, NC))
    if NC == 0:
        return out
    c = np.cos(theta)
    out[:, 0] = 1.0 / math.pi
    if NC == 1:
        return out
    out[:, 1] = c / math.pi
    for n in range(2, NC):
        out[:, n] = 2.0 * c * out[:, n - 1] - out[:, n - 2]
    return out


def green_seed(z, lam: float, branch: str):
    """(g0, q) of the retarded ("R") or advanced ("A") resolvent coefficients
    at real energies z (green_coefficients!, spectral_coefficients.jl:43-82):
    g[n] = g[n-1] q with g[0] = g0."""
    if not lam >= 0:
        raise ArgumentError(f"lambda must be nonnegative (got {lam})")
    if branch not in ("R", "A"):
        raise ArgumentError("branch must be 'R' or 'A'")
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    g0 = np.empty(z.size, dtype=np.complex128)
    q = np.empty(z.size, dtype=np.complex128)
    if lam == 0:
        if np.any(np.abs(z) == 1):
            raise DomainError("z = +/-1 is singular when lambda = 0")
        inside = np.abs(z) < 1
        phi = np.arccos(np.where(inside, z, 0.0))
        s = np.sin(phi)
        sign = -1.0 if branch == "R" else 1.0
        g0[inside] = sign * 1j / s[inside]
        q[inside] = np.exp(sign * 1j * phi[inside])
        up = z > 1
        lo = z < -1
        g0[up] = 1.0 / np.sqrt(z[up] ** 2 - 1)
        q[up] = np.exp(-np.arccosh(z[up]))
        g0[lo] = -1.0 / np.sqrt(z[lo] ** 2 - 1)
        q[lo] = -np.exp(-np.arccosh(-z[lo]))
    elif branch == "R":
        w = z + 1j * lam
        phi = np.arccos(w)
        g0[:] = -1j / np.sqrt(1 - w * w)
        q[:] = np.exp(-1j * phi)
    else:
        w = z - 1j * lam
        phi = np.arccos(w)
        g0[:] = 1j / np.sqrt(1 - w * w)
        q[:] = np.exp(1j * phi)
    return g0, q


def green_coefficients(z, NC: int, lam: float, branch: str) -> np.ndarray:
    """Chebyshev coefficients of the resolvent at real energies z
    (green_coefficients!, spectral_coefficients.jl:43-87): g[n] = g[0] q^n
    by repeated multiplication; z (N,) -> (N, NC) complex."""
    g0, q = green_seed(z, lam, branch)
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    if NC == 0:
        return np.empty((z.size, 0), dtype=np.complex128)
    seq = np.empty((z.size, NC), dtype=np.complex128)
    seq[:, 0] = g0
    seq[:, 1:] = q[:, None]
    return np.cumprod(seq, axis=1)


def spectral_fermi(x, xF: float, beta_a: float) -> np.ndarray:
    """_spectral_fermi (spectral_coefficients.jl:89-97): the step with value
    1/2 at xF when beta is infinite, else the overflow-safe Fermi function."""
    x = np.asarray(x, dtype=np.float64)
    if math.isinf(beta_a):
        return ((x < xF).astype(np.float64) + (x <= xF)) / 2
    y = beta_a * (x - xF)
    out = np.empty_like(y)
    pos = y > 0
    ey = np.exp(-y[pos])
    out[pos] = ey / (1 + ey)
    out[~pos] = 1.0 / (1 + np.exp(y[~pos]))
    return out


def deduplicate_sorted(points, tol: float = 1e-12) -> list:
    """_deduplicate_sorted (spectral_coefficients.jl:99-106)."""
    out = []
    for p in sorted(points):
        if not out or abs(p - out[-1]) > tol:
            out.append(p)
    return out


def lambda_coefficients(NC: int, xF: float) -> np.ndarray:
    """_lambda_coefficients (spectral_coefficients.jl:346-358)."""
    lam = np.zeros(NC, dtype=np.complex128)
    if xF <= -1:
        return lam
    tF = math.acos(min(max(xF, -1.0), 1.0))
    lam[0] = 1 - tF / math.pi
    n = np.arange(1, NC)
    lam[1:] = -np.sin(n * tF) / (n * math.pi)
    return lam


def _contract_error(detail: str = "") -> QuadratureError:
    suffix = f" ({detail})" if detail else ""
    return QuadratureError(
        f"spectral quadrature did not meet its error contract{suffix}; adjust quad_rtol, "
        "quad_atol, maxevals, or lambda. A nonzero quad_atol is needed for components "
        "that vanish by cancellation.")


class _Budget:
    def __init__(self, maxevals: int):
        self.maxevals = int(maxevals)
        self.used = 0

    def take(self, n: int):
        if self.used + n > self.maxevals:
            raise _contract_error("maxevals exhausted")
        self.used += n


class _Piece:
    """One quadgk call: integrand node(theta) * jac on [qa, qb] in the
    variable t, theta = t (plain), a + w t^2 (left edge) or b - w t^2
    (right edge), jac = 1 or 2 w t."""

    def __init__(self, kind: str, a: float, b: float):
        self.kind, self.a, self.b = kind, a, b
        self.width = b - a
        self.qa, self.qb = (0.0, 1.0) if kind != "plain" else (a, b)
        self.heap = []  # (-E, counter, seg)
        self.count = 0
        self.I = None
        self.E = 0.0
        self.numevals = 0
        self.done = False

    def map(self, t):
        if self.kind == "left":
            return self.a + self.width * t * t, 2.0 * self.width * t
        if self.kind == "right":
            return self.b - self.width * t * t, 2.0 * self.width * t
        return t, None


def _segment_nodes(a: float, b: float) -> np.ndarray:
    s = 0.5 * (b - a)
    return a + _OFFS * s


def _segment_rule(vals: np.ndarray, a: float, b: float, norm):
    """(I_K, E) of one segment from its 15 stored values (N_PER x ncomp)."""
    s = 0.5 * (b - a)
    pairs = vals[:14].reshape(7, 2, -1).sum(axis=1)  # f(a+(1-x)s) + f(a+(1+x)s)
    centre = vals[14]
    Ik = (pairs * _WGK[:7, None]).sum(axis=0) + centre * _WGK[7]
    Ig = (pairs[1:7:2] * _WG[:3, None]).sum(axis=0) + centre * _WG[3]
    Iks, Igs = Ik * s, Ig * s
    E = float(norm(Iks - Igs))
    if math.isnan(E) or math.isinf(E):
        raise _contract_error(f"DomainError: integrand produced {E} in the interval ({a}, {b})")
    return Iks, E


def _evaluate(segments, node_fn, ncomp: int, budget: _Budget):
    """Evaluate the 15 nodes of every (piece, a, b) in one node_fn call."""
    budget.take(NODES_PER_SEGMENT * len(segments))
    ts, thetas, jacs = [], [], []
    for piece, a, b in segments:
        t = _segment_nodes(a, b)
        th, jac = piece.map(t)
        ts.append(t)
        thetas.append(th)
        jacs.append(jac)
    vals = node_fn(np.concatenate(thetas))
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1, ncomp)
    out = []
    for k, (piece, a, b) in enumerate(segments):
        v = vals[k * NODES_PER_SEGMENT:(k + 1) * NODES_PER_SEGMENT]
        if jacs[k] is not None:
            v = v * jacs[k][:, None]
        out.append(_segment_rule(v, a, b, norm=piece.norm))
    return out


def _converged(p: _Piece) -> bool:
    return p.E <= p.atol or p.E <= p.rtol * p.norm(p.I)


def _integrate_pieces(pieces, node_fn, ncomp, rtol, atol, norm, budget):
    """Lockstep adaptive Gauss-Kronrod over independent pieces; returns the
    per-piece (I, E)."""
    for p in pieces:
        p.rtol, p.atol, p.norm = rtol, atol, norm
    live = [p for p in pieces if p.a != p.b]
    for p in pieces:
        if p.a == p.b:
            p.I, p.E, p.done = np.zeros(ncomp, dtype=np.complex128), 0.0, True
    if live:
        res = _evaluate([(p, p.qa, p.qb) for p in live], node_fn, ncomp, budget)
        for p, (I, E) in zip(live, res):
            p.I, p.E = I, E
            p.numevals = NODES_PER_SEGMENT
            heapq.heappush(p.heap, (-E, p.count, (p.qa, p.qb, I, E)))
            p.count += 1
            p.done = _converged(p)
    while True:
        active = [p for p in pieces if not p.done]
        if not active:
            break
        segs, popped = [], []
        for p in active:
            _, _, seg = heapq.heappop(p.heap)
            a, b = seg[0], seg[1]
            mid = (a + b) / 2
            segs += [(p, a, mid), (p, mid, b)]
            popped.append(seg)
        res = _evaluate(segs, node_fn, ncomp, budget)
        for k, p in enumerate(active):
            (I1, E1), (I2, E2) = res[2 * k], res[2 * k + 1]
            old = popped[k]
            p.I = (p.I - old[2]) + I1 + I2
            p.E = (p.E - old[3]) + E1 + E2
            p.numevals += 2 * NODES_PER_SEGMENT
            heapq.heappush(p.heap, (-E1, p.count, (segs[2 * k][1], segs[2 * k][2], I1, E1)))
            heapq.heappush(p.heap, (-E2, p.count + 1, (segs[2 * k + 1][1], segs[2 * k + 1][2], I2, E2)))
            p.count += 2
            if _converged(p):
                # re-sum over the segments (accumulated roundoff), as quadgk does
                order = sorted(p.heap, key=lambda h: h[1])
                p.I = order[0][2][2].copy()
                p.E = order[0][2][3]
                for h in order[1:]:
                    p.I = p.I + h[2][2]
                    p.E = p.E + h[2][3]
                p.done = _converged(p)


def spectral_integral(node_fn, ncomp: int, NC: int, shifts, xF: float, beta_a: float, lam: float, *,
                      rtol: float = 1e-8, atol: float = 0.0, maxevals: int = 10 ** 6, order: int = 7,
                      breakpoints=(), exact_zero=None):
    """_spectral_integral (spectral_coefficients.jl:135-344): integrate
    f(cos theta) F(theta) over the occupied part of (0, pi).

    Shifted real-axis Green edges split the range; with lambda == 0 they are
    regularised by the endpoint-squaring map. One pass under the Euclidean
    norm must satisfy error <= atol + rtol |I|; vector-valued integrands then
    take up to three weighted max-norm passes with scale-consistency checks.
    A shared maxevals budget covers every pass. Returns (integral, error)."""
    if not NC > 0:
        raise ArgumentError("NC must be positive")
    if not lam >= 0:
        raise ArgumentError(f"lambda must be nonnegative (got {lam})")
    if not rtol >= 0:
        raise ArgumentError("rtol must be nonnegative")
    if not atol >= 0:
        raise ArgumentError("atol must be nonnegative")
    if not maxevals > 0:
        raise ArgumentError("maxevals must be positive")
    if not order > 0:
        raise ArgumentError("order must be positive")
    if order != 7:
        raise ArgumentError("only the order-7 Gauss-Kronrod rule is provided")
    if not ncomp > 0:
        raise ArgumentError("the node function must have at least one component")
    empty = np.zeros(ncomp, dtype=np.complex128)
    if math.isinf(beta_a) and xF <= -1:
        return empty, 0.0

    theta_F = math.acos(min(max(xF, -1.0), 1.0))
    lo = theta_F if math.isinf(beta_a) else 0.0
    hi = math.pi
    points = [lo, hi]
    singular = []
    for shift in shifts:
        for edge in (-1.0, 1.0):
            xstar = edge - shift
            if -1 < xstar < 1:
                th = math.acos(xstar)
                if not (lo - 1e-12 <= th <= hi + 1e-12):
                    continue
                if abs(th - lo) <= 1e-12:
                    th = lo
                if abs(th - hi) <= 1e-12:
                    th = hi
                points.append(th)
                singular.append(th)
    if not math.isinf(beta_a) and beta_a > 50 and -1 < xF < 1:
        points.append(theta_F)
    for th in breakpoints:
        if lo < th < hi:
            points.append(float(th))
    points = deduplicate_sorted(points)
    singular = deduplicate_sorted(singular)

    def is_singular(th):
        return any(abs(th - s) <= 1e-12 for s in singular)

    piece_count = sum(2 if lam == 0 and is_singular(points[k]) and is_singular(points[k + 1]) else 1
                      for k in range(len(points) - 1))

    def fermi_node(theta):
        v = np.asarray(node_fn(theta), dtype=np.complex128).reshape(theta.size, ncomp)
        return v * spectral_fermi(np.cos(theta), xF, beta_a)[:, None]

    def make_pieces():
        groups = []
        for k in range(len(points) - 1):
            a, b = points[k], points[k + 1]
            ls, rs = is_singular(a), is_singular(b)
            if a == b:
                groups.append([_Piece("plain", a, a)])
            elif ls and rs:
                mid = (a + b) / 2
                groups.append([_sub_piece(a, mid, True, False), _sub_piece(mid, b, False, True)])
            else:
                groups.append([_sub_piece(a, b, ls, rs)])
        return groups

    def _sub_piece(a, b, ls, rs):
        if lam == 0 and ls:
            return _Piece("left", a, b)
        if lam == 0 and rs:
            return _Piece("right", a, b)
        return _Piece("plain", a, b)

    budget = _Budget(maxevals)

    def integrate_pass(p_rtol, p_atol, p_norm):
        groups = make_pieces()
        flat = [p for g in groups for p in g]
        _integrate_pieces(flat, fermi_node, ncomp, p_rtol, p_atol, p_norm, budget)
        integral = np.zeros(ncomp, dtype=np.complex128)
        total = 0.0
        for g in groups:
            I = g[0].I
            E = g[0].E
            for p in g[1:]:
                I = I + p.I
                E = E + p.E
            integral = integral + I
            total += E
        return integral, total

    piece_rtol = rtol / (10 * piece_count)
    piece_atol = atol / (10 * piece_count)
    first, first_err = integrate_pass(piece_rtol, piece_atol, lambda v: float(np.linalg.norm(v)))
    if not first_err <= atol + rtol * float(np.linalg.norm(first)):
        raise _contract_error(f"estimated error {first_err}")
    if ncomp == 1:
        return first, first_err

    exact_zero = np.zeros(ncomp, dtype=bool) if exact_zero is None else np.asarray(exact_zero, dtype=bool)
    if exact_zero.size != ncomp:
        raise ArgumentError("exact_zero mask length must equal ncomp")
    first[exact_zero] = 0
    if exact_zero.all():
        return first, first_err
    active = np.flatnonzero(~exact_zero)
    scales = atol + rtol * np.abs(first)
    if np.any(scales[active] == 0):
        raise _contract_error("quad_atol is required for components that vanish by cancellation")
    for weighted_pass in range(1, 4):
        weights = np.zeros(ncomp)
        weights[active] = 1.0 / scales[active]

        def component_norm(v, weights=weights):
            return float(np.max(np.abs(v[active]) * weights[active]))

        integral, werr = integrate_pass(0.0, 1.0 / (10 * piece_count), component_norm)
        integral[exact_zero] = 0
        if not werr <= 1:
            raise _contract_error(f"maximum scaled estimated error {werr}")
        returned = atol + rtol * np.abs(integral)
        if np.all(scales[active] <= 1.05 * returned[active]):
            return integral, werr
        if weighted_pass == 3:
            raise _contract_error("component scales remained inconsistent after 3 weighted passes")
        scales = returned
        if np.any(scales[active] == 0):
            raise _contract_error("quad_atol is required for components that vanish by cancellation")
    raise AssertionError("unreachable")
