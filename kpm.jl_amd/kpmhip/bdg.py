"""Bogoliubov-de Gennes operators on the device engine
(src/applications/bdg.jl:1-412,575-660,810-856; ext/KPMCUDAExt.jl:47-62).

The reference's GPU path runs BdG on the assembled sparse Nambu matrix
(bdg_assemble): the extension materialises the 2N x 2N matrix, which costs
O(nnz), and moves it as one CSR operator. Operators that cannot be assembled
stay on the host there. This engine has no host path, so the
device-resident form is the only one:
- BdGOperator and ScaledOperator assemble to a DeviceCSR through
  to_device;
- an operator with non-matrix blocks raises ArgumentError, the reference's
  _device_operator guard (bdg.jl:390-411);
- the power iteration (spectral_radius) and the stability probe run their
  SpMVs and norms on the device;
- the Gershgorin bound is O(nnz) host arithmetic on the assembled blocks,
  as in the reference.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, KPMUnstableError
from .device import DeviceBlock, DeviceCSR, get_context

_SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


def _is_matrix(x) -> bool:
    return sp.issparse(x) or isinstance(x, np.ndarray)


def _canonical_hole_convention(s: str, caller: str = "BdGOperator") -> str:
    """_canonical_hole_convention (bdg.jl:183-191): 'singlet' aliases 'conjugate'."""
    s = str(s).lstrip(":")
    if s == "singlet":
        return "conjugate"
    if s in ("conjugate", "intervalley"):
        return s
    raise ArgumentError(f"{caller}: hole_convention must be :intervalley or :conjugate "
                        f"(:singlet is an alias) (got {s})")


class ScaledOperator:
    """(op - b I) / a without forming it on the host
    (bdg.jl:8-35). Its device form is the CSR of (op - b I)/a."""

    def __init__(self, op, a: float, b: float = 0.0):
        self.op = op
        self.a = float(a)
        self.b = float(b)
        self._dev = None

    @property
    def shape(self):
        return self.op.shape

    def host_matrix(self):
        """The assembled (op - b I)/a, CSC complex."""
        A = self.op.host_matrix() if hasattr(self.op, "host_matrix") else self.op
        if not _is_matrix(A):
            raise ArgumentError("ScaledOperator: the wrapped operator has no assembled matrix form")
        A = sp.csc_matrix(A, dtype=np.complex128)
        if self.b != 0.0:
            A = A - self.b * sp.identity(A.shape[0], dtype=np.complex128, format="csc")
        return sp.csc_matrix(A / self.a)

    def device_operator(self, ctx=None) -> DeviceCSR:
        """maybe_to_device(::ScaledOperator) (ext/KPMCUDAExt.jl:58-59): the
        device CSR of (op - b I)/a, built once and kept."""
        if self._dev is None:
            self._dev = DeviceCSR(self.host_matrix(), ctx)
        return self._dev


class BdGOperator:
    """Nambu BdG operator [xi D; D^H hole] with xi = h - mu - U n / 2
    (bdg.jl:156-304). hole = -xi for hole_convention 'intervalley' (same h,
    the default), -conj(xi) for 'conjugate' (alias 'singlet'). Blocks are
    matrices (scipy sparse or dense); ``Delta`` is the onsite shorthand for
    D = Diagonal(Delta)."""

    def __init__(self, h, *, mu: float, U, n=None, Delta=None, D=None,
                 hole_convention: str = "intervalley", h_hole=None, assume_intervalley: bool = False):
        N = h.shape[0]
        if h.shape[1] != N:
            raise ArgumentError(f"BdGOperator: h must be square (got {h.shape})")
        hole_convention = _canonical_hole_convention(hole_convention)
        U_vec = np.full(N, float(U)) if np.isscalar(U) else np.asarray(U, dtype=np.float64).copy()
        n_vec = np.zeros(N) if n is None else np.asarray(n, dtype=np.float64).copy()
        if U_vec.shape != (N,):
            raise ArgumentError(f"BdGOperator: U has length {U_vec.size}; expected {N}")
        if n_vec.shape != (N,):
            raise ArgumentError(f"BdGOperator: n has length {n_vec.size}; expected {N}")
        if Delta is not None and D is not None:
            raise ArgumentError("BdGOperator: supply only one of Delta or D")
        self._onsite = None
        if Delta is not None:
            d = np.asarray(Delta)
            if d.ndim != 1:
                raise ArgumentError("BdGOperator: Delta must be a vector")
            d = d.astype(np.complex128)
            if d.size != N:
                raise ArgumentError(f"BdGOperator: Delta has length {d.size}; expected {N}")
            self._onsite = d
            D_stored = None
        elif D is not None:
            if D.shape != (N, N):
                raise ArgumentError(f"BdGOperator: D has size {D.shape}; expected ({N}, {N})")
            D_stored = D
        else:
            self._onsite = np.zeros(N, dtype=np.complex128)
            D_stored = None
        h_is_matrix = _is_matrix(h)
        if h_is_matrix:  # ishermitian: exact equality with the adjoint
            if sp.issparse(h):
                Hm = sp.csc_matrix(h)
                herm = (Hm != Hm.conj().T).nnz == 0
            else:
                herm = np.array_equal(np.asarray(h), np.asarray(h).conj().T)
            if not herm:
                raise ArgumentError("BdGOperator: matrix h must be Hermitian")
        if hole_convention == "intervalley":
            if h_hole is not None:
                raise ArgumentError("BdGOperator: h_hole must be nothing for hole_convention=:intervalley")
            if h_is_matrix and np.iscomplexobj(h.data if sp.issparse(h) else h):
                vals = h.data if sp.issparse(h) else np.asarray(h)
                max_imag = float(np.abs(vals.imag).max()) if vals.size else 0.0
                max_abs = float(np.abs(vals).max()) if vals.size else 0.0
                symmetric = max_imag <= _SQRT_EPS * max(max_abs, 1e-300)
                if not symmetric and not assume_intervalley:
                    raise ArgumentError(
                        "BdGOperator: for complex Hermitian h the reduced same-h hole block is only "
                        "the correct physics if h is the intervalley-identified operator "
                        "(h_{-K}^* = h_K); pass assume_intervalley=true to assert that, or use a "
                        "real-symmetric h.")
            hh = h
        else:
            if assume_intervalley:
                raise ArgumentError("BdGOperator: assume_intervalley is meaningless for "
                                    "hole_convention=:conjugate; leave it false")
            if h_is_matrix and h_hole is not None:
                bad = (not _is_matrix(h_hole)) or h_hole.shape != h.shape or \
                    abs(sp.csc_matrix(h_hole) - sp.csc_matrix(h).conj()).max() != 0
                if bad:
                    raise ArgumentError(
                        "BdGOperator: for an assembled h the conjugate hole block is conj(h); a "
                        "different h_hole would silently invalidate gershgorin_bound and the "
                        "Hermitian BdG action")
            if h_hole is None:
                if not h_is_matrix:
                    raise ArgumentError("BdGOperator: matrix-free h with hole_convention=:conjugate "
                                        "requires h_hole, the conjugated normal-state operator")
                hh = h.conj()
            else:
                if h_hole.shape != h.shape:
                    raise ArgumentError(f"BdGOperator: h_hole has size {h_hole.shape}; expected {h.shape}")
                hh = h_hole
        self.h, self.h_hole = h, hh
        self.mu = float(mu)
        self.U, self.n = U_vec, n_vec
        self._D = D_stored
        self.N = N
        self.hole_convention = hole_convention

    @property
    def D(self):
        """The pairing block: Diagonal(Delta) as a sparse diagonal for onsite
        pairing, else the stored matrix."""
        if self._onsite is not None:
            return sp.diags(self._onsite, format="csc")
        return self._D

    @property
    def Delta(self):
        """op.Δ (bdg.jl:167-178): only for onsite (diagonal) pairing."""
        if self._onsite is None:
            raise ArgumentError("BdGOperator: Δ is only defined for onsite (Diagonal) pairing; access op.D")
        return self._onsite

    @property
    def shape(self):
        return (2 * self.N, 2 * self.N)

    def assemblable(self) -> bool:
        """_bdg_assemblable (bdg.jl:341-345)."""
        return _is_matrix(self.h) and _is_matrix(self.h_hole) and (
            self._onsite is not None or _is_matrix(self._D))

    def host_matrix(self):
        return bdg_assemble(self)

    def device_operator(self, ctx=None) -> DeviceCSR:
        """to_device(::BdGOperator) (ext/KPMCUDAExt.jl:47-56) with the
        _device_operator guard (bdg.jl:390-411)."""
        if not (self.assemblable() and sp.issparse(self.h)):
            raise ArgumentError(
                "the active GPU device requires a device-assemblable operator (sparse matrix h and "
                "matrix D blocks); matrix-free or dense blocks would mix host and device residences "
                "here -- sparsify/assemble the blocks or deactivate the GPU")
        return DeviceCSR(bdg_assemble(self), ctx)


def bdg_assemble(op: BdGOperator):
    """bdg_assemble (bdg.jl:347-388): [h - (mu + U n/2) I, D;
    D^H, -(h_hole - (mu + U n/2) I)] as CSC complex."""
    if not op.assemblable():
        raise ArgumentError("bdg_assemble: requires assembled matrix blocks h, h_hole, and D "
                            "(matrix-free blocks cannot be assembled)")
    onsite = sp.diags((op.mu + (op.U / 2) * op.n).astype(np.complex128), format="csc")
    xi = sp.csc_matrix(op.h, dtype=np.complex128) - onsite
    hole = -(sp.csc_matrix(op.h_hole, dtype=np.complex128) - onsite)
    D = sp.csc_matrix(op.D, dtype=np.complex128)
    return sp.csc_matrix(sp.bmat([[xi, D], [D.conj().T, hole]], format="csc"))


def gershgorin_bound(op: BdGOperator) -> float:
    """gershgorin_bound (bdg.jl:581-660): max_i of the off-diagonal |h| row
    sum plus |h_ii - mu - U_i n_i / 2| plus the pairing row weight (|Delta_i|
    onsite; max of the |D| row and column sums for a matrix D)."""
    if not _is_matrix(op.h):
        raise ArgumentError("gershgorin_bound: BdGOperator must have an assembled matrix h")
    N = op.N
    if N == 0:
        return 0.0
    Hc = sp.coo_matrix(op.h)
    off = Hc.row != Hc.col
    rowsums = np.bincount(Hc.row[off], weights=np.abs(Hc.data[off]), minlength=N).astype(np.float64)
    diagonal = np.zeros(N, dtype=np.complex128)
    diagonal[Hc.row[~off]] = Hc.data[~off]
    if op._onsite is not None:
        pairing = np.abs(op._onsite)
    elif _is_matrix(op._D):
        Dc = sp.coo_matrix(op._D)
        av = np.abs(Dc.data)
        rs = np.bincount(Dc.row, weights=av, minlength=N)
        cs = np.bincount(Dc.col, weights=av, minlength=N)
        pairing = np.maximum(rs, cs)
    else:
        raise ArgumentError("gershgorin_bound: D must be an AbstractMatrix")
    rowsums += np.abs(diagonal - op.mu - op.U * op.n / 2) + pairing
    return float(rowsums.max())


def _randn_complex(rng, n: int) -> np.ndarray:
    """Standard complex normal values (variance 1/2 per component) from the
    generator's uniform stream by the Box-Muller transform."""
    u = rng.rand(n, 2)
    r = np.sqrt(-np.log1p(-u[:, 0]))  # sqrt(-2 ln(1-u1)) / sqrt(2)
    return r * np.exp(2j * np.pi * u[:, 1])


def _device_op(op, ctx):
    if isinstance(op, DeviceCSR):
        return op
    if hasattr(op, "device_operator"):
        return op.device_operator(ctx)
    return DeviceCSR(op, ctx)


def spectral_radius(op, *, tol: float = 1e-4, maxiter: int = 300, miniter: int = 32,
                    restarts: int = 2, rng=None, v0=None):
    """spectral_radius (bdg.jl:37-103): norm-ratio power iteration, one
    primary run (from v0 when given) and ``restarts`` fresh random runs; the
    largest estimate and its unit vector (host) are returned. Each step is
    one device SpMV and one device norm. The random starts come from the
    build's generator (default seed 0), not Julia's Xoshiro."""
    from .rng import Xoshiro256pp

    if not maxiter > 0:
        raise ArgumentError(f"spectral_radius: maxiter must be positive (got {maxiter})")
    if not (1 <= miniter <= maxiter):
        raise ArgumentError(f"spectral_radius: miniter must satisfy 1 <= miniter <= maxiter "
                            f"(got miniter={miniter}, maxiter={maxiter})")
    if not restarts >= 0:
        raise ArgumentError(f"spectral_radius: restarts must be nonnegative (got {restarts})")
    rng = Xoshiro256pp(0) if rng is None else rng
    ctx = get_context()
    A = _device_op(op, ctx)
    N = A.shape[0]
    blocks = [DeviceBlock(N, 1, "f64", ctx), DeviceBlock(N, 1, "f64", ctx)]
    tgt = np.zeros((1, 1), dtype=np.complex128, order="F")

    from .primitives import columnwise_dot, spmm

    def norm_of(b):
        columnwise_dot(tgt, 1, b, b)
        return math.sqrt(max(tgt[0, 0].real, 0.0))

    def power_run(v_start):
        v = np.asarray(v_start, dtype=np.complex128).reshape(-1)
        if v.size != N:
            raise ArgumentError(f"spectral_radius: v0 has length {v.size}; expected {N}")
        nv = np.linalg.norm(v)
        if nv == 0:
            raise ArgumentError("spectral_radius: v0 must be nonzero")
        blocks[0].upload((v / nv).reshape(N, 1))
        cur, nxt = 0, 1
        scale = 1.0  # blocks[cur] holds lam_prev * v
        lam_prev = math.inf
        for it in range(1, maxiter + 1):
            spmm(A, blocks[cur], blocks[nxt], alpha=scale, beta=0.0)  # w = A v
            lam = norm_of(blocks[nxt])
            if lam == 0:
                return 0.0, blocks[cur].download()[:, 0] * scale
            cur, nxt = nxt, cur
            scale = 1.0 / lam
            if it >= miniter and abs(lam - lam_prev) <= tol * max(lam, np.finfo(float).eps):
                return float(lam), blocks[cur].download()[:, 0] / lam
            lam_prev = lam
        return float(lam_prev), blocks[cur].download()[:, 0] * scale

    primary = _randn_complex(rng, N) if v0 is None else v0
    best, vbest = power_run(primary)
    for _ in range(restarts):
        r, v = power_run(_randn_complex(rng, N))
        if r > best:
            best, vbest = r, v
    return best, vbest


def rescale_bdg(op: BdGOperator, *, eps: float = 0.2, bound: str = "power", radius=None, **kwargs):
    """rescale(::BdGOperator) (src/frontend.jl:163-203): b = 0 and
    a = 2 radius / (2 - eps), with radius from the power iteration
    ("power"), the certified Gershgorin bound ("gershgorin"), or given."""
    from .frontend import RescaledHamiltonian

    if not (0 < eps < 2):
        raise ArgumentError(f"rescale: eps must satisfy 0 < eps < 2 (got {eps})")
    bound = str(bound).lstrip(":")
    if bound not in ("power", "gershgorin"):
        raise ArgumentError(f"rescale: bound must be :power or :gershgorin (got {bound})")
    if radius is not None:
        rad = float(radius)
    elif bound == "power":
        rad = spectral_radius(op, **kwargs)[0]
    else:
        if kwargs:
            raise ArgumentError("rescale: power-iteration keyword arguments are not used with "
                                "bound=:gershgorin")
        rad = gershgorin_bound(op)
    a = 2 * rad / (2 - eps)
    if not (math.isfinite(a) and a > 0):
        raise ArgumentError(f"rescale: BdG scale a must be finite and positive (got {a}); the zero "
                            "operator cannot be rescaled")
    return RescaledHamiltonian(ScaledOperator(op, a, 0.0), a, 0.0)


def chebyshev_stability_probe(Hs, NH: int, NC: int, *, rng=None) -> float:
    """chebyshev_stability_probe (bdg.jl:806-856): the largest norm of
    T_n(Hs) v, n < NC, for one random unit vector v (random from the build's
    generator, default seed 1).

    The norms come from one device moment pass: with the doubling identity
    <T_n v|T_n v> = (mu_2n + mu_0)/2, the 2 NC moments of kpm_1d on v give
    every norm. A recurrence that the engine's own guard stops counts as
    unstable (infinite norm)."""
    from .moment import kpm_1d
    from .rng import Xoshiro256pp

    if not NH > 0:
        raise ArgumentError(f"chebyshev_stability_probe: NH must be positive (got {NH})")
    if tuple(Hs.shape) != (NH, NH):
        raise ArgumentError(f"chebyshev_stability_probe: Hs has size {tuple(Hs.shape)}; expected "
                            f"({NH}, {NH})")
    if not NC > 0:
        raise ArgumentError(f"chebyshev_stability_probe: NC must be positive (got {NC})")
    rng = Xoshiro256pp(1) if rng is None else rng
    v = _randn_complex(rng, NH)
    v = (v / np.linalg.norm(v)).reshape(NH, 1)
    if NC == 1:
        return 1.0
    A = _device_op(Hs, get_context())
    try:
        mu = np.asarray(kpm_1d(A, 2 * NC, 1, NH, psi_in=np.asfortranarray(v)), dtype=np.float64)
    except KPMUnstableError:
        return math.inf
    mu = mu.reshape(-1)
    sq = (mu[0:2 * NC:2] + mu[0]) / 2  # <T_n v|T_n v>, n = 0 .. NC-1
    if not np.all(np.isfinite(sq)):
        return math.inf
    return float(max(1.0, np.sqrt(np.maximum(sq, 0.0)).max()))
