"""Typed front end (src/frontend.jl:1-356, 553-644, 984-1035): value types
carrying the rescaling metadata and the typed wrappers that delegate to the
raw engine unchanged."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ._lib import ArgumentError
from .applications import dos as _dos_raw
from .applications import dos0 as _dos0_raw
from .applications import kubo_bastin_cond as _kb_raw
from .moment import kpm_1d, kpm_1d_, kpm_1d_current, kpm_2d
from .primitives import normalize_by_col
from .rescale import normalizeH


@dataclass(frozen=True)
class RescaledHamiltonian:
    """H = (H_original - b I)/a (src/frontend.jl:11-15)."""
    H: object
    a: float
    b: float

    @property
    def shape(self):
        return self.H.shape


class AbstractMoments:
    pass


class DosMoments(AbstractMoments):
    """src/frontend.jl:32-44"""

    def __init__(self, mu, a, b, NH, NR):
        if not (NH > 0 and NR > 0):
            raise ArgumentError(f"DosMoments: NH and NR must be positive (got NH={NH}, NR={NR})")
        self.mu, self.a, self.b, self.NH, self.NR = np.asarray(mu), float(a), float(b), int(NH), int(NR)

    def __repr__(self):
        return f"DosMoments(NC={nc(self)}, NR={self.NR}, NH={self.NH}, a={self.a}, b={self.b})"


class ConductivityMoments(AbstractMoments):
    """src/frontend.jl:53-72"""

    def __init__(self, mu, a, b, NH, NR):
        mu = np.asarray(mu)
        if mu.shape[0] != mu.shape[1]:
            raise ArgumentError(
                f"ConductivityMoments: the moment matrix must be square (got {mu.shape})")
        if not (NH > 0 and NR > 0):
            raise ArgumentError(
                f"ConductivityMoments: NH and NR must be positive (got NH={NH}, NR={NR})")
        self.mu, self.a, self.b, self.NH, self.NR = mu, float(a), float(b), int(NH), int(NR)

    def __repr__(self):
        return (f"ConductivityMoments(NC={nc(self)}, NR={self.NR}, NH={self.NH}, a={self.a}, "
                f"b={self.b})")


class CurrentMoments(AbstractMoments):
    """src/frontend.jl:81-87"""

    def __init__(self, mu, a, b, NH, NR):
        self.mu, self.a, self.b, self.NH, self.NR = np.asarray(mu), float(a), float(b), int(NH), int(NR)


class GreenMoments(AbstractMoments):
    """src/frontend.jl:99-116: mu[n, p] = <u_p|T_{n-1}|v_p>, NC x npairs."""

    def __init__(self, mu, a, b, NH):
        mu = np.asarray(mu)
        if not NH > 0:
            raise ArgumentError(f"GreenMoments: NH must be positive (got NH={NH})")
        if mu.shape[0] < 1:
            raise ArgumentError(
                f"GreenMoments: mu must contain at least one moment (got size {mu.shape})")
        if not (np.isfinite(a) and a > 0):
            raise ArgumentError(f"GreenMoments: a must be finite and positive (got a={a})")
        if not np.isfinite(b):
            raise ArgumentError(f"GreenMoments: b must be finite (got b={b})")
        self.mu, self.a, self.b, self.NH = mu, float(a), float(b), int(NH)


def nc(m: AbstractMoments) -> int:
    return int(m.mu.shape[0])


def npairs(m: GreenMoments) -> int:
    return int(m.mu.shape[1])


def rescale(H, *, center: bool = False, eps: float | None = None, fixed_a: float = 0.0,
            method: str = "lanczos", **bdg_kwargs):
    """src/frontend.jl:149-161 (``method``: the normalizeH edge estimator).
    A BdGOperator takes the radial BdG rescaling instead
    (src/frontend.jl:163-203: eps default 0.2, ``bound``, ``radius`` and the
    power-iteration keywords)."""
    from .bdg import BdGOperator, rescale_bdg

    if isinstance(H, BdGOperator):
        return rescale_bdg(H, eps=0.2 if eps is None else eps, **bdg_kwargs)
    if bdg_kwargs:
        raise ArgumentError(f"rescale: unexpected keywords {sorted(bdg_kwargs)}")
    eps = 0.1 if eps is None else eps
    if center:
        a, b, Hn = normalizeH(H, eps=eps, fixed_a=fixed_a, center=True, method=method)
    else:
        a, Hn = normalizeH(H, eps=eps, fixed_a=fixed_a, method=method)
        b = 0.0
    af, bf = float(a), float(b)
    if not (np.isfinite(af) and np.isfinite(bf)):
        raise ArgumentError(f"rescale: a and b must be finite as Float64 (got a={a}, b={b})")
    return RescaledHamiltonian(Hn, af, bf)


def random_phase_vectors(rng, NH: int, NR: int) -> np.ndarray:
    """exp(2πi U) with unit-norm columns and no centering
    (src/frontend.jl:212-216); ``rng`` is a kpmhip.rng generator."""
    psi = np.exp(rng.rand(NH, NR) * (2j * np.pi))
    normalize_by_col(psi, NR)
    return np.asfortranarray(psi)


def dos_moments(h: RescaledHamiltonian, *, NC: int = 1024, NR: int = 12, rng=None, psi_in=None,
                verbose: int = 0, precision="f64") -> DosMoments:
    """src/frontend.jl:225-259"""
    if rng is not None and psi_in is not None:
        raise ArgumentError("pass either rng or psi_in, not both")
    if NC % 2 != 0:
        raise ArgumentError(
            "kpm_1d computes NC moments from NC/2 recurrence steps; NC must be even")
    NH = h.H.shape[0]
    NR_int = int(NR)
    if psi_in is not None:
        if psi_in.shape[0] != NH:
            raise ArgumentError(f"psi_in has {psi_in.shape[0]} rows; expected {NH}")
        NR_int = psi_in.shape[1]
    elif rng is not None:
        psi_in = random_phase_vectors(rng, NH, NR_int)
    if psi_in is None:
        mu = kpm_1d(h.H, int(NC), NR_int, verbose=verbose, precision=precision)
    else:
        mu = kpm_1d(h.H, int(NC), NR_int, psi_in=psi_in, verbose=verbose, precision=precision)
    return DosMoments(mu, h.a, h.b, NH, NR_int)


def cond_moments(h: RescaledHamiltonian, Ja, Jb, *, NC: int = 256, NR: int = 8, rng=None,
                 psi_in=None, arr_size="auto", right_block="auto", workspace_bytes="auto",
                 **kwargs) -> ConductivityMoments:
    """src/frontend.jl:270-327"""
    if rng is not None and psi_in is not None:
        raise ArgumentError("pass either rng or psi_in, not both")
    NH = h.H.shape[0]
    NR_int = int(NR)
    if psi_in is not None:
        if psi_in.shape[0] != NH:
            raise ArgumentError(f"psi_in has {psi_in.shape[0]} rows; expected {NH}")
        NR_int = psi_in.shape[1]
    elif rng is not None:
        psi_in = random_phase_vectors(rng, NH, NR_int)
    mu = kpm_2d(h.H, Ja, Jb, int(NC), NR_int, NH, psi_in=psi_in, arr_size=arr_size,
                right_block=right_block, workspace_bytes=workspace_bytes, **kwargs)
    return ConductivityMoments(mu, h.a, h.b, NH, NR_int)


def current_moments(h: RescaledHamiltonian, J, NC: int, NR: int, *, rng=None,
                    **kwargs) -> CurrentMoments:
    """src/frontend.jl:338-356"""
    NH = h.H.shape[0]
    if rng is None:
        mu = kpm_1d_current(h.H, J, int(NC), int(NR), NH, **kwargs)
    else:
        psi = random_phase_vectors(rng, NH, int(NR))
        mu = kpm_1d_current(h.H, J, int(NC), int(NR), NH, psi_in=psi, **kwargs)
    return CurrentMoments(mu, h.a, h.b, NH, int(NR))


def green_moments(h: RescaledHamiltonian, psi_l, psi_r=None, *, NC: int = 1024, verbose: int = 0,
                  precision="f64") -> GreenMoments:
    """green_moments(h, psi_l, psi_r) (src/frontend.jl:553-574) -- bra/ket
    recurrence -- or green_moments(h, psi) (:584-603) -- doubling."""
    if NC < 2:
        raise ArgumentError(f"NC must be at least 2 (got {NC})")
    NH = h.H.shape[0]
    if psi_r is None:
        if NC % 2 != 0:
            raise ArgumentError(
                "kpm_1d computes NC moments from NC/2 recurrence steps; NC must be even")
        psi = np.asarray(psi_l)
        psi = psi.reshape(-1, 1) if psi.ndim == 1 else psi
        if psi.shape[0] != NH:
            raise ArgumentError(f"psi has {psi.shape[0]} rows; expected {NH}")
        M = psi.shape[1]
        mu_all = np.zeros((M, int(NC)), dtype=np.complex128, order="F")
        kpm_1d_(h.H, int(NC), M, NH, mu_all, psi, precision=precision)
        return GreenMoments(mu_all.T.copy(), h.a, h.b, NH)
    pl = np.asarray(psi_l)
    pr = np.asarray(psi_r)
    pl = pl.reshape(-1, 1) if pl.ndim == 1 else pl
    pr = pr.reshape(-1, 1) if pr.ndim == 1 else pr
    if pl.shape[0] != NH:
        raise ArgumentError(f"psi_l has {pl.shape[0]} rows; expected {NH}")
    if pl.shape != pr.shape:
        raise ArgumentError(
            f"psi_l and psi_r must have equal size (got {pl.shape} and {pr.shape})")
    M = pl.shape[1]
    mu_all = np.zeros((M, int(NC)), dtype=np.complex128, order="F")
    kpm_1d_(h.H, int(NC), M, NH, mu_all, pl, pr, precision=precision)
    return GreenMoments(mu_all.T.copy(), h.a, h.b, NH)


def ldos_moments(h: RescaledHamiltonian, *, sites, NC: int = 1024, batch_size: int = 64,
                 verbose: int = 0, precision="f64") -> GreenMoments:
    """Site-diagonal moments <i|T_n|i> in batches of unit probes
    (src/frontend.jl:614-644); `sites` are 1-based like the reference."""
    if NC < 2:
        raise ArgumentError(f"NC must be at least 2 (got {NC})")
    if NC % 2 != 0:
        raise ArgumentError(
            "kpm_1d computes NC moments from NC/2 recurrence steps; NC must be even")
    if batch_size <= 0:
        raise ArgumentError(f"batch_size must be positive (got {batch_size})")
    NH = h.H.shape[0]
    sites = np.asarray(sites, dtype=np.int64)
    if not np.all((sites >= 1) & (sites <= NH)):
        raise ArgumentError(f"site indices must lie in 1:{NH}")
    mu = np.zeros((int(NC), len(sites)), dtype=np.complex128)
    for lo in range(0, len(sites), int(batch_size)):
        hi = min(lo + int(batch_size), len(sites))
        nb = hi - lo
        psi = np.zeros((NH, nb), dtype=np.complex128, order="F")
        for c, s in enumerate(sites[lo:hi]):
            psi[s - 1, c] = 1
        mb = np.zeros((nb, int(NC)), dtype=np.complex128, order="F")
        kpm_1d_(h.H, int(NC), nb, NH, mb, psi, precision=precision)
        mu[:, lo:hi] = mb.T
    return GreenMoments(mu, h.a, h.b, NH)


def dos(m, **kwargs):
    """dos(m::DosMoments) (src/frontend.jl:989-993) or
    dos(h::RescaledHamiltonian; NC, NR, rng, psi_in, ...) (:1001-1012)."""
    if isinstance(m, RescaledHamiltonian):
        keys = ("NC", "NR", "rng", "psi_in", "verbose", "precision")
        mk = {k: kwargs.pop(k) for k in keys if k in kwargs}
        m = dos_moments(m, **mk)
    if "b" in kwargs:
        raise ArgumentError("b is stored in the moments object; do not pass it separately")
    return _dos_raw(m.mu, m.a, b=m.b, **kwargs)


def dos0(m: DosMoments, **kwargs):
    """src/frontend.jl:1020"""
    return _dos0_raw(m.mu, m.a, **kwargs)


def kubo_bastin_cond(m: ConductivityMoments, Ef: float, *, area: float, **kwargs) -> float:
    """src/frontend.jl:1028-1035"""
    if "b" in kwargs:
        raise ArgumentError("b is stored in the moments object; do not pass it separately")
    if "NH" in kwargs:
        raise ArgumentError("NH is stored in the moments object; do not pass it separately")
    return _kb_raw(m.mu, m.a, Ef, b=m.b, NH=m.NH, area=area, **kwargs)
