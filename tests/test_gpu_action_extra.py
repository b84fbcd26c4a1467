"""More of the action family's reference tests on the device:

- evolution: the centre shift b only multiplies by the global phase
  e^{-i b t} (test/evolution_test.jl:41-52);
- the Fermi projector: an eigenvalue at a generic, asymmetric E_F takes the
  jump midpoint 1/2 (test/chern_marker_test.jl:137-143);
- the local Chern marker:
  - survives weak onsite disorder, against ED and the FHS Chern number
    (:292-307);
  - the stochastic regional estimator is unbiased at 512 probes on a small
    flake, its batching is exact bookkeeping, and it rejects bad regions,
    NR = 0 and a missing generator (:309-426);
  - a rigid spectral shift with E_F moved alongside leaves marker and
    projector unchanged (:428-446).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def marker_fixture(H, margin=0.95):
    ev = np.linalg.eigvalsh(H.toarray())
    b = (ev.max() + ev.min()) / 2
    a = (ev.max() - ev.min()) / 2 / margin
    return K.RescaledHamiltonian(((H - b * sp.identity(H.shape[0])) / a).tocsc(), a, b)


def cells(Ly, xs, ys):
    return [s for x in xs for y in ys
            for s in (2 * ((y - 1) + Ly * (x - 1)) + 1, 2 * ((y - 1) + Ly * (x - 1)) + 2)]


def test_evolution_center_shift_is_a_global_phase():
    N = 64
    h = marker_fixture(sp.csc_matrix(models.ring_model(N) + 0.7 * sp.identity(N)))
    h0 = K.RescaledHamiltonian(h.H, h.a, 0.0)
    rng = np.random.default_rng(193)
    psi = rng.normal(size=N) + 1j * rng.normal(size=N)
    psi /= np.linalg.norm(psi)
    for t in (1.3, -2.1):
        np.testing.assert_allclose(K.evolve(h, psi, t),
                                   np.exp(-1j * h.b * t) * K.evolve(h0, psi, t), atol=1e-13)


def test_projector_midpoint_at_generic_fermi_level():
    hg = K.RescaledHamiltonian(sp.csc_matrix(np.diag([-0.5, 0.3, 0.6]).astype(complex)), 1.0, 0.0)
    p = K.fermi_projector(hg, np.array([0, 1, 0], dtype=complex), Ef=0.3, NC=4096)
    assert abs(p[1].real - 0.5) <= 1e-2
    assert K.fermiFunctions(0.3, math.inf)(0.3) == 0.5


def test_marker_survives_weak_disorder():
    Lx = Ly = 10
    NC, W = 512, 0.5
    H, pos, Ac = models.haldane_open_model(Lx, Ly)
    r = Xoshiro256pp(227)
    Hd = (H + sp.diags(W * (r.rand(2 * Lx * Ly) - 0.5))).tocsc()
    h = marker_fixture(Hd)
    bulk = cells(Ly, range(4, 8), range(4, 8))
    mk = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=bulk, NC=NC)
    assert abs(K.chern_marker_average(mk, area=16 * Ac) - 1.0) <= 0.15
    mk_ed = oracle.ed_chern_marker(Hd, pos[:, 0], pos[:, 1], 0.0)
    np.testing.assert_allclose(mk, mk_ed[np.array(bulk) - 1], atol=0.05)


def test_regional_estimator_bias_batching_and_validation():
    NC, NR = 512, 32
    H, pos, Ac = models.haldane_open_model(10, 10)
    h = marker_fixture(H)
    region = cells(10, range(4, 8), range(4, 8))
    est = K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=region,
                                rng=Xoshiro256pp(31), NR=NR, NC=NC)
    est5 = K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=region,
                                 rng=Xoshiro256pp(31), NR=NR, NC=NC, batch_size=5)
    np.testing.assert_allclose(est5, est, rtol=1e-12, atol=1e-14)
    # 512 probes over the full basis of a 4 x 4 flake against the exact sum
    H4, pos4, _ = models.haldane_open_model(4, 4)
    h4 = marker_fixture(H4)
    det4 = K.chern_marker(h4, pos4[:, 0], pos4[:, 1], Ef=0.0, sites=list(range(1, 33)), NC=128).sum()
    est4 = K.chern_marker_region(h4, pos4[:, 0], pos4[:, 1], Ef=0.0, region=list(range(1, 33)),
                                 rng=Xoshiro256pp(41), NR=512, NC=128, batch_size=64)
    assert abs(est4.mean() - det4) < 4 * est4.std(ddof=1) / math.sqrt(512)
    with pytest.raises(K.ArgumentError):
        K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=[1, 1], rng=Xoshiro256pp(1), NC=NC)
    with pytest.raises(K.ArgumentError):
        K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=region, rng=Xoshiro256pp(1),
                              NR=0, NC=NC)
    with pytest.raises(TypeError):
        K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=region, NR=2, NC=NC)


def test_center_shift_invariance_of_marker_and_projector():
    Lx = Ly = 6
    NC, delta = 256, 3.7
    H, pos, Ac = models.haldane_open_model(Lx, Ly)
    h = marker_fixture(H)
    hs = K.RescaledHamiltonian(h.H, h.a, h.b + delta)
    sites = cells(Ly, range(3, 5), range(3, 5))
    mk0 = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=sites, NC=NC)
    mkd = K.chern_marker(hs, pos[:, 0], pos[:, 1], Ef=delta, sites=sites, NC=NC)
    np.testing.assert_allclose(mkd, mk0, rtol=0, atol=1e-12)
    v = (pos[:, 0] / np.linalg.norm(pos[:, 0])).astype(complex)
    np.testing.assert_allclose(K.fermi_projector(hs, v, Ef=delta, NC=NC),
                               K.fermi_projector(h, v, Ef=0.0, NC=NC), rtol=0, atol=1e-12)
