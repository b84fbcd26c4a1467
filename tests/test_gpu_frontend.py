"""The typed front end against the raw API it delegates to, on the device
(test/frontend_test.jl:15-202 on the open 64-site chain with Jx, Jy = 0.7 Jx).

Delegation is asserted bitwise: the engine is deterministic run to run, so
the typed wrapper and the raw call produce identical tables. The other
checks are:
- rescale provenance;
- the DOS sum rules with identity probes (Gauss-Chebyshev rule, exact);
- the trapezoid window deficit;
- grid / range equivalence;
- the validation errors.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu

NH, NR = 64, 4
H = sp.diags([np.ones(NH - 1), np.ones(NH - 1)], [-1, 1], format="csc").astype(np.complex128)
JX = sp.diags([-1j * np.ones(NH - 1), 1j * np.ones(NH - 1)], [-1, 1], format="csc")
JY = 0.7 * JX


@pytest.fixture(scope="module")
def h():
    return K.rescale(H)


def manual_probes(seed):
    psi = np.exp(Xoshiro256pp(seed).rand(NH, NR) * (2j * np.pi))
    K.normalize_by_col(psi, NR)
    return np.asfortranarray(psi)


def test_rescale_provenance(h):
    np.testing.assert_allclose(h.a * h.H.toarray() + h.b * np.eye(NH), H.toarray(), atol=1e-12)
    ev = np.linalg.eigvalsh(h.H.toarray())
    assert np.all((ev > -1) & (ev < 1))
    Hs = H + 0.7 * sp.identity(NH)
    hc = K.rescale(Hs, center=True)
    np.testing.assert_allclose(hc.a * hc.H.toarray() + hc.b * np.eye(NH), Hs.toarray(), atol=1e-12)
    evc = np.linalg.eigvalsh(hc.H.toarray())
    assert np.all((evc > -1) & (evc < 1))
    hf = K.rescale(H, fixed_a=3)
    assert hf.a == 3.0 and isinstance(hf.a, float) and hf.b == 0.0


def test_typed_dos_delegation_and_grids(h):
    psi = manual_probes(11)
    m = K.dos_moments(h, NC=64, psi_in=psi.copy())
    ref = K.kpm_1d(h.H, 64, NR, psi_in=psi.copy())
    assert np.array_equal(m.mu, ref)
    for a, b in zip(K.dos(m), K.dos_raw(ref, h.a, b=h.b)):
        assert np.array_equal(a, b)
    assert np.array_equal(np.asarray(K.dos0(m)), np.asarray(K.dos0_raw(ref, h.a)))
    assert np.array_equal(np.asarray(K.dos0(m, dE_order=2)), np.asarray(K.dos0_raw(ref, h.a, dE_order=2)))
    grid = np.linspace(-1.5, 1.5, 101)
    E, rho = K.dos(m, E_grid=grid)
    assert np.all(np.isfinite(rho))
    Er, rr = K.dos(m, E_range=[-1.5, 1.5], N_tilde=100)
    np.testing.assert_allclose(E, Er)
    np.testing.assert_allclose(rho, rr)
    Ee, re = K.dos(m, E_grid=Er)
    assert np.array_equal(Ee, Er) and np.array_equal(re, rr)
    _, ri = K.dos(m, E_grid=np.arange(-2, 3))
    assert ri.dtype == np.float64 and np.all(np.isfinite(ri))


def test_typed_conductivity_delegation(h):
    psi = manual_probes(11)
    m2 = K.cond_moments(h, JX, JY, NC=32, psi_in=psi.copy())
    ref = K.kpm_2d(h.H, JX, JY, 32, NR, NH, psi_in=psi.copy())
    assert np.array_equal(m2.mu, ref)
    assert K.kubo_bastin_cond(m2, 0.1, area=1.0) == K.kubo_bastin_cond_raw(ref, h.a, 0.1, b=h.b, NH=NH,
                                                                            area=1.0)
    with pytest.raises(K.ArgumentError):
        K.kubo_bastin_cond(m2, 0.1, area=1.0, NH=10)
    with pytest.raises(K.ArgumentError):
        K.ConductivityMoments(np.zeros((3, 2), dtype=complex), 1.0, 0.0, NH, 1)


def test_centered_rescaling_forwards_b():
    hc = K.rescale(H + 0.7 * sp.identity(NH), center=True)
    assert hc.b != 0.0
    psi = K.random_phase_vectors(Xoshiro256pp(21), NH, NR)
    mc = K.dos_moments(hc, NC=64, psi_in=psi.copy())
    ref = K.kpm_1d(hc.H, 64, NR, psi_in=psi.copy())
    assert np.array_equal(mc.mu, ref)
    for a, b in zip(K.dos(mc), K.dos_raw(ref, hc.a, b=hc.b)):
        assert np.array_equal(a, b)
    grid = np.linspace(hc.b - 0.5, hc.b + 0.5, 21)
    assert np.array_equal(K.dos(mc, E_grid=grid)[1], K.dos_raw(ref, hc.a, b=hc.b, E_grid=grid)[1])
    m2c = K.cond_moments(hc, JX, JY, NC=32, psi_in=psi.copy())
    r2 = K.kpm_2d(hc.H, JX, JY, 32, NR, NH, psi_in=psi.copy())
    assert K.kubo_bastin_cond(m2c, 0.3, area=1.0) == K.kubo_bastin_cond_raw(r2, hc.a, 0.3, b=hc.b, NH=NH,
                                                                             area=1.0)


def test_validation_reproducibility_and_sum_rules(h):
    m1 = K.dos_moments(h, NC=64, NR=4, rng=Xoshiro256pp(42))
    m2 = K.dos_moments(h, NC=64, NR=4, rng=Xoshiro256pp(42))
    m3 = K.dos_moments(h, NC=64, NR=4, rng=Xoshiro256pp(43))
    assert np.array_equal(m1.mu, m2.mu) and not np.allclose(m1.mu, m3.mu)
    assert abs(m1.mu[0] - 1.0) < 1e-12
    np.testing.assert_array_equal(K.random_phase_vectors(Xoshiro256pp(5), NH, NR), manual_probes(5))
    psi = K.random_phase_vectors(Xoshiro256pp(9), NH, NR)
    mom = K.dos_moments(h, NC=64, psi_in=psi.copy())
    cm = K.cond_moments(h, JX, JY, NC=32, psi_in=psi.copy())
    with pytest.raises(K.ArgumentError):
        K.dos_moments(h, NC=64, rng=Xoshiro256pp(1), psi_in=psi)
    with pytest.raises(K.ArgumentError):
        K.dos_moments(h, NC=63)
    with pytest.raises(K.ArgumentError):
        K.dos(mom, b=0.2)
    with pytest.raises(K.ArgumentError):
        K.dos_moments(h, NC=64, psi_in=np.zeros((NH - 1, NR), dtype=complex))
    with pytest.raises(K.ArgumentError):
        K.cond_moments(h, JX, JY, NC=32, psi_in=np.zeros((NH - 1, NR), dtype=complex))
    with pytest.raises(K.ArgumentError):
        K.kubo_bastin_cond(cm, 0.1, area=1.0, b=0.2)
    # identity probes: mu_0 = 1 exactly, and the Gauss-Chebyshev rule that
    # inverts 1/sqrt(1-x^2) integrates the NC-term DOS exactly for N > NC
    mf = K.dos_moments(h, NC=128, psi_in=np.eye(NH, dtype=complex))
    assert abs(mf.mu[0] - 1.0) <= 1e-12
    Nq = 512
    xq = np.cos((np.arange(1, Nq + 1) - 0.5) * math.pi / Nq)
    _, rq = K.dos(mf, E_grid=h.a * xq + h.b)
    assert abs(np.sum(h.a * math.pi * np.sqrt(1 - xq ** 2) * rq) / Nq - 1.0) <= 1e-12
    # the default window b +- 0.99 a misses the band-edge tails (2.3e-5)
    Et, rt = K.dos(mf, N_tilde=4096)
    assert abs(np.sum((rt[:-1] + rt[1:]) * np.diff(Et)) / 2 - 1.0) <= 1e-4
    E, rho = K.dos(h, NC=128, NR=4, rng=Xoshiro256pp(7))
    assert np.all(np.isfinite(rho))
    assert "DosMoments" in repr(mom) and "RescaledHamiltonian" in repr(h)
