"""GPU parity of the BdG operators and the superfluid-stiffness driver
(test/stiffness_test.jl:291-694 on the 3 x 3 periodic square lattice).

The moment tables run on the device engine, from the assembled rescaled
Nambu matrix. The checks are:
- the tables against the oracle's kpm_2d / current restatement on the same
  matrix (1e-12);
- the physics against ED: the Lehmann sum, the diamagnetic trace and the
  free-energy curvature, at the reference's tolerances;
- the exact identities:
  - zero-gap cancellation;
  - parity masks;
  - prefactor routing.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
from kpmhip.bdg import ScaledOperator, _randn_complex, rescale_bdg
from kpmhip.rng import Xoshiro256pp
import oracle as O

pytestmark = pytest.mark.gpu

H, POS, DISP = models.square_model(3, 3)
QY = np.array([0.0, 2 * math.pi / 3])
PSI = np.eye(18, dtype=np.complex128)
DF = np.full(9, 0.45 + 0.2j)


def op_with(Delta):
    return K.BdGOperator(H, mu=-0.7, U=2.0, n=np.full(9, 0.4), Delta=Delta)


def test_spectral_radius_and_stability_probe():
    op = op_with(DF)
    ev = np.linalg.eigvalsh(K.bdg_assemble(op).toarray())
    r, v = K.spectral_radius(op)
    rho = np.abs(ev).max()
    assert r <= rho * (1 + 1e-12) and r >= rho * (1 - 1e-3)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    rng = np.random.default_rng(5)
    A = rng.standard_normal((200, 200)) + 1j * rng.standard_normal((200, 200))
    Hr = sp.csc_matrix((A + A.conj().T) / 2)
    r2, _ = K.spectral_radius(Hr)
    rho2 = np.abs(np.linalg.eigvalsh(Hr.toarray())).max()
    assert r2 <= rho2 * (1 + 1e-12) and r2 >= rho2 * (1 - 1e-2)
    # stability probe: max_n ||T_n(Hs) v|| against a dense recurrence from
    # the same start vector
    for scale in (1.0 / 0.9, 1.0 / 1.3):
        a = rho * scale
        Hs = ScaledOperator(op, a, 0.0)
        got = K.chebyshev_stability_probe(Hs, 18, 40)
        M = K.bdg_assemble(op).toarray() / a
        v0 = _randn_complex(Xoshiro256pp(1), 18)
        prev = v0 / np.linalg.norm(v0)
        cur = M @ prev
        best = max(1.0, np.linalg.norm(cur))
        for _ in range(3, 41):
            prev = 2 * M @ cur - prev
            best = max(best, np.linalg.norm(prev))
            prev, cur = cur, prev
        if scale > 1:
            assert abs(got - best) <= 1e-9 * best, (got, best)
        else:  # a spectrum beyond [-1, 1] is flagged (an infinite norm if the engine stops it)
            assert best > 1.5 and got > 1.5
            if math.isfinite(got):
                assert abs(got - best) <= 1e-9 * best, (got, best)


def test_tables_match_oracle_on_assembled_bdg():
    op = op_with(DF)
    rh = K.rescale(op, eps=0.2)
    assert isinstance(rh.H, ScaledOperator) and rh.b == 0.0
    Jq = K.nambu_current_q(H, POS, QY, dir=1, disp=DISP)
    Jmq = K.nambu_current_q(H, POS, -QY, dir=1, disp=DISP)
    NC = 48
    mu = K.kpm_2d(rh.H, Jq, Jmq, NC, 18, 18, psi_in=PSI)
    ref = O.kpm_2d(rh.H.host_matrix(), Jq, Jmq, NC, PSI, PSI, NC, NC)
    np.testing.assert_allclose(mu, ref, rtol=1e-11, atol=1e-12)
    Dh = K.nambu_diamagnetic(H, POS, disp=DISP)
    g = K.kpm_1d_current(rh.H, Dh, NC, 18, 18, psi_in=PSI)
    gref = O.kpm_1d_current(rh.H.host_matrix(), Dh, NC, PSI).mean(axis=0)
    np.testing.assert_allclose(g, gref, rtol=1e-11, atol=1e-12)


def test_zero_gap_cancellation():
    op0 = op_with(np.zeros(9))
    r = K.superfluid_stiffness(op0, POS, QY, beta=10.0, eta=0.3, NC=128, psi_in=PSI, volume=9.0, disp=DISP)
    assert abs(r.Ds_over_pi) < 1e-10
    assert r.a_SC == r.a_N
    rs = K.superfluid_stiffness(op0, POS, QY, beta=10.0, eta=0.3, NC=128, NR=4, volume=9.0, disp=DISP)
    assert abs(rs.Ds_over_pi) < 1e-10


def test_stiffness_vs_ed():
    opd = op_with(DF)
    r = K.superfluid_stiffness(opd, POS, QY, beta=10.0, eta=0.3, NC=256, psi_in=PSI, volume=1.0,
                               g_J=1.0, disp=DISP)
    Jq = K.nambu_current_q(H, POS, QY, dir=1, disp=DISP)
    Jmq = K.nambu_current_q(H, POS, -QY, dir=1, disp=DISP)
    Hd = O.bdg_matrix(H, -0.7, np.full(9, 2.0), np.full(9, 0.4), DF)
    Hn = O.bdg_matrix(H, -0.7, np.full(9, 2.0), np.full(9, 0.4), np.zeros(9))
    sc = O.ed_two_energy_response(Hd, Jq, Jmq, beta=10.0, eta=0.3)
    nn = O.ed_two_energy_response(Hn, Jq, Jmq, beta=10.0, eta=0.3)
    ds = nn.real - sc.real
    print(f"Pi_SC {r.Pi_SC} vs {sc}; Pi_N {r.Pi_N} vs {nn}; Ds {r.Ds_over_pi} vs {ds}", flush=True)
    assert abs(r.Pi_SC - sc) <= 0.05 * abs(sc)
    assert abs(r.Pi_N - nn) <= 0.05 * abs(nn)
    assert abs(r.Ds_over_pi - ds) <= 0.05 * abs(ds)
    if ds > 0:
        assert r.Ds_over_pi > 0


def test_parity_opt_in_and_prefactors():
    opr = op_with(np.full(9, 0.5 + 0.0j))
    rh = K.rescale(opr, eps=0.2)
    Jq = K.nambu_current_q(H, POS, QY, dir=1, disp=DISP)
    Jmq = K.nambu_current_q(H, POS, -QY, dir=1, disp=DISP)
    mu = K.kpm_2d(rh.H, Jq, Jmq, 64, 18, 18, psi_in=PSI)
    n, m = np.indices(mu.shape)
    odd = (m - n) % 2 == 1
    assert np.abs(mu[odd]).max() < 1e-12 * np.abs(mu).max()
    common = dict(beta=10.0, eta=0.3, NC=128, psi_in=PSI, volume=1.0, disp=DISP)
    rn = K.superfluid_stiffness(opr, POS, QY, moment_parity="NONE", **common)
    re = K.superfluid_stiffness(opr, POS, QY, moment_parity="EVEN", **common)
    for f in ("Pi_SC", "Pi_N", "Ds_over_pi"):
        a, b = getattr(re, f), getattr(rn, f)
        assert abs(a - b) <= 1e-10 * abs(b)
    # kpm_2d! with EVEN parity overwrites skipped entries with zero
    pre = np.ones((8, 8), dtype=np.complex128, order="F")
    K.kpm_2d_(rh.H, Jq, Jmq, 8, 18, 18, pre, PSI, PSI, moment_parity="EVEN")
    n, m = np.indices(pre.shape)
    assert np.all(pre[(m - n) % 2 == 1] == 0)
    opd = op_with(DF)
    c2 = dict(beta=10.0, eta=0.3, NC=64, psi_in=PSI, disp=DISP)
    r1 = K.superfluid_stiffness(opd, POS, QY, volume=1.0, **c2)
    r4 = K.superfluid_stiffness(opd, POS, QY, volume=4.0, **c2)
    rg = K.superfluid_stiffness(opd, POS, QY, volume=1.0, g_J=2.0, **c2)
    assert abs(r4.Pi_SC - r1.Pi_SC / 4) <= 1e-14 and abs(r4.Pi_N - r1.Pi_N / 4) <= 1e-14
    assert abs(r4.Ds_over_pi - r1.Ds_over_pi / 4) <= 1e-14
    assert abs(rg.Pi_SC - 2 * r1.Pi_SC) <= 1e-14 and abs(rg.Ds_over_pi - 2 * r1.Ds_over_pi) <= 1e-14


def test_diamagnetic_term_vs_ed():
    opd = op_with(DF)
    dia = K.diamagnetic_term(opd, POS, disp=DISP, beta=10.0, NC=256, psi_in=PSI, volume=1.0)
    Dh = K.nambu_diamagnetic(H, POS, disp=DISP)
    ed = O.ed_diamagnetic(O.bdg_matrix(H, -0.7, np.full(9, 2.0), np.full(9, 0.4), DF), Dh, beta=10.0)
    print(f"diamagnetic {dia} vs ED {ed}", flush=True)
    assert abs(dia - ed) <= 5e-3 * abs(ed)


def test_complete_stiffness_vs_free_energy_curvature():
    beta, volume = 10.0, 9.0
    U, nn = np.full(9, 2.0), np.full(9, 0.4)
    opd, op0 = op_with(DF), op_with(np.zeros(9))
    r = K.superfluid_stiffness(opd, POS, QY, beta=beta, eta=0.0, NC=256, psi_in=PSI, volume=volume,
                               disp=DISP, include_diamagnetic=True)

    def F(Delta, A):
        return O.ed_bdg_free_energy(H, POS, DISP, -0.7, U, nn, Delta, A, q=QY, dir=1, beta=beta)

    def curv(Delta, d):
        return (F(Delta, d) - 2 * F(Delta, 0.0) + F(Delta, -d)) / d ** 2

    anchor = (2 / volume) * (curv(DF, 1e-3) - curv(np.zeros(9), 1e-3))
    print(f"complete stiffness {r.Ds_over_pi_complete} vs curvature {anchor}", flush=True)
    assert abs(r.Ds_over_pi_complete - anchor) <= 2e-2 * abs(anchor)
    assert abs(r.Ds_over_pi_complete - r.Ds_over_pi) > 1e-4
    r0 = K.superfluid_stiffness(op0, POS, QY, beta=beta, eta=0.0, NC=256, psi_in=PSI, volume=volume,
                                disp=DISP, include_diamagnetic=True)
    assert abs(r0.Ds_over_pi_complete) < 1e-10


def test_non_assemblable_operator_is_rejected():
    op = op_with(DF)
    op.h = np.asarray(H.toarray())  # dense blocks cannot go to the device
    with pytest.raises(K.ArgumentError):
        K.kpm_2d(op, H, H, 4, 1, 18)


def test_scaled_operator_radius_and_rescale_contract():
    """test/bdg_test.jl:383-449 on the 4-site ring."""
    h = models.ring_model(4)
    D = np.full(4, 0.3 + 0.4j)
    op = K.BdGOperator(h, mu=-0.5, U=2.0, n=np.full(4, 0.5), Delta=D)
    Hd = O.bdg_matrix(h, -0.5, np.full(4, 2.0), np.full(4, 0.5), D)
    exact = np.abs(np.linalg.eigvalsh(Hd)).max()
    rad, v = K.spectral_radius(op)
    assert abs(rad - exact) <= 2e-3 * exact and rad <= exact + 1e-8 and v.shape == (8,)
    rh = K.rescale(op)
    assert rh.b == 0.0 and np.all(np.abs(np.linalg.eigvalsh(Hd / rh.a)) < 1)
    # the device form of (op - b I)/a through a real-coefficient SpMM
    S2 = ScaledOperator(op, 2.0, 0.3)
    rng = np.random.default_rng(8)
    x = rng.standard_normal((8, 3)) + 1j * rng.standard_normal((8, 3))
    y = rng.standard_normal((8, 3)) + 1j * rng.standard_normal((8, 3))
    X, Y = K.DeviceBlock.from_host(np.asfortranarray(x)), K.DeviceBlock.from_host(np.asfortranarray(y))
    K.spmm(S2, X, Y, alpha=1.3, beta=0.7)
    np.testing.assert_allclose(Y.download(), 1.3 * ((Hd - 0.3 * np.eye(8)) / 2.0) @ x + 0.7 * y,
                               atol=1e-12)
    m = K.dos_moments(rh, NC=32, NR=2, rng=Xoshiro256pp(1))
    assert np.all(np.isfinite(m.mu)) and abs(m.mu[0] - 1) < 1e-10
    assert K.gershgorin_bound(op) >= exact
    # an isolated edge: one of the two bounds must cover it
    Ne, true_r = 20_000, 2.0
    he = sp.diags(np.concatenate([[true_r], np.ones(Ne - 1)]), format="csc")
    ope = K.BdGOperator(he, mu=0.0, U=0.0)
    need = true_r / (1 - 0.2 / 2)
    assert K.rescale(ope).a >= need or K.rescale(ope, bound="gershgorin").a >= need
    op0 = K.BdGOperator(sp.csc_matrix((2, 2)), mu=0.0, U=0.0)
    with pytest.raises(K.ArgumentError):
        K.rescale(op0)
    for bad in (0.0, 2.0):
        with pytest.raises(K.ArgumentError):
            K.rescale(op, eps=bad)
