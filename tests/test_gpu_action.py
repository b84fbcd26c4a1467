"""Parity of the matrix-function action (kpm_cheb_action through the C ABI)
and its consumers with the oracle, ED and the golden fixture.

FP64: 1e-12 against the oracle's per-step accumulation and the fixture
(test/chebyshev_action_test.jl uses 1e-12 against its direct recurrence).
FP32 recurrence vectors: bounded against the FP64 path at 2e-5 relative to
the output scale.  The physics checks follow test/evolution_test.jl and
test/chern_marker_test.jl with their tolerances.
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def prescaled(rng, NH, margin=0.95, density=None):
    A = rng.standard_normal((NH, NH)) + 1j * rng.standard_normal((NH, NH))
    if density is not None:
        mask = rng.random((NH, NH)) < density
        A = np.where(mask | mask.T | np.eye(NH, dtype=bool), A, 0)
    Hd = (A + A.conj().T) / 2
    Hd /= np.abs(np.linalg.eigvalsh(Hd)).max() / margin
    return Hd, sp.csc_matrix(Hd)


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_action_golden_fixture():
    g = dict(np.load(os.path.join(GOLD, "action_dense.npz"), allow_pickle=False))
    out = K.chebyshev_action(sp.csc_matrix(g["H"]), g["V"], g["C"])
    np.testing.assert_allclose(out, g["out"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("NH,NR,NC,Kc", [(48, 2, 16, 3), (64, 1, 1, 2), (64, 3, 2, 1),
                                         (64, 3, 3, 4), (100, 5, 17, 1), (100, 8, 40, 20),
                                         (96, 64, 33, 2), (80, 130, 20, 3)])
def test_action_vs_oracle(NH, NR, NC, Kc):
    print(f"action NH={NH} NR={NR} NC={NC} K={Kc}", flush=True)
    rng = np.random.default_rng(1000 + NH + NR + NC + Kc)
    Hd, Hs = prescaled(rng, NH, density=0.2)
    V = crand(rng, NH, NR)
    C = crand(rng, NC, Kc)
    out = K.chebyshev_action(Hs, V, C)
    ref = oracle.cheb_action(Hs, V, C)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_action_dense_ed_and_real_inputs():
    rng = np.random.default_rng(151)
    NH, NC, NR, Kc = 48, 16, 2, 3
    Hd, Hs = prescaled(rng, NH)
    V = crand(rng, NH, NR)
    C = crand(rng, NC, Kc)
    out = K.chebyshev_action(Hs, V, C)
    np.testing.assert_allclose(out, oracle.action_dense(Hd, V, C), atol=1e-12)
    np.testing.assert_allclose(out, oracle.action_ed(Hd, V, C), atol=1e-8)
    # real H (real-valued CSR path), real probes, real / complex coefficients
    Hr = np.real(Hd)
    Hr /= np.abs(np.linalg.eigvalsh(Hr)).max() / 0.95
    Vr = rng.standard_normal((NH, NR))
    Cr = rng.standard_normal(NC)
    got = K.chebyshev_action(sp.csc_matrix(Hr), Vr, Cr)
    np.testing.assert_allclose(got, oracle.action_dense(Hr, Vr, Cr)[:, :, 0], atol=1e-12)


def test_action_shapes_and_padding():
    rng = np.random.default_rng(163)
    NH, NC, NR, Kc = 32, 10, 3, 2
    _, Hs = prescaled(rng, NH)
    V = crand(rng, NH, NR)
    C = crand(rng, NC, Kc)
    full = K.chebyshev_action(Hs, V, C)
    assert full.shape == (NH, NR, Kc)
    np.testing.assert_allclose(K.chebyshev_action(Hs, V, C[:, 0]), full[:, :, 0], atol=1e-14)
    np.testing.assert_allclose(K.chebyshev_action(Hs, V[:, 0], C), full[:, 0, :], atol=1e-14)
    np.testing.assert_allclose(K.chebyshev_action(Hs, V[:, 0], C[:, 0]), full[:, 0, 0], atol=1e-14)
    out = np.zeros((NH, NR, Kc), dtype=complex)
    slots = (np.zeros((NH, NR), complex), np.zeros((NH, NR), complex))
    K.chebyshev_action_(out, Hs, V, C, slots=slots)
    np.testing.assert_allclose(out, full, atol=1e-14)
    Cpad = np.vstack([C, np.zeros((8, Kc))])
    np.testing.assert_allclose(K.chebyshev_action(Hs, V, Cpad), full, atol=1e-14)


def test_action_fp32_recurrence():
    rng = np.random.default_rng(167)
    NH, NC, NR, Kc = 2000, 64, 8, 3
    Hd, Hs = prescaled(rng, 200)
    Hs = sp.block_diag([Hs] * 10, format="csc")
    V = crand(rng, NH, NR)
    C = crand(rng, NC, Kc) / np.arange(1, NC + 1)[:, None]
    f64 = K.chebyshev_action(Hs, V, C)
    f32 = K.chebyshev_action(Hs, V, C, precision="f32")
    assert np.abs(f32 - f64).max() <= 2e-5 * np.abs(f64).max()


def test_action_guard():
    rng = np.random.default_rng(179)
    NH, NR = 32, 2
    Hd, Hs = prescaled(rng, NH)
    V = crand(rng, NH, NR)
    C = crand(rng, 64, 1)
    bad = sp.csc_matrix(1.2 * Hd)
    with pytest.raises(K.KPMUnstableError, match="Chebyshev recurrence is unstable") as ei:
        K.chebyshev_action(bad, V, C)
    with pytest.raises(oracle.OracleUnstable) as eo:
        oracle.cheb_action(bad, V, C)
    assert f"at iteration {eo.value.iteration} " in str(ei.value)
    assert K.chebyshev_action(bad, V, C, check_every=0).shape == (NH, NR, 1)
    np.testing.assert_allclose(K.chebyshev_action(Hs, 100 * V, C), 100 * K.chebyshev_action(Hs, V, C),
                               atol=1e-9)
    H1 = sp.csc_matrix(np.array([[1.1 + 0j]]))
    C15 = np.zeros(16)
    C15[15] = 1.0
    for amp in (1.0, 1e-8, 1e-12):
        with pytest.raises(K.KPMUnstableError):
            K.chebyshev_action(H1, np.array([[amp + 0j]]), C15)
    z = K.chebyshev_action(H1, np.array([[0j]]), C15)
    assert np.all(z == 0)


# ---- evolution (test/evolution_test.jl) ----------------------------------------------


def fixture(H):
    lam = np.linalg.eigvalsh(H.toarray())
    b = (lam.max() + lam.min()) / 2
    a = (lam.max() - lam.min()) / 2 / 0.95
    Hn = ((H - b * sp.identity(H.shape[0])) / a).tocsc()
    return K.RescaledHamiltonian(Hn, a, b), H.toarray()


def unit(rng, *shape):
    v = crand(rng, *shape)
    return v / np.linalg.norm(v, axis=0)


def test_evolve_vs_dense():
    rng = np.random.default_rng(191)
    N = 96
    for Hp in (models.ring_model(N) + 0.7 * sp.identity(N), models.flux_ring_model(N)):
        h, Hd = fixture(sp.csc_matrix(Hp))
        psi = unit(rng, N)
        for t in (0.0, 0.7, 3.3, -1.9):
            np.testing.assert_allclose(K.evolve(h, psi, t), oracle.ed_evolve(Hd, psi, t), atol=1e-10)


def test_evolve_norm_reversibility_group():
    rng = np.random.default_rng(199)
    N = 64
    h, _ = fixture(sp.csc_matrix(models.ring_model(N) + 0.7 * sp.identity(N)))
    psi = unit(rng, N)
    np.testing.assert_allclose(K.evolve(h, psi, 0.0), psi, atol=1e-13)
    for t in (0.5, 5.0, -11.0, 40.0):
        assert abs(np.linalg.norm(K.evolve(h, psi, t)) - 1) < 1e-11
    for t in (0.9, 4.7):
        np.testing.assert_allclose(K.evolve(h, K.evolve(h, psi, t), -t), psi, atol=1e-10)
    t1, t2 = 1.7, -0.6
    np.testing.assert_allclose(K.evolve(h, psi, t1 + t2), K.evolve(h, K.evolve(h, psi, t1), t2),
                               atol=1e-10)
    h0 = K.RescaledHamiltonian(h.H, h.a, 0.0)
    for t in (1.3, -2.1):
        np.testing.assert_allclose(K.evolve(h, psi, t), np.exp(-1j * h.b * t) * K.evolve(h0, psi, t),
                                   atol=1e-13)


def test_evolve_multitime_shapes_and_large_at():
    rng = np.random.default_rng(211)
    N, NR = 48, 3
    h, Hd = fixture(sp.csc_matrix(models.flux_ring_model(N)))
    psi = unit(rng, N)
    Psi = unit(rng, N, NR)
    ts = [-1.9, 0.0, 0.7, 3.3]
    assert K.evolve(h, psi, 0.7).shape == (N,)
    assert K.evolve(h, Psi, 0.7).shape == (N, NR)
    assert K.evolve(h, psi, ts).shape == (N, len(ts))
    multi = K.evolve(h, Psi, ts)
    assert multi.shape == (N, NR, len(ts))
    for k, t in enumerate(ts):
        np.testing.assert_allclose(multi[:, :, k], K.evolve(h, Psi, t), atol=5e-12)
    np.testing.assert_allclose(K.evolve(h, Psi, 3.3), oracle.ed_evolve(Hd, Psi, 3.3), atol=1e-10)
    N = 512
    h, Hd = fixture(sp.csc_matrix(models.ring_model(N)))
    t = 200.0 / h.a
    psi = unit(rng, N)
    NC = K.evolution_order(h.a, t)
    assert 200 < NC < 320
    pt = K.evolve(h, psi, t)
    assert abs(np.linalg.norm(pt) - 1) < 1e-10
    np.testing.assert_allclose(K.evolve(h, pt, -t), psi, atol=1e-9)
    np.testing.assert_allclose(pt, oracle.ed_evolve(Hd, psi, t), atol=1e-8)


def test_evolve_truncation_and_guard():
    rng = np.random.default_rng(227)
    N = 48
    h, _ = fixture(sp.csc_matrix(models.ring_model(N)))
    psi = unit(rng, N)
    with pytest.warns(UserWarning, match="evolution series tail"):
        pt = K.evolve(h, psi, 50.0 / h.a, NC=8)
    assert abs(np.linalg.norm(pt) - 1) > 1e-3
    bad = K.RescaledHamiltonian((1.2 * h.H).tocsc(), h.a, h.b)
    with pytest.raises(K.KPMUnstableError, match="Chebyshev recurrence is unstable"):
        K.evolve(bad, psi, 30.0 / h.a)
    for x, at in ((0.0, 1995.4), (0.3, 500.0), (-0.7, 3000.0)):
        a, b = 2.0, 0.4
        hx = K.RescaledHamiltonian(sp.csc_matrix(np.array([[x + 0j]])), a, b)
        got = K.evolve(hx, np.array([1.0 + 0j]), at / a)[0]
        assert abs(got - np.exp(-1j * (a * x + b) * at / a)) <= 1e-12


# ---- Fermi projector and Chern marker (test/chern_marker_test.jl) ---------------------


def test_fermi_projector_vs_dense():
    rng = np.random.default_rng(211)
    NH, NC = 64, 512
    a, b = 2.3, 0.4
    lam = np.concatenate([-0.9 + 0.7 * rng.random(NH // 2), 0.2 + 0.7 * rng.random(NH // 2)])
    Q, _ = np.linalg.qr(crand(rng, NH, NH))
    Hn = (Q * lam) @ Q.conj().T
    Hn = (Hn + Hn.conj().T) / 2
    h = K.RescaledHamiltonian(sp.csc_matrix(Hn), a, b)
    P = (Q * (lam < 0)) @ Q.conj().T
    V = crand(rng, NH, 3)
    PV = K.fermi_projector(h, V, Ef=b, NC=NC)
    np.testing.assert_allclose(PV, P @ V, atol=1e-3)
    # the same series through ED is exact to roundoff
    c = K.fermi_coefficients(a, b, b, NC=NC)
    np.testing.assert_allclose(PV, oracle.action_ed(Hn, V, c)[:, :, 0], atol=1e-10)
    v = V[:, 0]
    Pv = K.fermi_projector(h, v, Ef=b, NC=NC)
    assert Pv.shape == (NH,)
    np.testing.assert_allclose(Pv, PV[:, 0], atol=1e-12)
    u = crand(rng, NH)
    Pu = K.fermi_projector(h, u, Ef=b, NC=NC)
    assert abs(np.vdot(u, Pv) - np.vdot(Pu, v)) < 1e-10
    assert np.linalg.norm(K.fermi_projector(h, Pv, Ef=b, NC=NC) - Pv) < 2e-3
    beta = 20.0
    PT = (Q * (1 / (np.exp(a * lam * beta) + 1))) @ Q.conj().T
    np.testing.assert_allclose(K.fermi_projector(h, V, Ef=b, beta=beta, NC=NC), PT @ V, atol=1e-3)
    hd = K.RescaledHamiltonian(sp.csc_matrix(np.diag([-0.5, 0.0, 0.5]).astype(complex)), 1.0, 0.0)
    for NCc in (32, 512):
        p = K.fermi_projector(hd, np.array([0, 1, 0], dtype=complex), Ef=0.0, NC=NCc)
        assert abs(p[1].real - 0.5) < 1e-12


def marker_fixture(H, margin=0.95):
    ev = np.linalg.eigvalsh(H.toarray())
    b = (ev.max() + ev.min()) / 2
    a = (ev.max() - ev.min()) / 2 / margin
    return K.RescaledHamiltonian(((H - b * sp.identity(H.shape[0])) / a).tocsc(), a, b)


def cells(Ly, xs, ys):
    return [s for x in xs for y in ys
            for s in (2 * ((y - 1) + Ly * (x - 1)) + 1, 2 * ((y - 1) + Ly * (x - 1)) + 2)]


def test_chern_marker_topological_and_trivial():
    Lx = Ly = 10
    NC = 512
    bulk = cells(Ly, range(4, 8), range(4, 8))
    for m, C in ((0.0, 1.0), (1.6, 0.0)):
        H, pos, Ac = models.haldane_open_model(Lx, Ly, m=m)
        h = marker_fixture(H)
        mk = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=bulk, NC=NC)
        assert abs(K.chern_marker_average(mk, area=16 * Ac) - C) < 0.05
        mk_ed = oracle.ed_chern_marker(H, pos[:, 0], pos[:, 1], 0.0)
        np.testing.assert_allclose(mk, mk_ed[np.array(bulk) - 1], atol=0.02)
        if C:
            for k in range(16):
                assert abs(K.chern_marker_average(mk[2 * k:2 * k + 2], area=Ac) - C) < 0.1


def test_chern_marker_whole_sample_batching_and_region():
    Lx = Ly = 10
    NC = 512
    H, pos, Ac = models.haldane_open_model(Lx, Ly)
    h = marker_fixture(H)
    allsites = list(range(1, 2 * Lx * Ly + 1))
    mk = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=allsites, NC=NC)
    assert abs(mk.sum()) < 2e-3 * np.abs(mk).sum()
    mk2 = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=allsites, NC=2 * NC)
    assert abs(mk2.sum()) < abs(mk.sum()) / 2 and abs(mk2.sum()) < 0.1
    edge = cells(Ly, range(1, 2), range(1, Ly + 1))
    assert abs(mk[np.array(edge) - 1].sum() / (Ly * Ac) - 1.0) > 0.3
    for bs in (7, 200):
        mkb = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=allsites, NC=NC, batch_size=bs)
        np.testing.assert_allclose(mkb, mk, rtol=1e-12, atol=1e-13)
    region = cells(Ly, range(4, 8), range(4, 8))
    det = mk[np.array(region) - 1].sum()
    from kpmhip.rng import Xoshiro256pp
    est = K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=region,
                                rng=Xoshiro256pp(7), NR=32, NC=NC)
    assert est.shape == (32,)
    assert abs(est.mean() - det) < 4 * est.std(ddof=1) / math.sqrt(32)
    assert abs(K.chern_marker_average(est.mean(), area=16 * Ac) - 1.0) < 0.2
    # singleton region: every probe is exact (pins the |R| normalisation)
    i0 = region[0]
    m0 = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=[i0], NC=NC)
    est1 = K.chern_marker_region(h, pos[:, 0], pos[:, 1], Ef=0.0, region=[i0],
                                 rng=Xoshiro256pp(3), NR=4, NC=NC)
    np.testing.assert_allclose(est1, np.full(4, m0[0]), atol=1e-10)


def test_chern_marker_finite_temperature():
    Lx = Ly = 10
    NC = 512
    H, pos, Ac = models.haldane_open_model(Lx, Ly)
    h = marker_fixture(H)
    bulk = cells(Ly, range(4, 8), range(4, 8))
    mk = K.chern_marker(h, pos[:, 0], pos[:, 1], Ef=0.0, sites=bulk, beta=20.0, NC=NC)
    assert abs(K.chern_marker_average(mk, area=16 * Ac) - 1.0) < 0.1
    mk_ed = oracle.ed_chern_marker(H, pos[:, 0], pos[:, 1], 0.0, beta=20.0)
    np.testing.assert_allclose(mk, mk_ed[np.array(bulk) - 1], atol=0.02)
