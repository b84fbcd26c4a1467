"""Host-side logic of the package (no GPU): damping kernels, the DOS
derivative path, rescaling, the planner, the probe generator and the model
builders; and the oracle's reconstruction heads (dos, kubo_bastin_cond)
against the reference's physics checks.  Moments come from the oracle."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp, splitmix64_stream


# ---- probe generator known answers ------------------------------------------------

def test_splitmix64_known_answer():
    # splitmix64 seeded with 0: the first outputs of Vigna's reference
    # splitmix64.c (also the known-answer vector of rand_xoshiro's SplitMix64)
    assert [int(x) for x in splitmix64_stream(0, 3)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_xoshiro256pp_known_answer():
    # state (1, 2, 3, 4): the first ten outputs of Vigna's reference
    # xoshiro256plusplus.c, as published in rand_xoshiro's known-answer test
    # (the first is rotl(1 + 4, 23) + 1 = 41943041)
    st = np.array([[1], [2], [3], [4]], dtype=np.uint64)
    g = Xoshiro256pp(state=st)
    want = [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
            9973669472204895162, 14011001112246962877, 12406186145184390807,
            15849039046786891736, 10450023813501588000]
    assert [int(g.next_u64()[0]) for _ in range(10)] == want
    # every lane runs its own copy of the generator
    g2 = Xoshiro256pp(state=np.array([[1, 7], [2, 7], [3, 7], [4, 7]], dtype=np.uint64))
    assert [int(g2.next_u64()[0]) for _ in range(10)] == want


def test_random_phase_vectors_reproducible_and_unit():
    p1 = K.random_phase_vectors(Xoshiro256pp(7), 100, 4)
    p2 = K.random_phase_vectors(Xoshiro256pp(7), 100, 4)
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_allclose(np.linalg.norm(p1, axis=0), 1.0, rtol=1e-14)
    np.testing.assert_allclose(np.abs(p1) * 10, 1.0, rtol=1e-12)  # equal-modulus phases


# ---- damping kernels (test/test_kernel.jl) ----------------------------------------------

@pytest.mark.parametrize("kernel", [K.JacksonKernel, K.LorentzKernels(1.0), K.LorentzKernels(2.0)])
def test_kernel_properties(kernel):
    NC = 15
    assert K.hn(0) == 1 and K.hn(1) == K.hn(2) == K.hn(3) == 2
    m1 = K.muND_apply_kernel_and_h(np.ones(NC), NC, kernel, dims=[1])
    assert np.all(np.abs(m1.imag) < 1e-15)
    m1 = m1.real
    assert m1[0] == 1 < m1[1]
    assert m1[-1] < 2 / NC
    assert np.all(m1[1:-1] > m1[2:])
    m3 = K.muND_apply_kernel_and_h(np.ones((NC, NC, NC)), NC, K.JacksonKernel).real
    assert np.all(m3[0, :, :] == m3[:, 0, :]) and np.all(m3[:, 0, :] == m3[:, :, 0])
    np.testing.assert_allclose(m3[0] / m3[0, 0, 0], m3[1] / m3[1, 0, 0])


def test_jackson_closed_form_values():
    # g_0 = 1 exactly; g_n = 0 for n >= N
    assert K.JacksonKernel(0, 10) == pytest.approx(1.0, abs=1e-15)
    assert K.JacksonKernel(10, 10) == 0.0
    with pytest.raises(K.ArgumentError):
        K.LorentzKernels(0.0)


# ---- DOS reconstruction ------------------------------------------------------------------

def test_chain_dos_analytic_and_sum_rule():
    """test/dos_test.jl:214-249 with oracle moments."""
    N, t, shift = 4000, 1.0, 0.7
    H = models.chain(N, t=t, shift=shift)
    a, b, Hn = K.normalizeH(H, center=True, method="arpack")
    assert abs(b - shift) < 0.05
    NC, NR = 256, 24
    psi = K.random_phase_vectors(Xoshiro256pp(7), N, NR)
    mu = oracle.kpm_1d(Hn, NC, psi).mean(axis=0).real
    E, rho = oracle.dos(mu, a, b=b, N_tilde=1024)
    dE = E[1] - E[0]
    assert abs(np.sum(rho) * dE - 1.0) < 2e-2
    exact = lambda e: 1 / (np.pi * np.sqrt(4 * t * t - (e - shift) ** 2))  # noqa: E731
    idx = np.abs(E - shift) < 1.5 * t
    assert np.max(np.abs(rho[idx] - exact(E[idx]))) < 0.02


def test_dos_derivatives_vs_finite_differences_and_dos0():
    """test/dos_test.jl:143-193"""
    rng = Xoshiro256pp(23, lanes=4)
    NC, a, b = 48, 1.7, 0.35
    mu = rng.rand(NC) - 0.5
    Eg = b + a * np.array([-0.45, 0.2, 0.55])
    h = 1e-3 * a
    rho0 = lambda E: oracle.dos(mu, a, b=b, E_grid=[E], N_tilde=1)[1][0]  # noqa: E731
    stencil = [(-rho0(E + 2 * h) + 16 * rho0(E + h) - 30 * rho0(E) + 16 * rho0(E - h) -
                rho0(E - 2 * h)) / (12 * h * h) for E in Eg]
    _, d2 = K.dos_raw(mu, a, b=b, E_grid=Eg, N_tilde=3, dE_order=2)
    np.testing.assert_allclose(d2, stencil, rtol=1e-6)
    _, d1 = K.dos_raw(mu, a, b=b, E_grid=Eg, N_tilde=3, dE_order=1)
    hh = 1e-5 * a
    fd = [(rho0(E + hh) - rho0(E - hh)) / (2 * hh) for E in Eg]
    np.testing.assert_allclose(d1, fd, rtol=1e-5)
    _, d2c = K.dos_raw(mu, a, b=b, E_grid=[b], N_tilde=1, dE_order=2)
    assert d2c[0] == pytest.approx(K.dos0_raw(mu, a, dE_order=2), rel=1e-10)
    _, r0 = oracle.dos(mu, a, b=b, E_grid=[b], N_tilde=1)
    assert r0[0] == pytest.approx(K.dos0_raw(mu, a), rel=1e-10)


# ---- rescaling (test/dos_test.jl:195-212) -------------------------------------------------------

def test_normalizeH_hermiticity_and_center():
    rng = Xoshiro256pp(3, lanes=4)
    n = 100
    mask = rng.rand(n, n) < 0.05
    A = sp.csc_matrix(np.where(mask, (rng.rand(n, n) - 0.5) + 1j * (rng.rand(n, n) - 0.5), 0))
    with pytest.raises(K.ArgumentError):
        K.normalizeH(A + 2 * A.conj().T, method="arpack")
    H = A + A.conj().T
    a, Hn = K.normalizeH(H, method="arpack")
    assert np.max(np.abs(np.linalg.eigvalsh(Hn.toarray()))) < 1
    Hs = H + 0.7 * sp.identity(n)
    a, b, Hsn = K.normalizeH(Hs, center=True, method="arpack")
    ev = np.linalg.eigvalsh(Hsn.toarray())
    assert abs(b - 0.7) < 0.1 * a
    assert ev.min() > -1 and ev.max() < 1
    assert abs(ev.max() + ev.min()) < 0.02
    with pytest.raises(K.ArgumentError):
        K.normalizeH(H, fixed_a=2.0, center=True, method="arpack")


# ---- Kubo–Bastin (test/kubo_bastin_test.jl) ----------------------------------------------------

def _fhs_chern(t2, Nk=30):
    """Fukui–Hatsugai–Suzuki Berry flux of haldane_bloch (test/ed_reference.jl:256-305)."""
    nnn = ((1, 0), (-1, 1), (0, -1))

    def hk(k1, k2):
        ph = lambda off: np.exp(2j * np.pi * (k1 * off[0] + k2 * off[1]))  # noqa: E731
        hBA = -(1 + ph((1, 0)) + ph((0, 1)))
        hAA = 0j
        hBB = 0j
        for off in nnn:
            hAA += t2 * 1j * np.conj(ph(off)) + t2 * (-1j) * ph(off)
            hBB += t2 * (-1j) * np.conj(ph(off)) + t2 * 1j * ph(off)
        return np.array([[hAA, np.conj(hBA)], [hBA, hBB]])

    u = [[np.linalg.eigh(hk(i / Nk, j / Nk))[1][:, 0] for j in range(Nk)] for i in range(Nk)]
    tot = 0.0
    for i in range(Nk):
        for j in range(Nk):
            u00, u10 = u[i][j], u[(i + 1) % Nk][j]
            u11, u01 = u[(i + 1) % Nk][(j + 1) % Nk], u[i][(j + 1) % Nk]
            tot += np.angle(np.vdot(u00, u10) * np.vdot(u10, u11) * np.vdot(u11, u01) *
                            np.vdot(u01, u00))
    return -tot / (2 * np.pi)


def _ed_hall_T0(H, Jx, Jy, area, Ef):
    """TKNN form (test/ed_reference.jl:517-528)."""
    ev, U = np.linalg.eigh(H.toarray())
    Va = U.conj().T @ (Jx.toarray() / 1j) @ U
    Vb = U.conj().T @ (Jy.toarray() / 1j) @ U
    occ = ev <= Ef
    acc = 0.0
    for m in np.where(occ)[0]:
        for n in np.where(~occ)[0]:
            acc += np.imag(Va[m, n] * Vb[n, m]) / (ev[m] - ev[n]) ** 2
    return 2 * np.pi * (-2 / area) * acc


def test_haldane_sigma_xy_quantised():
    t2 = 0.2
    H, Jx, Jy, area = models.haldane_model(6, 6, t=1.0, t2=t2)
    D = H.shape[0]
    a, b, Hn = K.normalizeH(H, center=True, method="arpack")
    NC = 96
    psi = np.eye(D, dtype=complex)
    mu = oracle.kpm_2d(Hn, Jx, Jy, NC, psi, psi, NC, NC)
    C_fhs = _fhs_chern(t2)
    C = round(C_fhs)
    assert abs(C) == 1 and abs(C_fhs - C) < 1e-10
    s_ed = _ed_hall_T0(H, Jx, Jy, area, 0.0)
    assert abs(s_ed - C) < 0.02
    s = oracle.kubo_bastin_cond(mu, a, 0.0, b=b, NH=D, area=area)
    assert s == pytest.approx(s_ed, abs=0.03)
    assert s == pytest.approx(C, abs=0.05)
    sT = oracle.kubo_bastin_cond(mu, a, 0.0, b=b, NH=D, area=area, beta=50.0)
    assert sT == pytest.approx(s, abs=0.01)
    # a^2 counterfactual: currents from H_norm divide sigma by a^2
    Jxn, Jyn = Jx / a, Jy / a
    mu_n = oracle.kpm_2d(Hn, Jxn, Jyn, NC, psi, psi, NC, NC)
    sn = oracle.kubo_bastin_cond(mu_n, a, 0.0, b=b, NH=D, area=area)
    assert sn == pytest.approx(s / a ** 2, rel=1e-8)


def test_oracle_kubo_factorisation_vs_direct_gamma():
    """The oracle's factorised node contraction (two GEMMs per chunk,
    conductivity.jl:143-155) equals the direct double sum of Gamma_nm."""
    rng = np.random.default_rng(5)
    NC, a, b, NH, area = 7, 1.3, 0.1, 10, 2.0
    mu = rng.standard_normal((NC, NC)) + 1j * rng.standard_normal((NC, NC))
    quad_N = 16
    s = oracle.kubo_bastin_cond(mu, a, 0.2, b=b, NH=NH, area=area, quad_N=quad_N)
    w = oracle.jackson(np.arange(NC), NC) * oracle.hn(np.arange(NC))
    mt = mu * w[:, None] * w[None, :]
    x, wt = oracle.gauss_chebyshev(quad_N)
    E = a * x + b
    f = ((E < 0.2).astype(float) + (E <= 0.2).astype(float)) / 2
    keep = (np.abs(x) < 1 - 1e-3) & (f != 0)
    acc = sum(wt[k] * f[k] / (1 - x[k] ** 2) ** 1.5 * oracle.kubo_gamma_direct(mt, x[k])
              for k in np.where(keep)[0])
    assert s == pytest.approx(-2 * NH / (area * a ** 2) * acc.real, rel=1e-12)


def test_gauss_chebyshev_sum_rule():
    """Gauss–Chebyshev integrates T_n exactly: sum_k w_k T_n(x_k) = π δ_n0
    (frontend_test.jl:170-187 checks the same quadrature)."""
    x, w = K.gausschebyshevt(64)
    for n in range(0, 100):
        s = np.sum(w * np.cos(n * np.arccos(x)))
        assert s == pytest.approx(math.pi if n == 0 else 0.0, abs=1e-12)


# ---- planner ---------------------------------------------------------------------------------

def test_kpm_2d_blocking_mirror_matches_oracle():
    NH, NR, NC = 40, 3, 50
    one = NH * NR * 16
    for budget, arr, right in ((30 * one, "auto", 4), (10**6, "auto", "auto"), (10 * one, 27, 9),
                               (100 * one, "auto", 3), (50 * one, 2, "auto")):
        py = K.kpm_2d_blocking(NH, NR, NC, workspace_bytes=budget, arr_size=arr,
                               right_block=right)
        orc = oracle.kpm_2d_blocking(NH, NR, NC, budget, 0 if arr == "auto" else arr,
                                     0 if right == "auto" else right)
        assert py == orc
    with pytest.raises(K.ArgumentError):
        K.kpm_2d_blocking(NH, NR, NC, workspace_bytes=10 * one - 1)
    with pytest.raises(K.ArgumentError):
        K.kpm_2d_blocking(0, NR, NC, workspace_bytes=10 * one)
    with pytest.raises(K.ArgumentError):
        K.kpm_2d_blocking(NH, NR, NC, workspace_bytes=10 * one, arr_size=1)
    with pytest.raises(K.ArgumentError):
        K.kpm_2d_blocking(NH, NR, NC, workspace_bytes=-1)


def test_engine_blocking_prefers_resident_slots():
    p = K.engine_blocking(1000, 4, 64, workspace_bytes=10**9)
    assert p["arr_size"] == 64 and p["right_block"] == 64
    q = K.engine_blocking(1000, 4, 64, workspace_bytes=30 * 1000 * 4 * 16)
    assert 2 <= q["arr_size"] < 64 and q["right_block"] >= 1
    r = K.engine_blocking(1000, 4, 64, workspace_bytes=10**9, arr_size=5, right_block=3)
    assert (r["arr_size"], r["right_block"]) == (5, 3)


# ---- models --------------------------------------------------------------------------------------

def test_models_hermitian_and_sizes():
    H = models.graphene(8, 6, disorder=0.5, seed=1)
    assert H.shape == (96, 96) and H.nnz == 4 * 96
    assert abs(H - H.conj().T).max() == 0
    Hs, Jx, Jy = models.square_lattice(5, 4, currents=True)
    assert Hs.nnz == 4 * 20 and abs(Hs - Hs.conj().T).max() == 0
    assert abs(Jx + Jx.conj().T).max() < 1e-15  # bond currents are anti-Hermitian
    Hh, hx, hy, area = models.haldane_model(4, 4)
    assert abs(Hh - Hh.conj().T).max() < 1e-15
    assert area == pytest.approx(16 * 1.5 * math.sqrt(3))
    Hc, cx, cy, cz, vol = models.cubic_model(4, W=2.0)
    assert Hc.nnz == 7 * 64 and vol == 64.0
    ev = np.linalg.eigvalsh(models.chain(10, shift=0.3).toarray())
    np.testing.assert_allclose(np.sort(ev), np.sort(0.3 - 2 * np.cos(2 * np.pi * np.arange(10) / 10)),
                               atol=1e-12)


def test_bond_gradient_bookkeeping():
    """kpm_bond_gradient (host, in libkpm_hip) equals the per-pair formula
    Re(u M_ij + conj(u) M_ji)/a evaluated by scipy indexing."""
    H = models.square_disordered(6, seed=4).tolil()
    H[0, 1] = H[0, 1] * np.exp(0.7j)
    H[1, 0] = np.conj(H[0, 1])
    H = sp.csc_matrix(H)
    A = sp.csc_matrix(H, dtype=np.complex128)
    A.sort_indices()
    rng = np.random.default_rng(2)
    G = sp.csc_matrix((rng.random(A.nnz) + 1j * rng.random(A.nnz), A.indices.copy(),
                       A.indptr.copy()), shape=A.shape)
    got = K.bond_gradient(H, G, 2.5).tocsr()
    C = sp.coo_matrix(H)
    m = C.row < C.col
    i, j, h = C.row[m], C.col[m], C.data[m]
    Gr = G.tocsr()
    gij, gji = np.asarray(Gr[i, j]).ravel(), np.asarray(Gr[j, i]).ravel()
    u = h / np.abs(h)
    ref = sp.coo_matrix((np.real(u * gij + np.conj(u) * gji) / 2.5, (i, j)), shape=A.shape).tocsr()
    assert got.nnz == ref.nnz and abs(got - ref).max() < 1e-15
    with pytest.raises(K.ArgumentError):  # gradient on a different pattern
        K.bond_gradient(H, sp.identity(A.shape[0], format="csc"), 2.5)


def test_bond_gradient_rejects_asymmetric_pattern():
    """A stored (i, j) without (j, i) is refused: the transposed entry is
    looked up by the per-column cursor and checked."""
    H = models.square_disordered(5, seed=4).tolil()
    H[0, 7] = 0.3  # (7, 0) absent
    H = sp.csc_matrix(H)
    A = sp.csc_matrix(H, dtype=np.complex128)
    A.sort_indices()
    G = sp.csc_matrix((np.ones(A.nnz, dtype=np.complex128), A.indices.copy(), A.indptr.copy()),
                      shape=A.shape)
    with pytest.raises(K.ArgumentError, match="not symmetric"):
        K.bond_gradient(H, G, 1.0)
