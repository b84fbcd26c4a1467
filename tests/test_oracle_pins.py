"""Pins of the oracle (CPU restatement of the reference engine) against the
golden fixtures, the reference tests' dense T_n(H) references, exact
diagonalisation and exact integer arithmetic.  Each check is chosen so that a
plausible slip in the oracle (a wrong ring slot, a dropped doubling term, a
missing conjugate, a transposed operand) fails it; see also
test_oracle_counterfactuals.py."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def phases(rng, n, nr):
    p = np.exp(2j * np.pi * rng.rand(n, nr))
    return np.asfortranarray(p / np.linalg.norm(p, axis=0))


def blocking_model(rng, L):
    """test/kpm2d_blocking_test.jl:7-23 (square + 0.03 random Hermitian, /1.2||H||)."""
    N = L * L
    H = np.zeros((N, N), dtype=complex)
    for y in range(L):
        for x in range(L):
            i = x + L * y
            if x < L - 1:
                H[i, i + 1] = H[i + 1, i] = -1
            if y < L - 1:
                H[i, i + L] = H[i + L, i] = -1
    A = (rng.rand(N, N) - 0.5) + 1j * (rng.rand(N, N) - 0.5)
    H += 0.03 * (A + A.conj().T) / 2
    H /= 1.2 * np.linalg.norm(H, 2)
    B = (rng.rand(N, N) - 0.5) + 1j * (rng.rand(N, N) - 0.5)
    Cm = (rng.rand(N, N) - 0.5) + 1j * (rng.rand(N, N) - 0.5)
    return H, (B + B.conj().T) / 2, (Cm + Cm.conj().T) / 2


def test_kpm1d_golden():
    g = load("kpm1d_dense.npz")
    NC = int(g["NC"])
    H = sp.csc_matrix(g["H"])
    np.testing.assert_allclose(oracle.kpm_1d(H, NC, g["psi"]), g["mu_eq"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["mu_eq"], g["mu_ed"], atol=1e-10)  # fixture self-consistency
    np.testing.assert_allclose(oracle.kpm_1d_lr(H, NC, g["psi_l"], g["psi_r"]), g["mu_lr"],
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.kpm_1d_current(H, sp.csc_matrix(g["J"]), NC, g["psi"]),
                               g["mu_cur"], rtol=1e-12, atol=1e-12)


def test_kpm2d_golden_all_blockings():
    g = load("kpm2d_dense.npz")
    NC = int(g["NC"])
    args = [sp.csc_matrix(g[k]) for k in ("H", "Ja", "Jb")]
    for KL in (2, 3, 5, NC, NC + 7):
        for KR in (1, 3, 16, NC):
            got = oracle.kpm_2d(*args, NC, g["psi_l"], g["psi_r"], KL, KR)
            np.testing.assert_allclose(got, g["mu"], rtol=1e-12, atol=1e-12)


def test_kpm3d_golden():
    g = load("kpm3d_dense.npz")
    NC = int(g["NC"])
    got = oracle.kpm_3d(*[sp.csc_matrix(g[k]) for k in ("H", "Ja", "Jb", "Jg")], NC, g["psi_l"],
                        g["psi_r"], 3, 3)
    np.testing.assert_allclose(got, g["mu"], rtol=1e-12, atol=1e-12)


def test_doubling_vs_direct_and_ed():
    """test/dos_test.jl:11-44 at NH=64, NC=32 (doubling == direct == ED)."""
    rng = Xoshiro256pp(42, lanes=4)
    NH, NC = 64, 32
    A = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    Hd = (A + A.conj().T) / 2
    Hd /= np.max(np.abs(np.linalg.eigvalsh(Hd))) / 0.95
    psi = phases(rng, NH, 1)
    mu = oracle.kpm_1d(sp.csc_matrix(Hd), NC, psi)[0].real
    direct = np.zeros(NC)
    prev, cur = psi[:, 0].copy(), Hd @ psi[:, 0]
    direct[0] = np.vdot(psi[:, 0], prev).real
    direct[1] = np.vdot(psi[:, 0], cur).real
    for n in range(2, NC):
        prev, cur = cur, 2 * (Hd @ cur) - prev
        direct[n] = np.vdot(psi[:, 0], cur).real
    np.testing.assert_allclose(mu, direct, atol=1e-10)
    np.testing.assert_allclose(mu, oracle.ed_moments(Hd, psi, NC)[0], atol=1e-8)


def test_dense_references_square_currents():
    rng = Xoshiro256pp(805, lanes=4)
    NH, NR, NC = 20, 3, 24
    A = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    H = (A + A.conj().T) / 2
    H /= 1.2 * np.linalg.norm(H, 2)
    J = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    assert np.linalg.norm(J - J.conj().T) > 1
    psi, pl, pr = phases(rng, NH, NR), phases(rng, NH, NR), phases(rng, NH, NR)
    Hs = sp.csc_matrix(H)
    np.testing.assert_allclose(oracle.kpm_1d(Hs, NC, psi), oracle.kpm1d_dense(H, psi, NC),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.kpm_1d_lr(Hs, NC, pl, pr), oracle.kpm1d_lr_dense(H, pl, pr, NC),
                               rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.kpm_1d_current(Hs, sp.csc_matrix(J), NC, psi),
                               oracle.current_dense(H, J, psi, NC), rtol=1e-12, atol=1e-12)


def test_kpm2d_blocked_vs_dense_and_invariance():
    """test/kpm2d_blocking_test.jl:73-117 on the blocking model (6x6)."""
    rng = Xoshiro256pp(801, lanes=4)
    H, Ja, Jb = blocking_model(rng, 6)
    NH, NR = 36, 3
    pl, pr = phases(rng, NH, NR), phases(rng, NH, NR)
    for NC in (17, 24):
        ref = oracle.kpm2d_dense(H, Ja, Jb, pl, pr, NC)
        got = oracle.kpm_2d(sp.csc_matrix(H), sp.csc_matrix(Ja), sp.csc_matrix(Jb), NC, pl, pr, 5, 3)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_parity_and_mn_sym():
    """test/kpm2d_blocking_test.jl:154-207"""
    rng = Xoshiro256pp(803, lanes=4)
    H, Ja, Jb = blocking_model(rng, 6)
    NH, NR, NC = 36, 3, 17
    psi = phases(rng, NH, NR)
    S = [sp.csc_matrix(x) for x in (H, Ja, Jb)]
    full = oracle.kpm_2d(*S, NC, psi, psi, 5, 3)
    odd = oracle.kpm_2d(*S, NC, psi, psi, 5, 3, parity=1)
    even = oracle.kpm_2d(*S, NC, psi, psi, 5, 3, parity=2)
    n, m = np.meshgrid(np.arange(1, NC + 1), np.arange(1, NC + 1), indexing="ij")
    odd_mask = (n + m) % 2 == 1
    np.testing.assert_allclose(odd, full * odd_mask, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(even, full * ~odd_mask, rtol=1e-12, atol=1e-12)
    sym_ref = full.copy()
    for mm in range(NC):
        for nn in range(mm + 1, NC):
            sym_ref[mm, nn] = sym_ref[mm, nn].real
            sym_ref[nn, mm] = sym_ref[mm, nn]
    sym = oracle.kpm_2d(*S, NC, psi, psi, 5, 3, mn_sym=True)
    # mn_sym computes only n <= m2 per rep; entries below are filled by symmetry
    np.testing.assert_allclose(sym, sym_ref, rtol=1e-12, atol=1e-12)


def test_blocking_arithmetic_exact():
    """test/kpm2d_blocking_test.jl:119-152 (exact integers)."""
    NH, NR, NC = 40, 3, 50
    one = NH * NR * 16
    c = oracle.kpm_2d_blocking(NH, NR, NC, 30 * one, 0, 4)
    assert c["arr_size"] == 18 and c["right_block"] == 4
    assert c["bytes"] == 11 * one + 18 * (one + 4 * 16)
    assert c["bytes"] <= 30 * one
    with pytest.raises(ValueError):
        oracle.kpm_2d_blocking(NH, NR, NC, 10 * one - 1, 0, 0)
    e = oracle.kpm_2d_blocking(NH, NR, NC, 10 * one, 27, 9)
    assert e["arr_size"] == 27 and e["right_block"] == 9


def test_chebyshev_iter_manual():
    """test/chebyshev_iteration_test.jl:11-43"""
    rng = Xoshiro256pp(1234, lanes=4)
    NH, NR, n = 4, 2, 6
    A = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    H = (A + A.conj().T) / 2
    s = np.zeros((NH, NR, n), dtype=complex)
    s[:, :, 0] = (rng.rand(NH, NR) - 0.5) + 1j * rng.rand(NH, NR)
    s[:, :, 1] = (rng.rand(NH, NR) - 0.5) + 1j * rng.rand(NH, NR)
    got = oracle.cheb_iter(sp.csc_matrix(H), s)
    exp = [s[:, :, 0], s[:, :, 1]]
    for _ in range(2, n):
        exp.append(2 * H @ exp[-1] - exp[-2])
    for i in range(n):
        assert np.linalg.norm(got[:, :, i] - exp[i]) < 1e-10


def test_unnormalised_scaling():
    """test/dos_test.jl:80-93: mu scales as |c|^2."""
    rng = Xoshiro256pp(13, lanes=4)
    NH, NC = 32, 16
    A = (rng.rand(NH, NH) - 0.5) + 1j * (rng.rand(NH, NH) - 0.5)
    H = sp.csc_matrix((A + A.conj().T) / (2 * NH))
    psi = phases(rng, NH, 1)
    np.testing.assert_allclose(oracle.kpm_1d(H, NC, 2 * psi), 4 * oracle.kpm_1d(H, NC, psi),
                               rtol=1e-10)


def test_chain_identity_probes_closed_form():
    """Identity probes give the exact trace mu_n = (1/D) sum_i T_n(lambda_i);
    for the periodic chain lambda_k = -2t cos(2πk/N)/a (integration_test.jl:57-65)."""
    N, NC, a = 24, 20, 2.5
    H = models.chain(N) / a
    psi = np.eye(N, dtype=complex)
    mu = oracle.kpm_1d(H, NC, psi).mean(axis=0).real
    lam = -2 * np.cos(2 * np.pi * np.arange(N) / N) / a
    exact = np.cos(np.arange(NC)[:, None] * np.arccos(lam)[None, :]).mean(axis=1)
    np.testing.assert_allclose(mu, exact, rtol=1e-10, atol=1e-12)


def test_tangent_vs_finite_differences():
    """The oracle's forward-mode tangent (used to pin the GPU adjoint) equals
    central differences of the oracle's own moments."""
    rng = Xoshiro256pp(31, lanes=4)
    H = (models.square_anderson(5, 2.0, seed=2) / 5.5).tocsc()
    H.sort_indices()
    psi = phases(rng, 25, 2)
    d = np.zeros(H.nnz, dtype=complex)
    rows = H.indices
    cols = np.repeat(np.arange(25), np.diff(H.indptr))
    zz = {}
    for k, (r, c) in enumerate(zip(rows, cols)):
        key = (min(r, c), max(r, c))
        if key not in zz:
            zz[key] = (rng.rand() - 0.5) + (1j * (rng.rand() - 0.5) if r != c else 0)
        d[k] = zz[key] if r <= c else np.conj(zz[key])
    NC = 20
    mu, dmu = oracle.kpm_1d_tangent(H, d, NC, psi)
    np.testing.assert_allclose(mu, oracle.kpm_1d(H, NC, psi), atol=1e-12)
    h = 1e-6
    Hp = H.copy(); Hp.data = H.data + h * d
    Hm = H.copy(); Hm.data = H.data - h * d
    fd = (oracle.kpm_1d(Hp, NC, psi) - oracle.kpm_1d(Hm, NC, psi)) / (2 * h)
    np.testing.assert_allclose(dmu, fd, atol=1e-7)


def test_ed_hall_and_broadened_kubo_bastin_pins():
    """ED conductivity references (test/ed_reference.jl:349-385,517-528) on the
    6 x 6 Haldane torus: the T = 0 Hall value is quantised to the Chern number
    C = +1 of phi = pi/2 within 0.02 (kubo_bastin_test.jl:38-39), the broadened
    form agrees within 1e-2 at eta = 0.05 (:43-44), the trivial phase m = 1.6
    gives its finite-size value near -0.11 (:73-81), and exchanging the two
    currents flips the sign (sigma_xy = -sigma_yx)."""
    from kpmhip import models
    H, Jx, Jy, area = models.haldane_model(6, 6, t=1.0, t2=0.2)
    s0 = oracle.ed_hall_conductivity_T0(H, Jx, Jy, area, Ef=0.0)
    assert abs(s0 - 1.0) < 0.02
    assert abs(oracle.ed_hall_conductivity_T0(H, Jy, Jx, area, Ef=0.0) + s0) < 1e-12
    sb = oracle.ed_kubo_bastin_broadened(H, Jx, Jy, area, Ef=0.0, eta=0.05)
    assert abs(sb - s0) < 1e-2
    Ht, Jxt, Jyt, at = models.haldane_model(6, 6, t=1.0, t2=0.2, m=1.6)
    st = oracle.ed_hall_conductivity_T0(Ht, Jxt, Jyt, at, Ef=0.0)
    assert abs(st) < 0.2 and abs(st + 0.11) < 0.02


def phases_nr(seed, n, nr):
    return phases(Xoshiro256pp(seed), n, nr)


def test_kpm2d_tangent_matches_table_and_finite_differences():
    """oracle.kpm_2d_tangent: its table equals the C restatement's kpm_2d and
    its tangent equals central differences of that kpm_2d along random
    complex directions of H, Ja and Jb (the checker of the kpm_2d adjoint)."""
    from kpmhip import models
    H, Jx, Jy, area = models.haldane_model(4, 4, t2=0.2)
    Hn = (H / 4.4).tocsc()
    NC, NR = 16, 3
    pl = phases_nr(1, H.shape[0], NR)
    pr = phases_nr(2, H.shape[0], NR)
    rng = Xoshiro256pp(9)

    def rnd(A):
        D = sp.csc_matrix(A, dtype=complex).copy()
        D.data = (rng.rand(D.nnz) - 0.5) + 1j * (rng.rand(D.nnz) - 0.5)
        return D
    dH, dJa, dJb = rnd(Hn), rnd(Jx), rnd(Jy)
    mu, dmu = oracle.kpm_2d_tangent(Hn, Jx, Jy, dH, dJa, dJb, NC, pl, pr)
    np.testing.assert_allclose(mu, oracle.kpm_2d(Hn, Jx, Jy, NC, pl, pr, NC, NC), rtol=0, atol=1e-13)
    h = 1e-6
    fd = (oracle.kpm_2d(Hn + h * dH, Jx + h * dJa, Jy + h * dJb, NC, pl, pr, NC, NC) -
          oracle.kpm_2d(Hn - h * dH, Jx - h * dJa, Jy - h * dJb, NC, pl, pr, NC, NC)) / (2 * h)
    assert np.max(np.abs(fd - dmu)) <= 1e-8 * np.max(np.abs(dmu))
    # each operator's part alone (a dropped term would show here)
    for d in ((dH, 0 * dJa, 0 * dJb), (0 * dH, dJa, 0 * dJb), (0 * dH, 0 * dJa, dJb)):
        _, t = oracle.kpm_2d_tangent(Hn, Jx, Jy, *d, NC, pl, pr)
        f = (oracle.kpm_2d(Hn + h * d[0], Jx + h * d[1], Jy + h * d[2], NC, pl, pr, NC, NC) -
             oracle.kpm_2d(Hn - h * d[0], Jx - h * d[1], Jy - h * d[2], NC, pl, pr, NC, NC)) / (2 * h)
        assert np.max(np.abs(f - t)) <= 1e-8 * max(np.max(np.abs(t)), 1e-12)


@pytest.mark.parametrize("Ef,beta", [(0.0, np.inf), (1.3, np.inf), (-0.7, 40.0)])
def test_kubo_bastin_weights_reproduce_the_head(Ef, beta):
    """sigma(mu) = Re sum K * mu with K = kpmhip.kubo_bastin_weights, against
    the oracle's restatement of kubo_bastin_cond on random complex tables."""
    import kpmhip as K
    NC, a, NH, area = 40, 4.2, 72, 37.5
    rng = Xoshiro256pp(31)
    W = K.kubo_bastin_weights(NC, a, Ef, b=0.1, NH=NH, area=area, beta=beta)
    for _ in range(3):
        mu = (rng.rand(NC, NC) - 0.5) + 1j * (rng.rand(NC, NC) - 0.5)
        ref = oracle.kubo_bastin_cond(mu, a, Ef, b=0.1, NH=NH, area=area, beta=beta)
        got = float(np.real(np.sum(W * mu)))
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
