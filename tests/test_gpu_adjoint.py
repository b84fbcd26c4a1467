"""The reverse-mode gradient of the moments (GPU) against the forward-mode
tangent of the oracle and central finite differences (h = 1e-6)."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp
from matrices import BUCKETS, ragged_hermitian

pytestmark = pytest.mark.gpu


def hermitian_direction(H, seed):
    """A random Hermitian perturbation on H's (symmetric) sparsity pattern,
    as values in H's sorted CSC order; returns (A, dH_data)."""
    A = sp.csc_matrix(H, dtype=np.complex128)
    A.sort_indices()
    rng = Xoshiro256pp(seed)
    rows = A.indices
    cols = np.repeat(np.arange(A.shape[1]), np.diff(A.indptr))
    z = (rng.rand(A.nnz) - 0.5) + 1j * (rng.rand(A.nnz) - 0.5)
    val = {}
    for k, (r, c) in enumerate(zip(rows, cols)):
        if r < c:
            val[(r, c)] = z[k]
        elif r == c:
            val[(r, c)] = z[k].real
    out = np.zeros(A.nnz, dtype=np.complex128)
    for k, (r, c) in enumerate(zip(rows, cols)):
        out[k] = val[(r, c)] if r <= c else np.conj(val[(c, r)])
    return A, out


def model(L=8, seed=3):
    H = models.square_anderson(L, 2.0, seed=seed)
    Hc = sp.csc_matrix(H, dtype=np.complex128)
    # complex Peierls-like phases on the bonds keep it Hermitian but not real
    Hc = Hc.tolil()
    rng = Xoshiro256pp(seed + 1)
    n = Hc.shape[0]
    for i in range(0, n, 3):
        j = (i + 1) % n
        if Hc[i, j] != 0:
            ph = np.exp(1j * 2 * np.pi * rng.rand())
            Hc[i, j] = Hc[i, j] * ph
            Hc[j, i] = np.conj(Hc[i, j])
    return (sp.csc_matrix(Hc) / 5.5).tocsc()


@pytest.mark.parametrize("ckpt", [0, 1, 5, 64])
def test_adjoint_vs_tangent(ckpt):
    H = model()
    NC, NR = 48, 3
    psi = K.random_phase_vectors(Xoshiro256pp(7), H.shape[0], NR)
    g = Xoshiro256pp(8).rand(NC) - 0.5
    grad = K.kpm_1d_adjoint(H, NC, psi, g, ckpt_stride=ckpt)
    A, dH = hermitian_direction(H, 11)
    G = sp.csc_matrix(grad)
    G.sort_indices()
    pred = np.real(np.sum(G.data * dH))
    mu, dmu = oracle.kpm_1d_tangent(A, dH, NC, psi)
    tang = float(np.sum(g * np.real(dmu.mean(axis=0))))
    print(f"ckpt={ckpt} adjoint {pred:.15e} tangent {tang:.15e}", flush=True)
    assert abs(pred - tang) <= 1e-10 * max(1.0, abs(tang))


def test_adjoint_vs_finite_differences_dos_loss():
    H = model(L=10, seed=5)
    NC, NR = 64, 4
    psi = K.random_phase_vectors(Xoshiro256pp(21), H.shape[0], NR)
    a, b = 1.0, 0.0
    E = np.linspace(-0.8, 0.8, 32)
    target = np.full(len(E), 0.3)
    mu = oracle.kpm_1d(H, NC, psi).mean(axis=0).real
    L0, g = K.dos_loss(mu, a, b, E, target)
    grad = sp.csc_matrix(K.kpm_1d_adjoint(H, NC, psi, g))
    grad.sort_indices()
    A, dH = hermitian_direction(H, 12)
    pred = np.real(np.sum(grad.data * dH))
    h = 1e-6
    def loss(Hm):
        m = oracle.kpm_1d(Hm, NC, psi).mean(axis=0).real
        return K.dos_loss(m, a, b, E, target)[0]
    Hp = A.copy(); Hp.data = A.data + h * dH
    Hm = A.copy(); Hm.data = A.data - h * dH
    fd = (loss(Hp) - loss(Hm)) / (2 * h)
    print(f"dos loss {L0:.6e}: adjoint {pred:.10e} fd {fd:.10e}", flush=True)
    assert abs(pred - fd) <= 1e-5 * max(abs(fd), 1e-12)


def test_site_and_bond_gradients_vs_fd():
    """per-site potential and per-bond hopping gradients of the physical H
    through the fixed rescaling H_norm = (H - b)/a."""
    L, a, b = 6, 5.0, 0.1
    Hphys = models.square_anderson(L, 2.0, seed=9)
    Hn = ((Hphys - b * sp.identity(L * L)) / a).tocsc()
    NC, NR = 40, 2
    psi = K.random_phase_vectors(Xoshiro256pp(4), L * L, NR)
    g = Xoshiro256pp(5).rand(NC) - 0.5
    grad = K.kpm_1d_adjoint(Hn, NC, psi, g)
    gs = K.site_gradient(grad, a)
    gb = K.bond_gradient(Hn, grad, a).tocoo()
    def L_of(Hp):
        Hnp = ((Hp - b * sp.identity(L * L)) / a).tocsc()
        return float(np.sum(g * oracle.kpm_1d(Hnp, NC, psi).mean(axis=0).real))
    h = 1e-6
    for i in (0, 7, 20):
        E = sp.csc_matrix(([1.0], ([i], [i])), shape=Hphys.shape)
        fd = (L_of(Hphys + h * E) - L_of(Hphys - h * E)) / (2 * h)
        assert abs(gs[i] - fd) <= 1e-5 * max(abs(fd), 1e-9), (i, gs[i], fd)
    for k in (0, 3, 11):
        i, j = gb.row[k], gb.col[k]
        E = sp.csc_matrix(([1.0, 1.0], ([i, j], [j, i])), shape=Hphys.shape)
        u = np.sign(np.real(Hphys[i, j]))  # real hoppings: unit phase = sign
        fd = (L_of(Hphys + h * u * E) - L_of(Hphys - h * u * E)) / (2 * h)
        assert abs(gb.data[k] - fd) <= 1e-5 * max(abs(fd), 1e-9), (k, gb.data[k], fd)


@pytest.mark.parametrize("w", sorted(BUCKETS))
@pytest.mark.parametrize("NR", [1, 2, 8, 16, 64, 65])
def test_adjoint_ragged_rows_vs_tangent(w, NR):
    """Ragged rows (empty up to 41 nonzeros) across row widths of 1 to 64
    lanes and a two-chunk row (NR = 65): every shape of the gradient
    accumulation, checked along two random Hermitian directions."""
    H = ragged_hermitian(300, w, seed=40 + w)
    NC = 24
    print(f"adjoint ragged w={w} NR={NR} max_row={np.diff(H.tocsr().indptr).max()}", flush=True)
    psi = K.random_phase_vectors(Xoshiro256pp(60 + NR), H.shape[0], NR)
    g = Xoshiro256pp(61 + w).rand(NC) - 0.5
    grad = K.kpm_1d_adjoint(H, NC, psi, g)
    G = sp.csc_matrix(grad)
    G.sort_indices()
    for seed in (5, 6):
        A, dH = hermitian_direction(H, seed)
        pred = np.real(np.sum(G.data * dH))
        _, dmu = oracle.kpm_1d_tangent(A, dH, NC, psi)
        tang = float(np.sum(g * np.real(dmu.mean(axis=0))))
        assert abs(pred - tang) <= 1e-10 * max(1.0, abs(tang)), (seed, pred, tang)


@pytest.mark.parametrize("ckpt", [0, 1, 5])
@pytest.mark.parametrize("NR", [3, 65])
def test_tape_moments_and_pullback(ckpt, NR):
    """The split adjoint: the forward half's moments are the oracle's
    per-probe moments <psi_r|T_n(H)|psi_r> of every order (to 1e-12), and its
    pullback is bitwise the one-call kpm_1d_adjoint (same passes, same
    summation order)."""
    H = model(L=6, seed=8)
    NC = 30
    psi = K.random_phase_vectors(Xoshiro256pp(70 + NR), H.shape[0], NR)
    print(f"adjoint tape NR={NR} ckpt={ckpt}", flush=True)
    tape = K.kpm_1d_adjoint_forward(H, NC, psi, ckpt_stride=ckpt)
    want = oracle.kpm_1d(H, NC, psi)
    assert tape.mu_all.shape == (NR, NC)
    assert np.max(np.abs(tape.mu_all - want)) <= 1e-12
    np.testing.assert_allclose(tape.mu, K.kpm_1d(H, NC, NR, psi_in=psi), rtol=0, atol=1e-12)
    g = Xoshiro256pp(9).rand(NC) - 0.5
    G1 = tape.pullback(g)
    G2 = K.kpm_1d_adjoint(H, NC, psi, g, ckpt_stride=ckpt)
    assert np.array_equal(G1.data, G2.data)
    # the one-call adjoint replaced the kept pass: the old tape refuses
    with pytest.raises(K.ArgumentError):
        tape.pullback(g)


def test_tape_pullback_is_repeatable_and_checks_operator():
    H = model(L=6, seed=8)
    NC, NR = 20, 4
    psi = K.random_phase_vectors(Xoshiro256pp(3), H.shape[0], NR)
    mu, pb = K.moments_rrule(H, NC, psi)
    g1 = Xoshiro256pp(1).rand(NC) - 0.5
    g2 = Xoshiro256pp(2).rand(NC) - 0.5
    a1, b, a2 = pb(g1), pb(g2), pb(g1)  # several cotangents on one forward pass
    assert np.array_equal(a1.data, a2.data)
    # linear in g
    ab = pb(g1 + g2)
    np.testing.assert_allclose(ab.data, a1.data + b.data, rtol=0, atol=1e-13)


@pytest.mark.parametrize("NR", [65, 128])
def test_adjoint_bitwise_repeatable_wide_rows(NR):
    """Rows of two and four 64-word chunks: per-chunk gradient accumulators
    summed in a fixed order, so two runs agree bitwise."""
    H = ragged_hermitian(300, 9, seed=49)
    NC = 24
    psi = K.random_phase_vectors(Xoshiro256pp(80 + NR), H.shape[0], NR)
    g = Xoshiro256pp(81).rand(NC) - 0.5
    print(f"adjoint repeatability NR={NR}", flush=True)
    G1 = K.kpm_1d_adjoint(H, NC, psi, g)
    G2 = K.kpm_1d_adjoint(H, NC, psi, g)
    assert np.array_equal(G1.data, G2.data)


def test_tape_with_resident_operator():
    """A DeviceCSR passed as operator gives bitwise the gradient of a fresh
    upload; a mismatched one is refused."""
    H = model(L=6, seed=8)
    NC, NR = 20, 4
    psi = K.random_phase_vectors(Xoshiro256pp(3), H.shape[0], NR)
    g = Xoshiro256pp(1).rand(NC) - 0.5
    Hd = K.to_device(H)
    a = K.kpm_1d_adjoint_forward(H, NC, psi, operator=Hd).pullback(g)
    b = K.kpm_1d_adjoint(H, NC, psi, g)
    assert np.array_equal(a.data, b.data)
    with pytest.raises(K.ArgumentError):
        K.kpm_1d_adjoint_forward(H, NC, psi, operator=K.to_device(model(L=5, seed=8)))
