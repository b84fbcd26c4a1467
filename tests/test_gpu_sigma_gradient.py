"""Reverse mode through kpm_2d and the Kubo-Bastin head (GPU).

- kpm_2d_adjoint against the oracle's forward-mode tangent of the kpm_2d
  table (oracle.kpm_2d_tangent, pinned by central differences of the oracle's
  kpm_2d on the CPU): Re sum g dA = Re sum conj(gbar) dmu along random
  complex directions of H, Ja and Jb, to 1e-10 relative, on the Haldane
  torus and on ragged operators with differing patterns, at several probe
  widths.
- The sigma_xy gradient of the Haldane model (sigma_rrule + hopping_gradient)
  against central finite differences (h = 1e-6) of sigma from the oracle's
  kpm_2d and Kubo-Bastin head, for site potentials and bond hoppings (which
  move H and both current operators), to 1e-5 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp
from matrices import ragged_hermitian

pytestmark = pytest.mark.gpu


def phases(seed, n, nr):
    return K.random_phase_vectors(Xoshiro256pp(seed), n, nr)


def random_on_pattern(A, rng):
    D = sp.csc_matrix(A, dtype=np.complex128).copy()
    D.sort_indices()
    D.data = (rng.rand(D.nnz) - 0.5) + 1j * (rng.rand(D.nnz) - 0.5)
    return D


def pair(G, D):
    """Re sum_k G_k D_k over a shared pattern."""
    G = sp.csc_matrix(G)
    G.sort_indices()
    D = sp.csc_matrix(D)
    D.sort_indices()
    assert np.array_equal(G.indices, D.indices) and np.array_equal(G.indptr, D.indptr)
    return float(np.real(np.sum(G.data * D.data)))


def check_vs_tangent(H, Ja, Jb, NC, pl, pr, seed):
    rng = Xoshiro256pp(seed)
    gbar = (rng.rand(NC, NC) - 0.5) + 1j * (rng.rand(NC, NC) - 0.5)
    gH, gJa, gJb = K.kpm_2d_adjoint(H, Ja, Jb, NC, pl, pr, gbar)
    for d in range(3):
        dH, dJa, dJb = (random_on_pattern(X, rng) for X in (H, Ja, Jb))
        _, dmu = oracle.kpm_2d_tangent(H, Ja, Jb, dH, dJa, dJb, NC, pl, pr)
        want = float(np.real(np.sum(np.conj(gbar) * dmu)))
        got = pair(gH, dH) + pair(gJa, dJa) + pair(gJb, dJb)
        print(f"  direction {d}: adjoint {got:.15e} tangent {want:.15e}", flush=True)
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want))


@pytest.mark.parametrize("NR", [1, 3, 8])
def test_kpm2d_adjoint_haldane_vs_tangent(NR):
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    Hn = (H / 4.4).tocsc()
    NC = 32
    print(f"kpm_2d adjoint Haldane 6x6 NC={NC} NR={NR}", flush=True)
    check_vs_tangent(Hn, Jx, Jy, NC, phases(10 + NR, H.shape[0], NR), phases(20 + NR, H.shape[0], NR),
                     seed=30 + NR)


def test_kpm2d_adjoint_ragged_vs_tangent():
    n, NC, NR = 90, 20, 4
    H = ragged_hermitian(n, 3, seed=5)
    Ja = ragged_hermitian(n, 5, seed=6)
    Jb = ragged_hermitian(n, 12, seed=7)
    print("kpm_2d adjoint ragged operators", flush=True)
    check_vs_tangent(H, Ja, Jb, NC, phases(8, n, NR), phases(9, n, NR), seed=11)
    # NC = 1 and 2: the Clenshaw sums of one and two orders
    for nc in (1, 2):
        print(f"kpm_2d adjoint NC={nc}", flush=True)
        check_vs_tangent(H, Ja, Jb, nc, phases(8, n, NR), phases(9, n, NR), seed=12 + nc)


def sigma_oracle(H, Jx, Jy, a, Ef, NC, psi, area):
    mu = oracle.kpm_2d((H / a).tocsc(), Jx, Jy, NC, psi, psi, NC, NC)
    return oracle.kubo_bastin_cond(mu, a, Ef, NH=H.shape[0], area=area)


def test_sigma_xy_gradient_vs_finite_differences():
    Lc, t2, a, NC, NR, Ef = 6, 0.2, 4.4, 64, 4, 1.5   # Ef inside the upper band
    H, Jx, Jy, area = models.haldane_model(Lc, Lc, t2=t2)
    rng = Xoshiro256pp(5)
    # site disorder so that the sites differ
    eps = 0.3 * (rng.rand(H.shape[0]) - 0.5)
    H = (H + sp.diags(eps)).tocsc()
    H.sort_indices()
    psi = phases(3, H.shape[0], NR)
    Hn = (H / a).tocsc()
    sigma, pullback = K.sigma_rrule(Hn, Jx, Jy, NC, psi, a, Ef, area=area)
    mu = K.kpm_2d(Hn, Jx, Jy, NC, NR, psi_in=psi)
    assert abs(sigma - K.kubo_bastin_cond_raw(mu, a, Ef, NH=H.shape[0], area=area)) <= 1e-12
    assert abs(sigma - sigma_oracle(H, Jx, Jy, a, Ef, NC, psi, area)) <= 1e-10
    gH, gJx, gJy = pullback(1.0)
    site, bonds = K.hopping_gradient(H, Jx, Jy, gH, gJx, gJy, a)
    print(f"Haldane {Lc}x{Lc} sigma_xy(Ef={Ef}) = {sigma:.10f}", flush=True)
    h = 1e-6
    n = H.shape[0]
    for i in np.argsort(-np.abs(site))[:5]:
        E = sp.csc_matrix(([1.0], ([i], [i])), shape=(n, n))
        fd = (sigma_oracle(H + h * E, Jx, Jy, a, Ef, NC, psi, area) -
              sigma_oracle(H - h * E, Jx, Jy, a, Ef, NC, psi, area)) / (2 * h)
        print(f"  site {i}: adjoint {site[i]:.10e} fd {fd:.10e}", flush=True)
        assert abs(site[i] - fd) <= 1e-5 * abs(fd)
    Hc = sp.csc_matrix(H)
    for k in np.argsort(-np.abs(bonds.data))[:6]:
        i, j = int(bonds.row[k]), int(bonds.col[k])
        mag = abs(Hc[i, j])
        # t_b -> t_b + h moves H_ij, H_ji and the current entries in proportion
        dH = sp.csc_matrix(([Hc[i, j] / mag, Hc[j, i] / mag], ([i, j], [j, i])), shape=(n, n))
        dX = sp.csc_matrix(([Jx[i, j] / mag, Jx[j, i] / mag], ([i, j], [j, i])), shape=(n, n))
        dY = sp.csc_matrix(([Jy[i, j] / mag, Jy[j, i] / mag], ([i, j], [j, i])), shape=(n, n))
        fd = (sigma_oracle(H + h * dH, Jx + h * dX, Jy + h * dY, a, Ef, NC, psi, area) -
              sigma_oracle(H - h * dH, Jx - h * dX, Jy - h * dY, a, Ef, NC, psi, area)) / (2 * h)
        print(f"  bond ({i},{j}) |t|={mag:.2f}: adjoint {bonds.data[k]:.10e} fd {fd:.10e}", flush=True)
        assert abs(bonds.data[k] - fd) <= 1e-5 * abs(fd)


def test_kpm2d_adjoint_bitwise_repeatable_wide_rows():
    """NR = 65: rows of two 64-word chunks, one gradient accumulator per
    chunk summed in a fixed order, so two runs agree bitwise."""
    H, Jx, Jy, area = models.haldane_model(6, 6, t2=0.2)
    Hn = (H / 4.4).tocsc()
    NC, NR = 16, 65
    pl, pr = phases(40, H.shape[0], NR), phases(41, H.shape[0], NR)
    gbar = (Xoshiro256pp(42).rand(NC, NC) - 0.5) + 1j * (Xoshiro256pp(43).rand(NC, NC) - 0.5)
    print(f"kpm_2d adjoint repeatability NR={NR}", flush=True)
    a = K.kpm_2d_adjoint(Hn, Jx, Jy, NC, pl, pr, gbar)
    b = K.kpm_2d_adjoint(Hn, Jx, Jy, NC, pl, pr, gbar)
    for x, y in zip(a, b):
        assert np.array_equal(sp.csc_matrix(x).data, sp.csc_matrix(y).data)
