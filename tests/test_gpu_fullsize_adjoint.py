"""The DOS-loss gradient at the benchmark's settings.

Model: periodic L x L square lattice, eps_i ~ U[-1, 1], t_ij ~ 1 + U[-0.1, 0.1]
(every site potential and bond hopping trainable), fixed rescaling a = 5.5,
NC = 1024, NR = 16, Jackson kernel; loss = sum over 256 energies of
(rho(E) - rho*(E))^2, rho* the DOS of a second disorder seed.

- L = 32: the per-site and per-bond gradients of the reverse Chebyshev
  recurrence against central finite differences (h = 1e-6) of the oracle's
  FP64 forward, on the 10 sites and the 10 bonds with the largest gradients,
  to 1e-5 relative.
- L = 512 (NH = 262,144): the gradient's directional derivative along a
  random direction over all sites and bonds against a central finite
  difference (h = 1e-6) of the device's FP64 forward, to 1e-5 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu

NC, NR, A = 1024, 16, 5.5
H_FD = 1e-6


def setup(L):
    H = models.square_disordered(L, seed=42)
    Ht = models.square_disordered(L, seed=43)
    psi = K.random_phase_vectors(Xoshiro256pp(42), L * L, NR)
    E = np.linspace(-0.9 * A, 0.9 * A, 256)
    x = E / A
    ns = np.arange(NC)
    w = K.JacksonKernel(ns, NC) * K.hn(ns)
    T = np.cos(ns[None, :] * np.arccos(x)[:, None])
    D = T * w[None, :] / (A * np.pi * np.sqrt(1 - x * x))[:, None]
    return H, Ht, psi, E, D


def loss_of(Hphys, psi, E, rho_t, forward):
    mu = forward((Hphys / A).tocsc(), psi)
    return K.dos_loss(mu, A, 0.0, E, rho_t)[0]


def gpu_forward(Hn, psi):
    return K.kpm_1d(Hn, NC, NR, psi_in=psi)


def oracle_forward(Hn, psi):
    return oracle.kpm_1d(Hn, NC, psi).mean(axis=0).real


def gradients(H, psi, E, rho_t):
    Hn = (H / A).tocsc()
    mu = gpu_forward(Hn, psi)
    loss, g = K.dos_loss(mu, A, 0.0, E, rho_t)
    grad = K.kpm_1d_adjoint(Hn, NC, psi, g)
    return loss, K.site_gradient(grad, A), K.bond_gradient(H, grad, A).tocoo()


def test_adjoint_l32_sites_and_bonds_vs_fd():
    L = 32
    H, Ht, psi, E, D = setup(L)
    rho_t = D @ oracle_forward((Ht / A).tocsc(), psi)
    loss, gs, gb = gradients(H, psi, E, rho_t)
    print(f"L={L} NC={NC} NR={NR}: loss {loss:.6e}", flush=True)
    n = L * L
    f = lambda Hp: loss_of(Hp, psi, E, rho_t, oracle_forward)  # noqa: E731
    for i in np.argsort(-np.abs(gs))[:10]:
        Ei = sp.csc_matrix(([1.0], ([i], [i])), shape=(n, n))
        fd = (f(H + H_FD * Ei) - f(H - H_FD * Ei)) / (2 * H_FD)
        print(f"  site {i}: adjoint {gs[i]:.10e} fd {fd:.10e}", flush=True)
        assert abs(gs[i] - fd) <= 1e-5 * abs(fd)
    for k in np.argsort(-np.abs(gb.data))[:10]:
        i, j = gb.row[k], gb.col[k]
        # t_ij enters as H_ij = H_ji = -t_ij
        Eb = sp.csc_matrix(([-1.0, -1.0], ([i, j], [j, i])), shape=(n, n))
        fd = (f(H + H_FD * Eb) - f(H - H_FD * Eb)) / (2 * H_FD)
        print(f"  bond ({i},{j}): adjoint {gb.data[k]:.10e} fd {fd:.10e}", flush=True)
        # bond_gradient is dL/dt for H_ij = t u, u the stored unit phase (here -1)
        assert abs(gb.data[k] - fd) <= 1e-5 * abs(fd)


def test_adjoint_512_directional_derivative_vs_fd():
    L = 512
    H, Ht, psi, E, D = setup(L)
    rho_t = D @ gpu_forward((Ht / A).tocsc(), psi)
    print(f"L={L} NH={L * L} NC={NC} NR={NR}: gradient step", flush=True)
    loss, gs, gb = gradients(H, psi, E, rho_t)
    n = L * L
    rng = Xoshiro256pp(77)
    de = rng.rand(n) - 0.5                      # site potentials
    dt = rng.rand(gb.nnz) - 0.5                 # bond amplitudes along each stored phase
    pred = float(np.dot(gs, de) + np.dot(gb.data, dt))
    u = np.sign(np.real(np.asarray(H[gb.row, gb.col]).ravel()))  # real bonds: phase = sign
    Dm = sp.csc_matrix((np.concatenate([de, u * dt, u * dt]),
                        (np.concatenate([np.arange(n), gb.row, gb.col]),
                         np.concatenate([np.arange(n), gb.col, gb.row]))), shape=(n, n))
    f = lambda Hp: loss_of(Hp, psi, E, rho_t, gpu_forward)  # noqa: E731
    fd = (f(H + H_FD * Dm) - f(H - H_FD * Dm)) / (2 * H_FD)
    print(f"  loss {loss:.6e}: directional adjoint {pred:.10e} fd {fd:.10e} "
          f"rel {abs(pred - fd) / abs(fd):.2e}", flush=True)
    assert abs(pred - fd) <= 1e-5 * abs(fd)
