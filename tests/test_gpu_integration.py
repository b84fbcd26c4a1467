"""End-to-end pipeline on a model whose spectrum is known analytically,
restating test/integration_test.jl on the device: a dimerised (SSH) ring of
128 cells with hoppings t1 = 1, t2 = 0.4 and bond currents
(J_x)_ij = H_ij (r_i - r_j)_x.  Bands ±sqrt(t1² + t2² + 2 t1 t2 cos k): gap
2|t1 - t2| = 1.2 at E = 0 and bandwidth 2(t1 + t2) = 2.8.

- rescale: the spectral-edge estimate of normalizeH gives a within 2e-3 of
  2 Emax/(2 - 0.1); identity probes then give mu_n = Tr T_n(H_norm)/D from the
  dense spectrum at rtol 1e-10, with mu_0 = 1 (integration_test.jl:46-66);
- DOS: rho(0) vanishes in the gap (< 1e-4), its second derivative at 0 is
  finite, and the KPM counting function follows the ED one within 2% of the
  states in the bands (integration_test.jl:68-96).  The probes come from this
  build's generator, not Julia's Xoshiro(2718) stream;
- kpm_3d and cpge run end to end to finite values (integration_test.jl:98-118).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu

T1, T2, NCELL = 1.0, 0.4, 128
NH = 2 * NCELL
GAP = 2 * abs(T1 - T2)
BANDWIDTH = 2 * (T1 + T2)


def ssh_ring():
    rows, cols, vals, dvals = [], [], [], []

    def bond(i, j, amp, d):
        rows.extend([i, j])
        cols.extend([j, i])
        vals.extend([amp, np.conj(amp)])
        dvals.extend([amp * d, -np.conj(amp) * d])

    for n in range(NCELL):
        bond(2 * n + 1, 2 * n, -T1 + 0j, 0.5)                  # A(n) -> B(n)
        bond((2 * n + 2) % NH, 2 * n + 1, -T2 + 0j, 0.5)       # B(n) -> A(n+1)
    H = sp.csc_matrix((vals, (rows, cols)), shape=(NH, NH))
    Jx = sp.csc_matrix((dvals, (rows, cols)), shape=(NH, NH))
    return H, Jx


def test_rescaling_and_identity_probe_moments():
    H, _ = ssh_ring()
    Hd = H.toarray()
    assert np.allclose(Hd, Hd.conj().T)
    lam = np.linalg.eigvalsh(Hd)
    assert abs(lam.max() - BANDWIDTH / 2) <= 1e-10 * BANDWIDTH / 2
    assert abs(np.min(np.abs(lam)) - GAP / 2) <= 1e-10 * GAP / 2
    a, Hn = K.normalizeH(H)
    assert abs(a - 2 * (BANDWIDTH / 2) / (2 - 0.1)) <= 2e-3 * a
    NC = 64
    psi = np.eye(NH, dtype=np.complex128)  # exact trace
    print("SSH identity-probe kpm_1d NH=256 NC=64", flush=True)
    mu = K.kpm_1d(Hn, NC, NH, psi_in=psi)
    ln = np.clip(np.linalg.eigvalsh(Hn.toarray()), -1, 1)
    mu_ed = np.array([np.sum(np.cos(n * np.arccos(ln))) / NH for n in range(NC)])
    np.testing.assert_allclose(mu, mu_ed, rtol=1e-10, atol=1e-13)
    assert abs(mu[0] - 1.0) <= 1e-12


def test_dos_gap_and_counting_function():
    H, _ = ssh_ring()
    a, Hn = K.normalizeH(H)
    NC, NR = 2048, 21
    psi = K.random_phase_vectors(Xoshiro256pp(2718), NH, NR)
    print("SSH DOS kpm_1d NC=2048 NR=21", flush=True)
    mu = K.kpm_1d(Hn, NC, NR, psi_in=psi)
    assert abs(K.dos0_raw(mu, a)) < 1e-4
    assert np.isfinite(K.dos0_raw(mu, a, dE_order=2))
    E, rho = K.dos_raw(mu, a, N_tilde=4096)
    dE = E[1] - E[0]
    N_kpm = np.cumsum(rho) * dE
    lam = np.linalg.eigvalsh(H.toarray())
    N_ed = np.array([np.count_nonzero(lam <= e) / NH for e in E])
    inband = (np.abs(np.abs(E) - GAP / 2) > 0.15) & (np.abs(E) < 0.9 * BANDWIDTH / 2)
    assert np.max(np.abs(N_kpm[inband] - N_ed[inband])) < 0.02
    assert abs(N_kpm[-1] - 1.0) <= 0.02


def test_third_order_moments_and_cpge_end_to_end():
    H, Jx = ssh_ring()
    a, Hn = K.normalizeH(H)
    NC, NR = 16, 2
    psi = K.random_phase_vectors(Xoshiro256pp(31415), NH, NR)
    print("SSH kpm_3d NC=16 NR=2 + cpge", flush=True)
    Gamma = np.zeros((NC, NC, NC), dtype=np.complex128, order="F")
    K.kpm_3d_(Hn, Jx, Jx, Jx, NC, NR, NH, Gamma, psi, psi)
    assert np.all(np.isfinite(Gamma))
    val = K.cpge(Gamma, NC, 0.1, 0.05, lambda_=0.02)
    assert np.isfinite(val)
