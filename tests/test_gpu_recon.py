"""Device reconstruction heads (FP64) against the oracle's host restatement:
chebyshev_lin_trans / dos (src/applications/dos.jl:76-149) and the
Kubo–Bastin node contraction (src/applications/conductivity.jl:101-157).
Bound: 1e-12 relative (only the summation order differs)."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def test_lin_trans_vs_oracle():
    rng = np.random.default_rng(3)
    for Nx, Nn in ((1, 1), (5, 63), (257, 64), (1000, 1537)):
        print(f"lin_trans Nx={Nx} Nn={Nn}", flush=True)
        x = np.concatenate([[-1.0, 1.0, 0.0], rng.uniform(-1, 1, Nx)])
        n = rng.permutation(np.arange(3 * Nn))[:Nn]  # an arbitrary integer grid
        mu = rng.standard_normal(Nn) + 1j * rng.standard_normal(Nn)
        got = K.chebyshev_lin_trans(x, n, mu)
        ref = oracle.lin_trans(x, n, mu)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * max(1.0, np.abs(mu).sum()))
    muR = rng.standard_normal(40)
    y = K.chebyshev_lin_trans(0.3, np.arange(40), muR)
    assert np.isscalar(y) or np.ndim(y) == 0
    assert abs(y - oracle.lin_trans([0.3], np.arange(40), muR)[0]) < 1e-12
    with pytest.raises(K.ArgumentError):
        K.chebyshev_lin_trans([1.5], np.arange(4), np.ones(4))


def test_dos_vs_oracle_chain():
    N, t, shift = 4000, 1.0, 0.7
    H = models.chain(N, t=t, shift=shift)
    a, b, Hn = K.normalizeH(H, center=True)
    NC, NR = 256, 24
    psi = K.random_phase_vectors(Xoshiro256pp(7), N, NR)
    mu = K.kpm_1d(Hn, NC, NR, psi_in=psi)
    E, rho = K.dos_raw(mu, a, b=b, N_tilde=1024)
    E2, rho2 = oracle.dos(mu, a, b=b, N_tilde=1024)
    np.testing.assert_allclose(E, E2, rtol=0, atol=1e-15)
    np.testing.assert_allclose(rho, rho2, rtol=1e-12, atol=1e-13)
    # points outside the band stay zero, in-band grid may be empty
    _, r = K.dos_raw(mu, a, b=b, E_grid=[b + 2 * a, b - 3 * a])
    assert np.all(r == 0)


@pytest.mark.parametrize("NC,quad_N,beta,Ef", [(7, 16, np.inf, 0.2), (37, 300, np.inf, -0.1),
                                               (64, 512, 50.0, 0.0), (130, 1040, np.inf, 0.35)])
def test_kubo_bastin_vs_oracle_random_tables(NC, quad_N, beta, Ef):
    print(f"kubo NC={NC} quad_N={quad_N} beta={beta} Ef={Ef}", flush=True)
    rng = np.random.default_rng(NC)
    mu = rng.standard_normal((NC, NC)) + 1j * rng.standard_normal((NC, NC))
    kw = dict(b=0.05, NH=100, area=7.0, beta=beta, quad_N=quad_N)
    got = K.kubo_bastin_cond_raw(mu, 1.7, Ef, **kw)
    ref = oracle.kubo_bastin_cond(mu, 1.7, Ef, **kw)
    assert got == pytest.approx(ref, rel=1e-11, abs=1e-13)


def test_kubo_bastin_haldane_quantised_and_empty():
    """kubo_bastin_test.jl:19-56: sigma_xy = C = +1 from identity-probe moments."""
    H, Jx, Jy, area = models.haldane_model(6, 6, t=1.0, t2=0.2)
    D = H.shape[0]
    a, b, Hn = K.normalizeH(H, center=True)
    NC = 96
    psi = np.eye(D, dtype=complex)
    mu = oracle.kpm_2d(Hn, Jx, Jy, NC, psi, psi, NC, NC)
    s = K.kubo_bastin_cond_raw(mu, a, 0.0, b=b, NH=D, area=area)
    s_ref = oracle.kubo_bastin_cond(mu, a, 0.0, b=b, NH=D, area=area)
    assert s == pytest.approx(s_ref, rel=1e-12)
    assert s == pytest.approx(1.0, abs=0.05)
    # Fermi level below the band: no node carries weight
    assert K.kubo_bastin_cond_raw(mu, a, b - 2 * a, b=b, NH=D, area=area) == 0.0
