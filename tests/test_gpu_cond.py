"""Conductivity tables and their reconstruction on the device.

- kpm_2d's moment_parity selects the documented index parity
  (test/cond_test.jl:92-131). On a random Hermitian H with dense random
  currents and identity probes (an exact trace), :ODD keeps mu_nm with
  (n + m) odd (0-based) and :EVEN the complement. The kept entries agree
  with :NONE at rtol 1e-12, the rest are exactly zero, and an unknown parity
  is an ArgumentError.
- The host Kubo–Bastin integrand d_dc_cond (Γ contraction per energy,
  conductivity.jl:199-245), integrated over the Gauss–Chebyshev nodes, equals
  kubo_bastin_cond (conductivity.jl:101-157). That one runs the factorised
  node contraction on the device, so two independent implementations of the
  same Σ_nm Γnm(x) μ̃nm are tied together.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def rand_c(rng, *shape):
    n = int(np.prod(shape))
    return ((rng.rand(n) - 0.5) + 1j * (rng.rand(n) - 0.5)).reshape(shape)


def test_kpm2d_moment_parity_selects_index_parity():
    rng = Xoshiro256pp(31)
    D, NC = 10, 8
    A = rand_c(rng, D, D)
    Hd = (A + A.conj().T) / 2
    Hd /= 2 * np.max(np.abs(np.linalg.eigvalsh(Hd)))
    H = sp.csc_matrix(Hd)
    Ja = sp.csc_matrix(rand_c(rng, D, D))
    Jb = sp.csc_matrix(rand_c(rng, D, D))
    psi = np.eye(D, dtype=np.complex128)  # exact trace
    print("kpm_2d parity masks D=10 NC=8", flush=True)
    mu_none = K.kpm_2d(H, Ja, Jb, NC, D, D, psi_in=psi, moment_parity="NONE")
    mu_odd = K.kpm_2d(H, Ja, Jb, NC, D, D, psi_in=psi, moment_parity="ODD")
    mu_even = K.kpm_2d(H, Ja, Jb, NC, D, D, psi_in=psi, moment_parity="EVEN")
    n = np.arange(NC)
    odd = ((n[:, None] + n[None, :]) % 2) == 1
    np.testing.assert_allclose(mu_odd, mu_none * odd, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(mu_even, mu_none * ~odd, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(mu_odd + mu_even, mu_none, rtol=1e-12, atol=1e-14)
    assert np.all(mu_odd[~odd] == 0) and np.all(mu_even[odd] == 0)
    assert np.count_nonzero(mu_odd) == NC * NC // 2
    with pytest.raises(K.ArgumentError):
        K.kpm_2d(H, Ja, Jb, NC, D, D, psi_in=psi, moment_parity="SOMETHING")


@pytest.mark.parametrize("Ef,beta", [(0.0, np.inf), (0.3, 40.0), (-0.5, np.inf)])
def test_kubo_bastin_equals_integrated_d_dc_cond(Ef, beta):
    rng = Xoshiro256pp(77)
    NC, a, b, NH, area = 40, 2.2, 0.1, 500, 37.0
    mu = rand_c(rng, NC, NC)
    quad_N, cut = 8 * NC, 1e-3
    got = K.kubo_bastin_cond_raw(mu, a, Ef, b=b, NH=NH, area=area, beta=beta, quad_N=quad_N,
                                 edge_cutoff=cut)
    x, w = K.gausschebyshevt(quad_N)
    f = K.fermiFunction(a * x + b, Ef, beta)
    keep = (np.abs(x) < 1 - cut) & (f != 0)
    ds = K.d_dc_cond(mu, a, a * x[keep] + b, b=b, NC=NC)
    ref = -2 * NH / area * np.sum(w[keep] * f[keep] * ds * np.sqrt(1 - x[keep] ** 2))
    print(f"kubo_bastin Ef={Ef} beta={beta}: device {got:.15e} host {ref:.15e}", flush=True)
    assert abs(got - ref) <= 1e-10 * max(1.0, abs(ref))
