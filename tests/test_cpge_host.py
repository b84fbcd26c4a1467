"""CPU pins of the three-current (CPGE) reconstruction (no GPU).

The oracle side:
- cpge_forms against a brute-force triple sum;
- the eigenbasis reference against the paper-tensor formula
  i sum Gamma kernel3 Lambda (test/cpge_test.jl:409-449).

The product's host logic:
- the 15-point Gauss-Kronrod rule (Legendre nodes, polynomial exactness);
- the resolvent and delta coefficient recurrences against closed forms
  (test/cpge_test.jl:67-96);
- the adaptive spectral integral;
- cpge's input validation (test/cpge_test.jl:451-475,602-607).

End to end on the CPU, the product quadrature and coefficient assembly,
driven by the oracle's contraction, reproduce the eigenbasis reference
(test/cpge_test.jl:345-407) to the reference's rtol 1e-8.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
from kpmhip.cpge import _check_cpge_inputs, cpge_exclusion_breakpoints, cpge_node_coefficients
from kpmhip.kernels import JacksonKernel, mu3D_apply_kernel_and_h
from kpmhip.spectral import (chebyshev_delta_theta, gauss_kronrod_15, green_coefficients,
                             spectral_integral)
import oracle as O


def helix():
    H, J, _ = models.helix_cluster()
    ev = np.linalg.eigvalsh(H.toarray())
    b = (ev.max() + ev.min()) / 2
    a = (ev.max() - ev.min()) / (2 - 0.1)
    Hn = sp.csc_matrix((H.toarray() - b * np.eye(6)) / a)
    return Hn, J


def helix_mu3(NC, Ja=0, Jb=1, Jg=2):
    Hn, J = helix()
    psi = np.eye(6, dtype=np.complex128)
    return O.kpm_3d(Hn, J[Ja], J[Jb], J[Jg], NC, psi, psi)


def cpu_node_fn(mu3, NC, w1, w2, lam, delta=0.0):
    mt = mu3D_apply_kernel_and_h(mu3[:NC, :NC, :NC], NC, JacksonKernel)

    def F(theta):
        a, b, c = cpge_node_coefficients(theta, NC, w1, w2, lam, delta)
        return O.cpge_forms(mt, a, b, c).reshape(-1, 3).sum(axis=1)[:, None]

    return F


def cpu_cpge(mu3, NC, w1, w2, *, E_f, beta, lam, delta=0.0, rtol=1e-8, atol=2e-11):
    _check_cpge_inputs(mu3, NC, w1, w2, lam, delta)
    Om = w1 + w2
    shifts = (Om, w2, w1, -w2, -w1, -Om)
    I, _ = spectral_integral(cpu_node_fn(mu3, NC, w1, w2, lam, delta), 1, NC, shifts, E_f, beta, lam,
                             rtol=rtol, atol=atol, breakpoints=cpge_exclusion_breakpoints(shifts, delta))
    return 1j * I[0]


def test_gauss_kronrod_rule():
    x, wk, wg = gauss_kronrod_15()
    xg, wgl = np.polynomial.legendre.leggauss(7)
    np.testing.assert_allclose(x[wg > 0], xg, atol=1e-15)
    np.testing.assert_allclose(wg[wg > 0], wgl, atol=1e-15)
    for d in range(24):  # Kronrod rule exact to degree 3*7+2 = 23
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(wk, x ** d) - exact) < 1e-14
    for d in range(14):  # Gauss part exact to degree 13
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert abs(np.dot(wg, x ** d) - exact) < 1e-14
    assert abs(np.dot(wk, x ** 24) - 2.0 / 25) > 1e-9  # and no further


def test_coefficients_vs_closed_forms():
    NC = 40
    th = np.array([0.1, 1.0, 2.5, 3.0])
    np.testing.assert_allclose(chebyshev_delta_theta(th, NC),
                               np.cos(np.outer(th, np.arange(NC))) / math.pi, atol=1e-13)
    for lam in (0.0, 0.03, 0.5):
        for br in ("R", "A"):
            zs = [-0.97, -0.3, 0.0, 0.41, 0.99] + ([-2.5, 1.3, 3.0] if lam == 0 else [1.0, -1.0, 1.7])
            got = green_coefficients(np.array(zs), NC, lam, br)
            for k, z in enumerate(zs):
                ref = O.cpge_ref_green(z, NC, lam, br)
                np.testing.assert_allclose(got[k], ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    with pytest.raises(K.DomainError):
        green_coefficients(np.array([1.0]), 4, 0.0, "R")
    with pytest.raises(K.ArgumentError):
        green_coefficients(np.array([0.2]), 4, -0.1, "R")


def test_spectral_integral_closed_forms():
    # int_0^pi cos(3 theta)^2 f dtheta over the occupied part; beta = inf,
    # xF = cos(theta_F): exact (pi - theta_F)/2 - sin(6 theta_F)/12
    xF = 0.23
    tF = math.acos(xF)
    F = lambda th: (np.cos(3 * th) ** 2)[:, None].astype(complex)
    I, E = spectral_integral(F, 1, 8, (), xF, math.inf, 0.0, rtol=1e-12)
    exact = (math.pi - tF) / 2 - math.sin(6 * tF) / 12
    assert abs(I[0] - exact) < 1e-11 and E < 1e-11
    # integrable inverse-square-root edge singularity at a shifted Chebyshev
    # edge: 1/sqrt(|cos theta - x*|) with x* = 1 - s, regularised by the
    # endpoint map; scipy's QUADPACK with the same break point as the check
    from scipy.integrate import quad
    s = 0.4
    xs = 1 - s
    g = lambda th: 1.0 / np.sqrt(np.abs(np.cos(th) - xs))
    I, _ = spectral_integral(lambda th: g(th)[:, None].astype(complex), 1, 8, (s,), 0.5, 7.0, 0.0,
                             rtol=1e-10)
    fermi = lambda x: 1 / (1 + math.exp(7.0 * (x - 0.5)))
    ts = math.acos(xs)
    ref = sum(quad(lambda t: g(t) * fermi(math.cos(t)), a, b, epsabs=1e-13, epsrel=1e-12, limit=200)[0]
              for a, b in ((0.0, ts), (ts, math.pi)))
    assert abs(I[0] - ref) < 1e-9 * abs(ref)
    # vector-valued integrand: every component to its own relative tolerance
    Fv = lambda th: np.stack([np.cos(th) ** 2, 1e-6 * np.sin(2 * th) ** 2], axis=1).astype(complex)
    I, _ = spectral_integral(Fv, 2, 8, (), 0.0, 30.0, 0.0, rtol=1e-10)
    ref = [quad(lambda t, k=k: Fv(np.array([t]))[0, k].real / (1 + math.exp(30 * math.cos(t))),
                0, math.pi, epsabs=0, epsrel=1e-13)[0] for k in range(2)]
    np.testing.assert_allclose(I.real, ref, rtol=1e-9)
    # budget
    with pytest.raises(K.QuadratureError):
        spectral_integral(F, 1, 8, (), xF, math.inf, 0.0, maxevals=5)
    assert np.all(spectral_integral(F, 1, 8, (), -1.0, math.inf, 0.0)[0] == 0)


def test_oracle_forms_and_orientation():
    rng = np.random.default_rng(3)
    NC = 5
    mt = rng.standard_normal((NC, NC, NC)) + 1j * rng.standard_normal((NC, NC, NC))
    a, b, c = (rng.standard_normal((4, NC)) + 1j * rng.standard_normal((4, NC)) for _ in range(3))
    brute = np.array([sum(mt[n, m, p] * a[q, n] * b[q, m] * c[q, p] for n in range(NC)
                          for m in range(NC) for p in range(NC)) for q in range(4)])
    np.testing.assert_allclose(O.cpge_forms(mt, a, b, c), brute, rtol=1e-13)
    # eigenbasis reference == i sum Gamma kernel3 Lambda with Gamma the
    # dense trace table (test/cpge_test.jl:409-434)
    Hn, J = helix()
    NC = 6
    w1, w2 = 0.3, 0.2
    Tm = O.chebyshev_matrices(Hn, NC)
    D = 6
    Jd = [j.toarray() for j in J]
    Gamma = np.array([[[np.trace(Jd[0] @ Tm[n] @ Jd[1] @ Tm[m] @ Jd[2] @ Tm[p]) / D for p in range(NC)]
                       for m in range(NC)] for n in range(NC)])
    Lam = O.cpge_lambda_table(NC, w1, w2, xF=0.15, beta=12.0, lam=0.03)
    kh = O.jackson(np.arange(NC), NC) * O.hn(np.arange(NC))
    direct = 1j * np.sum(Gamma * np.einsum("n,m,p->nmp", kh, kh, kh) * Lam)
    eig = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=0.15, beta=12.0, lam=0.03)
    assert abs(eig - direct) <= 1e-8 * abs(direct) + 2e-10


@pytest.mark.parametrize("w1,w2", [(0.31, -0.29), (0.3, 0.2)])
@pytest.mark.parametrize("xF,beta", [(0.15, math.inf), (0.15, 12.0)])
@pytest.mark.parametrize("lam", [0.0, 0.03])
def test_cpu_cpge_vs_eigen_reference(w1, w2, xF, beta, lam):
    NC = 12
    Hn, J = helix()
    mu3 = helix_mu3(NC)
    val = cpu_cpge(mu3, NC, w1, w2, E_f=xF, beta=beta, lam=lam)
    ref = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=xF, beta=beta, lam=lam)
    assert abs(val - ref) <= 1e-8 * abs(ref) + 2e-9, (val, ref)


@pytest.mark.parametrize("w1,w2", [(0.0, 0.2), (0.3, 0.0), (0.7, -0.7)])
def test_cpu_cpge_coincident_regularised(w1, w2):
    NC = 12
    Hn, J = helix()
    mu3 = helix_mu3(NC)
    v = cpu_cpge(mu3, NC, w1, w2, E_f=0.15, beta=math.inf, lam=0.05, rtol=1e-9)
    r = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=0.15, beta=math.inf, lam=0.05)
    assert abs(v - r) <= 1e-8 * abs(r) + 2e-9
    v4 = cpu_cpge(mu3, NC, w1, w2, E_f=0.15, beta=math.inf, lam=0.0, delta=1e-4, rtol=2e-8, atol=2e-10)
    r4 = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=0.15, beta=math.inf, lam=0.0,
                                delta=1e-4)
    assert abs(v4 - r4) <= 1e-8 * abs(r4) + 2e-9


def test_cpu_cpge_permuted_currents_differ():
    NC = 6
    w1, w2 = 0.3, 0.2
    a = cpu_cpge(helix_mu3(NC), NC, w1, w2, E_f=0.15, beta=12.0, lam=0.03, rtol=1e-9)
    b = cpu_cpge(helix_mu3(NC, 0, 2, 1), NC, w1, w2, E_f=0.15, beta=12.0, lam=0.03, rtol=1e-9)
    assert not abs(a - b) <= 1e-8 + 1e-3 * abs(b)


def test_cpge_input_errors():
    mu3 = np.ones((5, 5, 5), dtype=np.complex128)
    with pytest.raises(K.ArgumentError):
        K.cpge(mu3, 5, 0.3, 0.2, lambda_=-0.1)
    with pytest.raises(K.ArgumentError):
        K.cpge(mu3, 5, 0.3, 0.2, delta=-0.1)
    with pytest.raises(K.ArgumentError):
        K.cpge(mu3, 6, 0.3, 0.2)
    with pytest.raises(K.ArgumentError):
        K.cpge(np.ones((5, 5)), 5, 0.3, 0.2)
    for d in (-2.0, 0.0, 2.0):
        for args in ((d, 0.2), (0.3, d), (0.7, d - 0.7)):
            with pytest.raises(K.ArgumentError):
                K.cpge(mu3, 5, *args)
    with pytest.raises(K.ArgumentError) as e:
        K.cpge(mu3, 5, 0.31, -0.31)
    msg = str(e.value)
    assert "lambda > 0" in msg and "delta > 0" in msg and "0 or +/-2" in msg
    with pytest.raises(K.DomainError):
        K.d_cpge(mu3, 5, 0.3, 0.2, 1.0)
    with pytest.raises(K.ArgumentError):
        K.d_cpge(mu3, 5, 0.3, 0.2, N_int=0)
    with pytest.raises(K.DomainError):
        K.d_cpge(mu3, 5, 0.3, 0.2, e_range=(-1.5, 0.5))
