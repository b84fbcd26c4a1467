"""Kubo–Bastin reconstruction helpers on the host, restating test/cond_test.jl:13-90
and :133-140: dc_cond0 against dc_cond_single at E_F = 0 (independent mod-4
slicing vs T_n(0) weights) for a non-symmetric table, the factorised Γ
contraction against the direct double sum, d_dc_cond zero outside the band,
its first and second energy derivatives against fourth-order
finite-difference stencils, and kernel application leaving its input intact.
Plus: the plain d_dc_cond against the direct Γ double sum, and the Taylor
derivatives against the closed-form derivative of a single-moment table.
"""
import math

import numpy as np

import kpmhip as K
from kpmhip.rng import Xoshiro256pp


def cmat(seed, n):
    r = Xoshiro256pp(seed)
    return (r.rand(n * n) - 0.5).reshape(n, n) + 1j * (r.rand(n * n) - 0.5).reshape(n, n)


def test_dc_cond0_equals_dc_cond_single_at_zero():
    NC, a = 32, 2.3
    mu = cmat(5, NC)  # generic: mu_nm != mu_mn
    v0 = K.dc_cond0(mu, a, NC=NC)
    v1 = K.dc_cond_single(mu, a, 0.0, NC=NC)
    assert abs(v0 - v1) <= 1e-12


def test_gamma_contraction_factorised_vs_direct():
    NC = 24
    mu = cmat(6, NC)
    n = np.arange(NC)
    for eps in (-0.83, 0.0, 0.37):
        direct = np.sum(K.Gamma_nm(n[:, None], n[None, :], eps) * mu)
        got = K.Gamma_nm_mu(mu, eps, NC)
        assert abs(got - direct) <= 1e-11 * abs(direct)


def test_d_dc_cond_zero_outside_band_and_direct_sum():
    NC, a, b = 16, 1.7, 0.2
    mu = cmat(8, NC)
    dsig = K.d_dc_cond(mu, a, [0.3, 2.5 * a])
    assert np.isfinite(dsig[0]) and dsig[1] == 0.0
    # the plain integrand is the direct Γ double sum over the kernel-improved table
    mt = K.mu2D_apply_kernel_and_h(mu, NC, K.JacksonKernel)
    n = np.arange(NC)
    for E in (b - 0.6 * a, b + 0.1 * a, b + 0.8 * a):
        x = (E - b) / a
        direct = np.real(np.sum(K.Gamma_nm(n[:, None], n[None, :], x) * mt)) / ((1 - x * x) ** 2 * a * a)
        got = K.d_dc_cond(mu, a, E, b=b, NC=NC)[0]
        assert abs(got - direct) <= 1e-11 * max(1.0, abs(direct))


def test_d_dc_cond_first_derivative_vs_fd():
    NC, a, b = 16, 1.7, 0.35
    mu = cmat(8, NC)

    def f0(E):
        return K.d_dc_cond(mu, a, [E], b=b, NC=NC)[0]

    h = 1e-3 * a
    for E in (b - 0.5 * a, b + 0.3 * a, b + 0.72 * a):
        d1 = K.d_dc_cond(mu, a, [E], b=b, NC=NC, dE_order=1)[0]
        fd = (f0(E - 2 * h) - 8 * f0(E - h) + 8 * f0(E + h) - f0(E + 2 * h)) / (12 * h)
        assert abs(d1 - fd) <= 1e-7 * abs(fd)
    # the scalar energy goes through the array form
    E = b + 0.3 * a
    assert np.array_equal(K.d_dc_cond(mu, a, E, b=b, NC=NC, dE_order=1),
                          K.d_dc_cond(mu, a, [E], b=b, NC=NC, dE_order=1))


def test_d_dc_cond_second_derivative_vs_stencil():
    NC, a, b = 16, 1.7, 0.35
    mu = cmat(8, NC)

    def f0(E):
        return K.d_dc_cond(mu, a, [E], b=b, NC=NC)[0]

    E, h = b, 1e-3 * a
    stencil = (-f0(E + 2 * h) + 16 * f0(E + h) - 30 * f0(E) + 16 * f0(E - h) - f0(E - 2 * h)) / (12 * h * h)
    d2 = K.d_dc_cond(mu, a, [E], b=b, NC=NC, dE_order=2)[0]
    assert abs(d2 - stencil) <= 1e-7 * abs(stencil)


def test_d_dc_cond_derivatives_closed_form():
    """A table with the single entry μ̃[0, 0] = 1 (kernel factor g_0 h_0 = 1)
    has Σ Γ μ̃ = Γ00(x) = 2x, so dσ(E) = 2x / ((1-x²)² a²): its first and
    third E derivatives in closed form pin the Taylor path exactly."""
    NC, a, b = 6, 1.3, -0.1
    mu = np.zeros((NC, NC), dtype=complex)
    mu[0, 0] = 1.0

    def f(x):
        return 2 * x / ((1 - x * x) ** 2 * a * a)

    def f1(x):  # d/dx of 2x (1-x²)^-2
        return (2 * (1 - x * x) ** -2 + 8 * x * x * (1 - x * x) ** -3) / (a * a)

    def f3(x):
        u = 1 - x * x
        return (24 * u ** -3 + 288 * x * x * u ** -4 + 384 * x ** 4 * u ** -5) / (a * a)

    for E in (b - 0.4 * a, b + 0.25 * a, b + 0.6 * a):
        x = (E - b) / a
        assert abs(K.d_dc_cond(mu, a, E, b=b, NC=NC)[0] - f(x)) <= 1e-12 * abs(f(x))
        d1 = K.d_dc_cond(mu, a, E, b=b, NC=NC, dE_order=1)[0]
        assert abs(d1 - f1(x) / a) <= 1e-11 * abs(f1(x) / a)
        d3 = K.d_dc_cond(mu, a, E, b=b, NC=NC, dE_order=3)[0]
        assert abs(d3 - f3(x) / a ** 3) <= 1e-10 * abs(f3(x) / a ** 3)


def test_kernel_application_leaves_input_intact():
    NC = 15
    mu2 = np.ones((NC, NC))
    t_mut = K.mu2D_apply_kernel_and_h(mu2.copy(), NC, K.JacksonKernel)
    t_nomut = K.kernels.mu2D_apply_kernel_and_h_no_mutate(mu2, NC, K.JacksonKernel)
    np.testing.assert_allclose(t_mut, t_nomut, rtol=1e-15)
    assert np.all(mu2 == 1)


def test_chebyshevT_accurate():
    for x in (0.0, 0.3, -0.91, 1.4):
        for n in range(9):
            ref = math.cos(n * math.acos(x)) if abs(x) <= 1 else math.cosh(n * math.acosh(x))
            assert abs(K.chebyshevT_accurate(n, x) - ref) <= 1e-12 * max(1.0, abs(ref))
    assert [K.chebyshevT_accurate(n, 0.0) for n in range(6)] == [1, 0, -1, 0, 1, 0]
