"""CPU pins of the matrix-function action and its consumers (no GPU).

- the oracle's chebyshev_action! against the golden fixture, the
  column-by-column three-term recurrence and ED
  (test/chebyshev_action_test.jl:44-160), its stability guard
  (test/chebyshev_action_test.jl:163-199);
- host logic: evolution order / tail / coefficients
  (test/evolution_test.jl:133-250, scalar-oracle form), Fermi coefficients
  (test/chern_marker_test.jl:27-73), argument validation, and the
  open-boundary Haldane fixture against its ED marker
  (test/chern_marker_test.jl:150-175).
"""
import math
import os
import warnings

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.special import jv

import kpmhip as K
from kpmhip import models
import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def prescaled(rng, NH, margin=0.95):
    A = rng.standard_normal((NH, NH)) + 1j * rng.standard_normal((NH, NH))
    Hd = (A + A.conj().T) / 2
    Hd /= np.abs(np.linalg.eigvalsh(Hd)).max() / margin
    return Hd, sp.csc_matrix(Hd)


def test_action_golden():
    g = np.load(os.path.join(GOLD, "action_dense.npz"), allow_pickle=False)
    out = O.cheb_action(sp.csc_matrix(g["H"]), g["V"], g["C"])
    np.testing.assert_allclose(out, g["out"], rtol=0, atol=1e-12)


def test_action_dense_and_ed():
    rng = np.random.default_rng(151)
    NH, NC, NR, Kc = 48, 16, 2, 3
    Hd, Hs = prescaled(rng, NH)
    V = rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR))
    C = rng.standard_normal((NC, Kc)) + 1j * rng.standard_normal((NC, Kc))
    out = O.cheb_action(Hs, V, C)
    np.testing.assert_allclose(out, O.action_dense(Hd, V, C), rtol=0, atol=1e-12)
    np.testing.assert_allclose(out, O.action_ed(Hd, V, C), rtol=0, atol=1e-8)


def test_action_promotion_padding_and_K():
    rng = np.random.default_rng(157)
    NH, NC, NR = 40, 12, 2
    Hd, Hs = prescaled(rng, NH)
    Vr = rng.standard_normal((NH, NR))
    Cr = rng.standard_normal(NC)
    out = O.cheb_action(Hs, Vr, Cr)
    np.testing.assert_allclose(out[:, :, 0], O.action_dense(Hd, Vr.astype(complex), Cr)[:, :, 0],
                               atol=1e-12)
    # zero padding of the coefficient rows is exact
    C = rng.standard_normal((9, 2)) + 1j * rng.standard_normal((9, 2))
    Cpad = np.vstack([C, np.zeros((8, 2))])
    np.testing.assert_allclose(O.cheb_action(Hs, Vr, Cpad), O.cheb_action(Hs, Vr, C), atol=1e-14)
    for Kc in (1, 9):
        C = rng.standard_normal((8, Kc)) + 1j * rng.standard_normal((8, Kc))
        np.testing.assert_allclose(O.cheb_action(Hs, Vr, C), O.action_dense(Hd, Vr, C), atol=1e-12)


def test_action_guard():
    rng = np.random.default_rng(179)
    NH, NR = 32, 2
    Hd, Hs = prescaled(rng, NH)
    V = rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR))
    C = rng.standard_normal((64, 1)) + 1j * rng.standard_normal((64, 1))
    with pytest.raises(O.OracleUnstable):
        O.cheb_action(sp.csc_matrix(1.2 * Hd), V, C)
    # check_every = 0 disables the guard: the call runs (its result is garbage)
    assert O.cheb_action(sp.csc_matrix(1.2 * Hd), V, C, check_every=0).shape == (NH, NR, 1)
    np.testing.assert_allclose(O.cheb_action(Hs, 100 * V, C), 100 * O.cheb_action(Hs, V, C),
                               atol=1e-9)
    H1 = sp.csc_matrix(np.array([[1.1 + 0j]]))
    C15 = np.zeros(16)
    C15[15] = 1.0
    for amp in (1.0, 1e-8, 1e-12):
        with pytest.raises(O.OracleUnstable):
            O.cheb_action(H1, np.array([[amp + 0j]]), C15)
    z = O.cheb_action(H1, np.array([[0j]]), C15)
    assert np.all(z == 0)


# ---- evolution host logic -----------------------------------------------------


def test_evolution_order_and_tail():
    orders = [K.evolution_order(2.0, t) for t in (0.5, 2.0, 10.0, 50.0)]
    assert orders == sorted(orders)
    assert K.evolution_order(2.0, 10.0, tol=1e-6) <= K.evolution_order(2.0, 10.0)
    assert K.evolution_order(2.0, 0.0) == 8
    assert K.evolution_order(2.0, 0.0, NC_min=32) == 32
    z = 1995.4
    NC = K.evolution_order(1.0, z)
    assert K.evolution_tail(z, NC) < 1e-12
    assert K.evolution_tail(z, NC) >= 2 * abs(jv(NC, z))
    for args, kw in (((1.0, 1e20), {}), ((1e308, 2.0), {}), ((1.0, 100.0), {"NC_cap": 50}),
                     ((1.0, 1.0), {"NC_min": 100, "NC_cap": 10}), ((-1.0, 1.0), {})):
        with pytest.raises(K.ArgumentError):
            K.evolution_order(*args, **kw)
    assert K.evolution_order(1.0, 0.0, NC_min=8, NC_cap=8) == 8


def test_evolution_scalar_oracle():
    """1x1 H_norm = [x]: sum_n C[n] T_n(x) must equal e^{-i(ax+b)t} to the
    default tol (test/evolution_test.jl:168-191), with T_n(x) = cos(n acos x)."""
    for x, at in ((0.0, 1995.4), (0.0, 2989.2), (0.3, 500.0), (-0.7, 3000.0)):
        a, b = 2.0, 0.4
        t = at / a
        C, NC, tail = K.evolution_coefficients(a, b, [t])
        Tn = np.cos(np.arange(NC) * math.acos(x))
        val = C[:, 0] @ Tn
        assert abs(val - np.exp(-1j * (a * x + b) * t)) <= 1e-12
        assert tail[0] < 1e-12


def test_evolution_coefficients_table_and_warnings():
    a, b, ts = 1.7, 0.3, [-1.2, 0.0, 2.5]
    with pytest.warns(UserWarning, match="evolution series tail"):  # |at| = 4.25 needs NC > 20
        C, NC, _ = K.evolution_coefficients(a, b, ts, NC=20)
    for k, t in enumerate(ts):
        for n in range(NC):
            ref = np.exp(-1j * b * t) * (1 if n == 0 else 2) * (-1j) ** n * jv(n, a * t)
            assert abs(C[n, k] - ref) <= 1e-15 * max(1.0, abs(ref)) + 1e-17
    with pytest.warns(UserWarning, match="evolution series tail"):
        K.evolution_coefficients(1.0, 0.0, [2989.2], NC=3116)
    NC_ok = K.evolution_order(1.0, 2989.2)
    assert NC_ok > 3116
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        K.evolution_coefficients(1.0, 0.0, [2989.2], NC=NC_ok)
    with pytest.raises(K.ArgumentError):
        K.evolution_coefficients(0.0, 0.0, [1.0])
    with pytest.raises(K.ArgumentError):
        K.evolution_coefficients(1.0, 0.0, [1.0], NC=1)
    with pytest.raises(K.ArgumentError):
        K.evolution_coefficients(1.0, 0.0, [np.nan])
    with pytest.raises(K.ArgumentError):
        K.evolution_coefficients(1.0, 0.0, [])


# ---- Fermi coefficients ---------------------------------------------------------


def cheb_eval(c, x):
    return float(np.cos(np.arange(len(c)) * math.acos(x)) @ c)


def test_fermi_coefficients():
    a, b, Ef, NC = 2.3, 0.4, 0.7, 256
    xF = (Ef - b) / a
    c_step = K.fermi_coefficients(a, b, Ef, NC=NC)
    c_quad = K.fermi_coefficients(a, b, Ef, beta=1e8, NC=NC)
    assert np.max(np.abs(c_step - c_quad)) < 2e-3
    c_fine = K.fermi_coefficients(a, b, Ef, beta=1e8, NC=NC, Np=8192)
    assert np.max(np.abs(c_step - c_fine)) < 1e-4
    c_bare = K.fermi_coefficients(a, b, Ef, NC=NC, kernel=lambda n, N: np.ones_like(n, dtype=float))
    for x in (-0.9, -0.5, xF - 0.2, xF + 0.2, 0.6, 0.9):
        step = 1.0 if x < xF else 0.0
        assert abs(cheb_eval(c_bare, x) - step) < 1e-2
        assert abs(cheb_eval(c_step, x) - step) < 1e-3
    beta = 20.0
    c_T = K.fermi_coefficients(a, b, Ef, beta=beta, NC=NC)
    for x in (-0.9, -0.3, xF, 0.5, 0.9):
        f = 1 / (math.exp((a * x + b - Ef) * beta) + 1)
        assert abs(cheb_eval(c_T, x) - f) < 1e-3
    for kw in ({"Ef": b + 1.5 * a}, {"Ef": b - a}):
        with pytest.raises(K.ArgumentError):
            K.fermi_coefficients(a, b, kw["Ef"], NC=NC)
    with pytest.raises(K.ArgumentError):
        K.fermi_coefficients(-a, b, Ef, NC=NC)
    with pytest.raises(K.ArgumentError):
        K.fermi_coefficients(a, b, Ef, NC=0)
    with pytest.raises(K.ArgumentError):
        K.fermi_coefficients(a, b, Ef, NC=NC, beta=0.0)
    with pytest.raises(K.ArgumentError):
        K.fermi_coefficients(a, b, Ef, NC=NC, Np=0)
    with pytest.raises(TypeError):
        K.fermi_coefficients(a, b, Ef)  # NC is required


# ---- argument validation (raised before any device work) -----------------------


def test_action_argument_validation():
    rng = np.random.default_rng(181)
    NH, NC, NR, Kc = 24, 8, 2, 2
    _, Hs = prescaled(rng, NH)
    V = rng.standard_normal((NH, NR)) + 0j
    C = rng.standard_normal((NC, Kc)) + 0j
    bad = [
        lambda: K.chebyshev_action(Hs, np.zeros((NH + 1, NR), complex), C),
        lambda: K.chebyshev_action_(np.zeros((NH, NR, Kc + 1), complex), Hs, V, C),
        lambda: K.chebyshev_action_(np.zeros((NH, NR + 1, Kc), complex), Hs, V, C),
        lambda: K.chebyshev_action(Hs, V, np.zeros((0, Kc), complex)),
        lambda: K.chebyshev_action(Hs, V, C, check_every=-1),
        lambda: K.chebyshev_action(Hs, np.zeros((NH, 0), complex), C),
        lambda: K.chebyshev_action(Hs, V, np.zeros((NC, 0), complex)),
    ]
    Vc = rng.standard_normal((NH, NR)) + 0j
    bad.append(lambda: K.chebyshev_action_(Vc.reshape(NH, NR, 1), Hs, Vc, C[:, :1]))
    w = np.zeros((NH, NR), complex)
    bad.append(lambda: K.chebyshev_action_(np.zeros((NH, NR, 1), complex), Hs, V, C[:, :1],
                                           slots=(w, w)))
    bad.append(lambda: K.chebyshev_action_(np.zeros((NH, NR, 1), complex), Hs, V, C[:, :1],
                                           slots=(w, np.zeros((NH + 1, NR), complex))))
    for f in bad:
        with pytest.raises(K.ArgumentError):
            f()


def test_open_haldane_fixture_marker():
    """The open flake's exact marker: bulk average = +1 (the sign of sigma_xy),
    whole-sample sum = 0, trivial phase ~ 0 (test/chern_marker_test.jl:150-175)."""
    Lx = Ly = 10
    H, pos, Ac = models.haldane_open_model(Lx, Ly)
    assert H.shape == (200, 200) and abs(Ac - 3 * math.sqrt(3) / 2) < 1e-12
    bulk = [s for x in range(4, 8) for y in range(4, 8)
            for s in (2 * ((y - 1) + Ly * (x - 1)), 2 * ((y - 1) + Ly * (x - 1)) + 1)]
    mk = O.ed_chern_marker(H, pos[:, 0], pos[:, 1], 0.0)
    assert abs(mk[bulk].sum() / (16 * Ac) - 1.0) < 0.01
    assert abs(mk.sum()) < 1e-8
    Ht, post, _ = models.haldane_open_model(Lx, Ly, m=1.6)
    mkt = O.ed_chern_marker(Ht, post[:, 0], post[:, 1], 0.0)
    assert abs(mkt[bulk].sum() / (16 * Ac)) < 0.01
