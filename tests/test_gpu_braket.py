"""Bra/ket moments <psi_l|T_n(H)|psi_r> on the device (test/greens_test.jl:
22-110): odd NC against the direct three-term recurrence (1e-12) and the ED
spectral sum (1e-8), genuinely complex values, left == right against the
doubling path, and the kpm_1d wrapper's complex average / shapes."""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K

pytestmark = pytest.mark.gpu


def prescaled(rng, NH, margin=0.95):
    A = rng.standard_normal((NH, NH)) + 1j * rng.standard_normal((NH, NH))
    Hd = (A + A.conj().T) / 2
    Hd /= np.abs(np.linalg.eigvalsh(Hd)).max() / margin
    return Hd, sp.csc_matrix(Hd)


def unit_cols(x):
    return np.asfortranarray(x / np.linalg.norm(x, axis=0))


def test_bra_ket_odd_nc_vs_direct_and_ed():
    rng = np.random.default_rng(7)
    NH, NC, NR = 64, 33, 2
    Hd, H = prescaled(rng, NH)
    pl = unit_cols(rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR)))
    pr = unit_cols(rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR)))
    mu = K.kpm_1d(H, NC, NR, psi_in_l=pl, psi_in_r=pr, avg_output=False)
    direct = np.zeros((NR, NC), dtype=complex)
    for j in range(NR):
        prev, cur = pr[:, j], Hd @ pr[:, j]
        direct[j, 0], direct[j, 1] = np.vdot(pl[:, j], prev), np.vdot(pl[:, j], cur)
        for n in range(2, NC):
            prev, cur = cur, 2 * (Hd @ cur) - prev
            direct[j, n] = np.vdot(pl[:, j], cur)
    np.testing.assert_allclose(mu, direct, rtol=0, atol=1e-12)
    lam, V = np.linalg.eigh(Hd)
    for j in range(NR):
        w = np.conj(V.conj().T @ pl[:, j]) * (V.conj().T @ pr[:, j])
        ed = np.array([np.sum(w * np.cos(n * np.arccos(np.clip(lam, -1, 1)))) for n in range(NC)])
        np.testing.assert_allclose(mu[j], ed, rtol=0, atol=1e-8)
    assert np.abs(mu.imag).max() > 1e-3


def test_left_equals_right_matches_doubling():
    rng = np.random.default_rng(11)
    NH, NC, NR = 48, 32, 3
    _, H = prescaled(rng, NH)
    psi = np.exp(2j * np.pi * rng.random((NH, NR)))
    K.normalize_by_col(psi, NR)
    psi = np.asfortranarray(psi)
    dbl = K.kpm_1d(H, NC, NR, psi_in=psi.copy(), avg_output=False)
    lr = K.kpm_1d(H, NC, NR, psi_in_l=psi.copy(), psi_in_r=psi.copy(), avg_output=False)
    np.testing.assert_allclose(lr, dbl, rtol=0, atol=1e-12)
    assert np.abs(lr.imag).max() < 1e-12


def test_wrapper_complex_average_and_shapes():
    rng = np.random.default_rng(23)
    NH, NC, NR = 32, 17, 2
    _, H = prescaled(rng, NH)
    pl = np.asfortranarray(rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR)))
    pr = np.asfortranarray(rng.standard_normal((NH, NR)) + 1j * rng.standard_normal((NH, NR)))
    avg = K.kpm_1d(H, NC, NR, psi_in_l=pl.copy(), psi_in_r=pr.copy())
    assert np.iscomplexobj(avg) and avg.shape == (NC,)
    allm = K.kpm_1d(H, NC, NR, psi_in_l=pl.copy(), psi_in_r=pr.copy(), avg_output=False)
    np.testing.assert_allclose(allm.sum(axis=0) / NR, avg, rtol=0, atol=1e-12)
    ser = K.kpm_1d(H, NC, NR, psi_in_l=pl.copy(), psi_in_r=pr.copy(), NR_parallel=False)
    np.testing.assert_allclose(ser, avg, rtol=0, atol=1e-12)
    psi = np.exp(2j * np.pi * rng.random((NH, NR)))
    K.normalize_by_col(psi, NR)
    eq = K.kpm_1d(H, 16, NR, psi_in=np.asfortranarray(psi))
    assert not np.iscomplexobj(eq)
