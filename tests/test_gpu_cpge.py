"""GPU parity of the three-current (CPGE) reconstruction.

- kpm_trilinear_forms (FP64 MFMA) against the oracle's _cpge_contract!
  restatement on seeded tables, ragged NC and Q, rtol 1e-12 of the scale;
- the reference's own CPGE tests (test/cpge_test.jl:319-607) on the helix
  cluster with the device kpm_3d table:
  - the eigenbasis reference (rtol 1e-8, atol 2e-9);
  - the paper-tensor orientation;
  - coincident-edge regularisation by lambda and delta;
  - the resolved integrand d_cpge (rtol 1e-10);
  - the input and quadrature-budget errors.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import models
from kpmhip.cpge import Table3
import oracle as O

pytestmark = pytest.mark.gpu


def helix():
    H, J, _ = models.helix_cluster()
    ev = np.linalg.eigvalsh(H.toarray())
    b = (ev.max() + ev.min()) / 2
    a = (ev.max() - ev.min()) / (2 - 0.1)
    return sp.csc_matrix((H.toarray() - b * np.eye(6)) / a), J


def device_mu3(NC, order=(0, 1, 2)):
    Hn, J = helix()
    psi = np.eye(6, dtype=np.complex128)
    return K.kpm_3d(Hn, J[order[0]], J[order[1]], J[order[2]], NC, 6, 6, psi_in=psi)


@pytest.mark.parametrize("NC,Q", [(1, 1), (5, 7), (12, 33), (31, 96), (64, 100), (96, 40)])
def test_trilinear_forms_vs_oracle(NC, Q):
    rng = np.random.default_rng(NC * 1000 + Q)
    mt = rng.standard_normal((NC, NC, NC)) + 1j * rng.standard_normal((NC, NC, NC))
    a, b, c = (rng.standard_normal((Q, NC)) + 1j * rng.standard_normal((Q, NC)) for _ in range(3))
    print(f"trilinear NC={NC} Q={Q}", flush=True)
    with Table3(mt, NC) as t:
        got = t.forms(a, b, c)
        again = t.forms(a, b, c)
    ref = O.cpge_forms(mt, a, b, c)
    scale = np.sqrt(NC ** 3) * 3.0  # typical |F| of unit-variance random inputs
    assert np.abs(got - ref).max() <= 1e-12 * scale
    assert np.array_equal(got, again)  # fixed-order reductions: run to run identical


@pytest.mark.parametrize("NC", [3, 12, 64])
def test_device_rows_and_weighting_match_host(NC):
    """Device row expansion (kpm_trilinear_forms_rec) and device kernel
    weighting (kpm_table3_create_weighted) against the host coefficient rows
    and host-weighted table through the same contraction."""
    from kpmhip.cpge import _kernel_vector, cpge_node_coefficients, cpge_node_descriptors
    rng = np.random.default_rng(NC)
    mu = rng.standard_normal((NC, NC, NC)) + 1j * rng.standard_normal((NC, NC, NC))
    kv = _kernel_vector(K.JacksonKernel, NC)
    mt = K.mu3D_apply_kernel_and_h(mu, NC, K.JacksonKernel)
    theta = np.linspace(0.05, 3.1, 23)
    for lam, delta in ((0.0, 0.0), (0.02, 0.0), (0.0, 0.05)):
        a, b, c = cpge_node_coefficients(theta, NC, 0.3, -0.45, lam, delta)
        da, db, dc = cpge_node_descriptors(theta, 0.3, -0.45, lam, delta)
        with Table3(mt, NC) as th, Table3(mu, NC, kvec=kv) as td:
            host = th.forms(a, b, c)
            weighted = td.forms(a, b, c)
            dev = td.forms_rec(da, db, dc)
        assert np.array_equal(host, weighted)  # same weighting order, bit for bit
        scale = np.abs(host).max()
        assert np.abs(dev - host).max() <= 1e-12 * scale


def test_kpm_3d_helix_vs_dense():
    NC = 8
    Hn, J = helix()
    mu3 = device_mu3(NC)
    Tm = O.chebyshev_matrices(Hn, NC)
    Jd = [j.toarray() for j in J]
    ref = np.array([[[np.trace(Tm[m] @ Jd[1] @ Tm[n] @ Jd[2] @ Tm[p] @ Jd[0]) / 6 for p in range(NC)]
                     for m in range(NC)] for n in range(NC)])
    np.testing.assert_allclose(mu3, ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("w1,w2", [(0.31, -0.29), (0.3, 0.2)])
@pytest.mark.parametrize("xF,beta", [(0.15, math.inf), (0.15, 12.0)])
@pytest.mark.parametrize("lam", [0.0, 0.03])
def test_cpge_vs_eigen_reference(w1, w2, xF, beta, lam):
    NC = 12
    Hn, J = helix()
    mu3 = device_mu3(NC)
    print(f"cpge w=({w1},{w2}) xF={xF} beta={beta} lambda={lam}", flush=True)
    val = K.cpge(mu3, NC, w1, w2, E_f=xF, beta=beta, lambda_=lam, quad_rtol=1e-8, quad_atol=2e-11)
    ref = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=xF, beta=beta, lam=lam)
    assert abs(val - ref) <= 1e-8 * abs(ref) + 2e-9, (val, ref)


def test_cpge_paper_tensor_orientation():
    NC = 6
    w1, w2 = 0.3, 0.2
    Hn, J = helix()
    Tm = O.chebyshev_matrices(Hn, NC)
    Jd = [j.toarray() for j in J]
    Gamma = np.array([[[np.trace(Jd[0] @ Tm[n] @ Jd[1] @ Tm[m] @ Jd[2] @ Tm[p]) / 6 for p in range(NC)]
                       for m in range(NC)] for n in range(NC)])
    Lam = O.cpge_lambda_table(NC, w1, w2, xF=0.15, beta=12.0, lam=0.03)
    kh = O.jackson(np.arange(NC), NC) * O.hn(np.arange(NC))
    direct = 1j * np.sum(Gamma * np.einsum("n,m,p->nmp", kh, kh, kh) * Lam)
    kw = dict(E_f=0.15, beta=12.0, lambda_=0.03, quad_rtol=1e-9, quad_atol=2e-11)
    val = K.cpge(device_mu3(NC), NC, w1, w2, **kw)
    assert abs(val - direct) <= 1e-8 * abs(direct) + 2e-10
    permuted = K.cpge(device_mu3(NC, (0, 2, 1)), NC, w1, w2, **kw)
    assert not abs(permuted - direct) <= 1e-8 + 1e-3 * abs(direct)


@pytest.mark.parametrize("w1,w2", [(0.0, 0.2), (0.3, 0.0), (0.7, -0.7)])
def test_cpge_coincident_edges_regularised(w1, w2):
    NC = 12
    Hn, J = helix()
    mu3 = device_mu3(NC)
    v_lam = K.cpge(mu3, NC, w1, w2, E_f=0.15, lambda_=0.05, quad_rtol=1e-9, quad_atol=2e-11)
    r_lam = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=0.15, beta=math.inf, lam=0.05)
    assert np.isfinite(v_lam) and abs(v_lam - r_lam) <= 1e-8 * abs(r_lam) + 2e-9
    v4 = K.cpge(mu3, NC, w1, w2, E_f=0.15, delta=1e-4, quad_rtol=2e-8, quad_atol=2e-10)
    v5 = K.cpge(mu3, NC, w1, w2, E_f=0.15, delta=1e-5, quad_rtol=2e-8, quad_atol=2e-10)
    r4 = O.cpge_eigen_reference(Hn, J[0], J[1], J[2], NC, w1, w2, xF=0.15, beta=math.inf, lam=0.0,
                                delta=1e-4)
    assert np.isfinite(v4) and np.isfinite(v5)
    assert abs(v4 - r4) <= 1e-8 * abs(r4) + 2e-9
    assert abs(v4 - v5) / abs(v5) < 5e-2


def test_d_cpge_resolved_integrand():
    NC = 12
    w1, w2, x, lam = 0.3, 0.2, -0.23, 0.03
    Hn, J = helix()
    mu3 = device_mu3(NC)
    setup = O.cpge_eigen_setup(Hn, J[0], J[1], J[2], NC)
    ref = O.cpge_eigen_integrand(setup, NC, w1, w2, x, lam)
    got = K.d_cpge(mu3, NC, w1, w2, x, lambda_=lam)
    assert abs(got - ref) <= 1e-10 * abs(ref) + 1e-12
    grid, vals = K.d_cpge(mu3, NC, w1, w2, lambda_=lam, N_int=7, e_range=(-0.8, 0.8))
    point = np.array([K.d_cpge(mu3, NC, w1, w2, g, lambda_=lam) for g in grid])
    np.testing.assert_allclose(vals, point, rtol=1e-12)
    ref_d = O.cpge_eigen_integrand(setup, NC, w1, w2, 0.5, 0.0, delta=1e-3)
    got_d = K.d_cpge(mu3, NC, w1, w2, 0.5, delta=1e-3)
    assert abs(got_d - ref_d) <= 1e-10 * abs(ref_d) + 1e-12


def test_cpge_quadrature_budget_error():
    mu3 = np.ones((5, 5, 5), dtype=np.complex128)
    with pytest.raises(K.QuadratureError):
        K.cpge(mu3, 5, 0.3, 0.2, maxevals=5)
    # an all-zero table integrates to exactly zero
    assert K.cpge(np.zeros((5, 5, 5), dtype=np.complex128), 5, 0.3, 0.2) == 0
