"""Ragged sparsity on the device: Hermitian operators whose rows run from
empty to 41 nonzeros, so every nonzeros-per-row bucket of the recurrence pass
(4, 8, 12, 16 and the long-row tail beyond 16) is taken, crossed with probe
widths whose rows span 1, 2, 4, 8, 16 and 64 lanes.  kpm_1d is compared with the
oracle at 1e-12 (FP64 vectors) and bounded at FP32 rounding (FP32 vectors),
and the left/right form with distinct probes and the probe-averaged form are
checked the same way.
CUSPARSE's SpMM (ext/KPMCUDAExt.jl:84-88) takes any CSR; the lattice models of
the other tests have near-uniform rows, so this is where empty and long rows
are exercised, and banded operators fill each gather width exactly.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip.rng import Xoshiro256pp
from matrices import BUCKETS, ragged_hermitian

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w", sorted(BUCKETS))
@pytest.mark.parametrize("NR", [1, 2, 4, 8, 16, 64])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ragged_rows_vs_oracle(w, NR, prec):
    n, NC = 600, 32
    H = ragged_hermitian(n, w, seed=w)
    counts = np.diff(H.tocsr().indptr)
    lo, hi = BUCKETS[w]
    assert lo <= counts.max() <= hi and counts.min() == 0
    print(f"ragged kpm_1d prec={prec} NR={NR} w={w} max_row={counts.max()}", flush=True)
    psi = K.random_phase_vectors(Xoshiro256pp(7 + NR), n, NR)
    ref = oracle.kpm_1d(H, NC, psi)
    got = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
    if prec == "f64":
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
    else:  # FP32-rounded vectors, FP64 arithmetic: a few ulp of 2^-24 per step
        assert np.max(np.abs(got - ref)) < 2e-5
    pl = K.random_phase_vectors(Xoshiro256pp(100 + NR), n, NR)
    ref_lr = oracle.kpm_1d_lr(H, 17, pl, psi)
    got_lr = K.kpm_1d(H, 17, NR, psi_in_l=pl, psi_in_r=psi, avg_output=False, precision=prec)
    if prec == "f64":
        np.testing.assert_allclose(got_lr, ref_lr, rtol=1e-12, atol=1e-12)
    else:
        assert np.max(np.abs(got_lr - ref_lr)) < 2e-5
    # the probe-averaged form (kpm_1d's default, the probe-summed pass) on
    # the same rows: the oracle's per-probe table averaged over probes
    avg = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
    ref_avg = ref.mean(axis=0).real
    if prec == "f64":
        np.testing.assert_allclose(avg, ref_avg, rtol=0, atol=1e-12)
    else:
        assert np.max(np.abs(avg - ref_avg)) < 2e-5


def banded(n, offsets, seed):
    """Hermitian operator whose interior rows hold exactly the given offsets
    (0 = the diagonal), scaled into (-1, 1)."""
    rng = np.random.default_rng(seed)
    H = sp.lil_matrix((n, n), dtype=np.complex128)
    for d in sorted({abs(o) for o in offsets if o != 0}):
        v = rng.standard_normal(n - d) + 1j * rng.standard_normal(n - d)
        H.setdiag(v, d)
        H.setdiag(np.conj(v), -d)
    if 0 in offsets:
        H.setdiag(rng.uniform(0.5, 1.5, n))
    H = H.tocsc()
    bound = np.max(np.asarray(abs(H).sum(axis=1)).ravel())
    return (H / (1.05 * bound)).tocsc()


# the pass's gather widths (4, 5, 6, 7, 8, 10, 12, 16), each hit exactly by
# the operator's longest row
WIDTHS = {
    4: [-2, -1, 1, 2],
    5: [-2, -1, 0, 1, 2],
    6: [-3, -2, -1, 1, 2, 3],
    7: [-3, -2, -1, 0, 1, 2, 3],
    8: [-4, -3, -2, -1, 1, 2, 3, 4],
    10: [-5, -4, -3, -2, -1, 1, 2, 3, 4, 5],
    12: [-6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6],
    16: [-8, -7, -6, -5, -4, -3, -2, -1, 1, 2, 3, 4, 5, 6, 7, 8],
}


@pytest.mark.parametrize("width", sorted(WIDTHS))
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_exact_gather_widths(width, prec):
    """Operators whose longest row fills each gather width exactly, on rows of
    1, 4 and 32 lanes: per-probe and probe-averaged moments against the
    oracle."""
    n, NC = 400, 30
    H = banded(n, WIDTHS[width], seed=width)
    assert np.diff(H.tocsr().indptr).max() == width
    for NR in (1, 8, 64):
        print(f"gather width {width} {prec} NR={NR}", flush=True)
        psi = K.random_phase_vectors(Xoshiro256pp(3 + width + NR), n, NR)
        ref = oracle.kpm_1d(H, NC, psi)
        tol = 1e-12 if prec == "f64" else 2e-5
        got = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
        assert np.max(np.abs(got - ref)) < tol
        avg = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
        assert np.max(np.abs(avg - ref.mean(axis=0).real)) < tol


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_tiny_operators(n, prec):
    """Operators of 1-3 rows (nnz from 1 to 9): a lane's run of nonzeros
    reads far past nnz into the arrays' zero tail; the moments still match
    the oracle on rows of 1, 4 and 32 lanes."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    D = (A + A.conj().T) / 2
    D /= 1.05 * np.max(np.abs(np.linalg.eigvalsh(D)))
    H = sp.csc_matrix(D)
    for NR in (1, 8, 64):
        print(f"tiny operator n={n} {prec} NR={NR}", flush=True)
        psi = K.random_phase_vectors(Xoshiro256pp(40 + n + NR), n, NR)
        ref = oracle.kpm_1d(H, 16, psi)
        tol = 1e-12 if prec == "f64" else 2e-5
        got = K.kpm_1d(H, 16, NR, psi_in=psi, avg_output=False, precision=prec)
        assert np.max(np.abs(got - ref)) < tol
        avg = K.kpm_1d(H, 16, NR, psi_in=psi, precision=prec)
        assert np.max(np.abs(avg - ref.mean(axis=0).real)) < tol
