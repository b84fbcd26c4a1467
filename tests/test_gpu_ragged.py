"""Ragged sparsity on the device: Hermitian operators whose rows run from
empty to 41 nonzeros, so every nonzeros-per-row bucket of the recurrence pass
(4, 8, 12, 16 and the long-row tail beyond 16) is taken, crossed with probe
widths whose rows span 1, 2, 4, 8, 16 and 64 lanes.  kpm_1d is compared with the
oracle at 1e-12 (FP64 vectors) and bounded at FP32 rounding (FP32 vectors),
and the left/right form with distinct probes and the probe-averaged form are
checked the same way.
CUSPARSE's SpMM (ext/KPMCUDAExt.jl:84-88) takes any CSR; the lattice models of
the other tests have near-uniform rows, so this is where empty and long rows
are exercised.
"""
import numpy as np
import pytest

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
