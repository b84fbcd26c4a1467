"""The doubling passes' own-row words taken from the diagonal's gather slot.

When every row of an operator stores its diagonal among the first KB
nonzeros (KB: the pass's gather slots), the doubling steps of kpm_1d
(src/moment.jl:339-418, both the per-probe table and the probe average) read
x[row] for their moment dots from that slot's gather instead of loading it a
second time.  The operators here put the diagonal inside and outside the
slots, drop it from some rows or from row 0 alone, and add long-row tails, so
both the reuse and the plain load run on every lane layout (1, 4, 32 and 64
lanes per row); each result must match the oracle as the other kpm_1d tests
do (FP64 to 1e-12, FP32 vectors within FP32 rounding).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def band(n, k, seed, drop=None):
    """Hermitian band operator with offsets -k..k (2k + 1 entries per interior
    row, the diagonal at slot k + 1 of a sorted row), scaled into (-1, 1).
    drop: rows whose diagonal is not stored."""
    rng = np.random.default_rng(seed)
    H = sp.lil_matrix((n, n), dtype=np.complex128)
    for d in range(1, k + 1):
        v = rng.standard_normal(n - d) + 1j * rng.standard_normal(n - d)
        H.setdiag(v, d)
        H.setdiag(np.conj(v), -d)
    diag = rng.uniform(0.5, 1.5, n) * rng.choice([-1.0, 1.0], n)
    if drop is not None:
        diag[drop] = 0.0
    H.setdiag(diag)
    H = H.tocsc()
    H.eliminate_zeros()
    bound = np.max(np.asarray(abs(H).sum(axis=1)).ravel())
    return (H / (1.05 * bound)).tocsc()


def pruned(H):
    H = H.copy()
    H.eliminate_zeros()
    return H


def diag_slot(H):
    """1-based slot of each row's diagonal in sorted CSR order (0: not stored)."""
    C = H.tocsr()
    C.sort_indices()
    out = np.zeros(H.shape[0], dtype=int)
    for i in range(H.shape[0]):
        cols = C.indices[C.indptr[i]:C.indptr[i + 1]]
        hit = np.nonzero(cols == i)[0]
        out[i] = hit[0] + 1 if hit.size else 0
    return out


CASES = {
    # name: (operator, whether every row stores its diagonal within KB slots)
    "tridiag": (lambda: band(700, 1, 1), True),
    "tridiag_some_rows_without": (lambda: band(700, 1, 2, drop=np.arange(0, 700, 5)), False),
    "tridiag_row0_without": (lambda: band(700, 1, 3, drop=[0]), False),
    "band8_tail": (lambda: band(500, 8, 4), True),         # 17 entries: KB 16 + tail
    "band16_diag_in_tail": (lambda: band(500, 16, 5), False),  # the diagonal at slot 17
    "graphene_disorder": (lambda: (models.graphene(14, 12, disorder=0.5, seed=6) / 3.6).tocsc(), True),
    # the clean lattice stores its zero on-site terms: a stored zero is a slot too
    "graphene_clean_stored_zeros": (lambda: (models.graphene(14, 12) / 3.1).tocsc(), True),
    "graphene_clean_no_diagonal": (lambda: pruned((models.graphene(14, 12) / 3.1).tocsc()), False),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_operator_layouts(name):
    make, inside = CASES[name]
    H = make()
    slots = diag_slot(H)
    counts = np.diff(H.tocsr().indptr)
    kb = 4 if counts.max() <= 4 else 6 if counts.max() <= 6 else 8 if counts.max() <= 8 else 16
    assert ((slots > 0) & (slots <= kb)).all() == inside, "the case does not cover what it names"
    n = H.shape[0]
    NC = 36
    for NR in (1, 8, 64):
        psi = K.random_phase_vectors(Xoshiro256pp(31 + NR), n, NR)
        ref = oracle.kpm_1d(H, NC, psi)
        for prec in ("f64", "f32"):
            print(f"diag gather {name} NR={NR} {prec}", flush=True)
            tol = 1e-12 if prec == "f64" else 2e-5
            per = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
            assert np.max(np.abs(per - ref)) < tol
            avg = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
            assert np.max(np.abs(avg - ref.mean(axis=0).real)) < tol


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_launch_shapes_with_reuse(prec):
    """The reuse under every schedule the tuner may record (persistent grids,
    one-shot chunks with each XCD mapping): the same moments to roundoff."""
    H = (models.graphene(40, 36, disorder=0.5, seed=8) / 3.6).tocsc()
    NR, NC = 64, 48
    psi = K.random_phase_vectors(Xoshiro256pp(77), H.shape[0], NR)
    ref = oracle.kpm_1d(H, NC, psi)
    tol = 1e-12 if prec == "f64" else 2e-5
    try:
        for unroll in (1, 2):
            for grid in (8, 256):
                K.tune_set(H, NR, prec, unroll=unroll, grid=grid)
                got = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
                assert np.max(np.abs(got - ref)) < tol
            for chunk, xm in ((1, 0), (4, 0), (2, 1), (4, 4)):
                print(f"diag gather shapes {prec} unroll={unroll} chunk={chunk} xmap={xm}", flush=True)
                K.tune_set_chunked(H, NR, prec, unroll=unroll, chunk=chunk, xcd_grouped=xm)
                avg = K.kpm_1d(H, NC, NR, psi_in=psi, precision=prec)
                assert np.max(np.abs(avg - ref.mean(axis=0).real)) < tol
                per = K.kpm_1d(H, NC, NR, psi_in=psi, avg_output=False, precision=prec)
                assert np.max(np.abs(per - ref)) < tol
    finally:
        K.tune_set(H, NR, prec, unroll=2, grid=0)
