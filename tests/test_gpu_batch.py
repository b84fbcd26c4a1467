"""Batched ensembles on the GPU: every realisation's moments from one
batched call equal the unbatched kpm_1d of that realisation (FP64: to the
reference's 1e-12 parity bound), and HIP-graph replays equal the eager run."""
import numpy as np
import pytest

import kpmhip as K
import oracle
from kpmhip import models
from kpmhip.rng import Xoshiro256pp

pytestmark = pytest.mark.gpu


def ensemble(seed=0):
    rng = Xoshiro256pp(seed)
    Hs = []
    for i, L in enumerate((3, 5, 4, 7, 6, 9, 3, 8)):
        W = 2.0 + 18.0 * rng.rand()
        H, *_ = models.cubic_model(L, W=W, seed=100 + i)
        a = 6.0 + W / 2 + 0.1  # Gershgorin bound of the cubic Anderson model
        Hs.append((H / a).tocsc())
    Hs.append((models.graphene(20, 13, disorder=0.4, seed=9) / 3.6).tocsc())
    return Hs


@pytest.mark.parametrize("NR", [1, 4, 16])
def test_batch_equals_unbatched_fp64(NR):
    Hs = ensemble()
    psis = [K.random_phase_vectors(Xoshiro256pp(7 + i), H.shape[0], NR) for i, H in enumerate(Hs)]
    NC = 64
    B = K.BatchOperator(Hs)
    got = K.kpm_batch_1d(B, NC, NR, psis, avg_output=False)
    for i, (H, p) in enumerate(zip(Hs, psis)):
        ref = oracle.kpm_1d(H, NC, p)
        np.testing.assert_allclose(got[i], ref, rtol=0, atol=1e-12, err_msg=f"realisation {i}")
    print("batch info", B.info(), flush=True)


def test_batch_fp32_close_to_fp64():
    Hs = ensemble(1)
    NR = 4
    psis = [K.random_phase_vectors(Xoshiro256pp(50 + i), H.shape[0], NR) for i, H in enumerate(Hs)]
    B = K.BatchOperator(Hs)
    m64 = K.kpm_batch_1d(B, 128, NR, psis)
    m32 = K.kpm_batch_1d(B, 128, NR, psis, precision="f32")
    assert np.max(np.abs(m32 - m64)) < 1e-6


def test_batch_graph_replays_match_eager():
    Hs = ensemble(2)
    NR = 4
    psis = [K.random_phase_vectors(Xoshiro256pp(80 + i), H.shape[0], NR) for i, H in enumerate(Hs)]
    B = K.BatchOperator(Hs)
    eager = K.kpm_batch_1d(B, 96, NR, psis, avg_output=False)
    for rep in range(3):
        g = K.kpm_batch_1d(B, 96, NR, psis, avg_output=False, use_graph=True)
        np.testing.assert_array_equal(g, eager)
        # eager work and a different shape in between replays
        _ = K.kpm_1d(Hs[0], 16, NR, psi_in=psis[0])
        e2 = K.kpm_batch_1d(B, 96, NR, psis, avg_output=False)
        np.testing.assert_array_equal(e2, eager)


def test_batch_graph_rerecorded_when_probe_width_changes():
    """A graph recorded at NR = 4 copies from the resident probe block.  An
    upload at NR = 16 reallocates that block, and an upload back at NR = 4
    restores the original shape key: the graph must be recorded again (it
    would otherwise copy from the freed block), and every replay must equal
    the eager run at that width."""
    Hs = ensemble(3)
    B = K.BatchOperator(Hs)
    NC = 64
    ps4 = [K.random_phase_vectors(Xoshiro256pp(90 + i), H.shape[0], 4) for i, H in enumerate(Hs)]
    ps16 = [K.random_phase_vectors(Xoshiro256pp(120 + i), H.shape[0], 16) for i, H in enumerate(Hs)]
    e4 = K.kpm_batch_1d(B, NC, 4, ps4, avg_output=False)
    e16 = K.kpm_batch_1d(B, NC, 16, ps16, avg_output=False)
    builds = B.graph_builds()
    g4 = K.kpm_batch_1d(B, NC, 4, ps4, avg_output=False, use_graph=True)
    np.testing.assert_array_equal(g4, e4)
    assert B.graph_builds() == builds + 1
    g16 = K.kpm_batch_1d(B, NC, 16, ps16, avg_output=False, use_graph=True)
    np.testing.assert_array_equal(g16, e16)
    assert B.graph_builds() == builds + 2
    g4b = K.kpm_batch_1d(B, NC, 4, ps4, avg_output=False, use_graph=True)
    np.testing.assert_array_equal(g4b, e4)
    assert B.graph_builds() >= builds + 2
    # same width, same buffers: the recorded graph is replayed as it is
    n = B.graph_builds()
    g4c = K.kpm_batch_1d(B, NC, 4, ps4, avg_output=False, use_graph=True)
    np.testing.assert_array_equal(g4c, e4)
    assert B.graph_builds() == n
    # the autotuner may rebuild the launch schedule: the graph follows it
    B.autotune(reps=2)
    g4d = K.kpm_batch_1d(B, NC, 4, ps4, avg_output=False, use_graph=True)
    np.testing.assert_allclose(g4d, e4, rtol=0, atol=1e-12)


def test_batch_rejects_coupled_blocks():
    import scipy.sparse as sp
    H = models.chain(10)
    A = sp.block_diag([H, H], format="csc").tolil()
    A[0, 15] = 0.1
    A[15, 0] = 0.1
    ctx = K.get_context()
    with pytest.raises(K.ArgumentError):
        import ctypes as C
        A = A.tocsc()
        off = np.array([0, 10, 20], dtype=np.int64)
        h = C.c_void_p()
        K._lib.check(ctx.lib.kpm_batch_create(ctx.handle, 2, K._lib.i64ptr(off),
                                              K._lib.i64ptr(A.indptr.astype(np.int64)),
                                              K._lib.i64ptr(A.indices.astype(np.int64)),
                                              K._lib.dptr(A.data.astype(np.complex128)), 0,
                                              C.byref(h)))
