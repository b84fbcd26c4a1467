"""Batched disorder ensembles: kpm_1d of many independent realisations in
one fused launch per recurrence step (libkpm_hip's kpm_batch_*).

The reference computes an ensemble by running kpm_1d! once per realisation,
one process per task (scripts/anderson_seebeck_ensemble.jl:1-20;
src/moment.jl:420-492).  ``BatchOperator`` packs the realisations into one
block-diagonal device operator; ``kpm_batch_1d`` returns every
realisation's moments, each equal to the unbatched kpm_1d of that
realisation up to the summation order of the FP64 dot partials.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError, check, dptr, i64ptr
from .device import get_context, prec_code


class BatchOperator:
    def __init__(self, Hs, ctx=None):
        if len(Hs) == 0:
            raise ArgumentError("need at least one realisation")
        self.ctx = ctx or get_context()
        sizes = [int(H.shape[0]) for H in Hs]
        for H in Hs:
            if H.shape[0] != H.shape[1]:
                raise ArgumentError(f"operator must be square (got {H.shape})")
        self.sizes = sizes
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        A = sp.block_diag([sp.csc_matrix(H, dtype=np.complex128) for H in Hs], format="csc")
        A.sort_indices()
        colptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        rowval = np.ascontiguousarray(A.indices, dtype=np.int64)
        nz = np.ascontiguousarray(A.data, dtype=np.complex128)
        h = C.c_void_p()
        check(self.ctx.lib.kpm_batch_create(self.ctx.handle, len(Hs), i64ptr(self.offsets),
                                            i64ptr(colptr), i64ptr(rowval), dptr(nz), 0,
                                            C.byref(h)), "kpm_batch_create")
        self.handle = h
        self.nreal = len(Hs)

    def upload(self, psi, NR: int, precision="f64"):
        """Stage the probes (per-realisation list or stacked block) on the device."""
        NR = int(NR)
        if isinstance(psi, (list, tuple)):
            if len(psi) != self.nreal:
                raise ArgumentError("one probe block per realisation is required")
            psi = np.concatenate([np.asarray(p).reshape(p.shape[0], -1) for p in psi], axis=0)
        psi = np.asfortranarray(psi, dtype=np.complex128)
        if psi.shape != (int(self.offsets[-1]), NR):
            raise AssertionError(
                f"probe block has shape {psi.shape}; expected ({self.offsets[-1]}, {NR})")
        check(self.ctx.lib.kpm_batch_upload(self.ctx.handle, self.handle, NR, dptr(psi),
                                            prec_code(precision)), "kpm_batch_upload")
        self.NR = NR

    def run(self, NC: int, *, use_graph: bool = False) -> np.ndarray:
        """Moments of every realisation from the resident probes: (nreal, NR, NC)."""
        NC = int(NC)
        if NC % 2 != 0:
            raise AssertionError("Invalid NC: NC should be even.")
        mu = np.zeros((self.nreal, NC, self.NR), dtype=np.complex128)
        check(self.ctx.lib.kpm_batch_run(self.ctx.handle, self.handle, NC, dptr(mu),
                                         1 if use_graph else 0), "kpm_batch_run")
        return np.transpose(mu, (0, 2, 1)).copy()

    def autotune(self, reps: int = 5) -> dict:
        u, g, ms = C.c_int(), C.c_int(), C.c_float()
        check(self.ctx.lib.kpm_batch_autotune(self.ctx.handle, self.handle, int(reps), C.byref(u),
                                              C.byref(g), C.byref(ms)), "kpm_batch_autotune")
        return {"unroll": u.value, "grid": g.value, "ms_per_pass": ms.value}

    def info(self) -> dict:
        pr, nnz, ns = C.c_int64(), C.c_int64(), C.c_int()
        check(self.ctx.lib.kpm_batch_info(self.handle, C.byref(pr), C.byref(nnz), C.byref(ns)),
              "kpm_batch_info")
        return {"padded_rows": pr.value, "nnz": nnz.value, "segments": ns.value}

    def graph_builds(self) -> int:
        """How many times run(use_graph=True) has recorded its step graph."""
        n = C.c_int64()
        check(self.ctx.lib.kpm_batch_graph_builds(self.handle, C.byref(n)), "graph builds")
        return int(n.value)

    def __del__(self):
        from .device import _shutdown
        try:
            if getattr(self, "handle", None) and not _shutdown[0]:
                self.ctx.lib.kpm_batch_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def kpm_batch_1d(batch, NC: int, NR: int, psi, *, precision="f64", use_graph: bool = False,
                 avg_output: bool = True):
    """Moments of every realisation.  ``batch`` is a BatchOperator or a list
    of operators; ``psi`` a list of per-realisation NH_r x NR probe blocks or
    one stacked (sum NH_r) x NR block.  Returns an (nreal, NC) array of
    probe-averaged real moments, or (nreal, NR, NC) complex with
    ``avg_output=False``."""
    if not isinstance(batch, BatchOperator):
        batch = BatchOperator(batch)
    NC, NR = int(NC), int(NR)
    if NC % 2 != 0:
        raise AssertionError("Invalid NC: NC should be even.")
    batch.upload(psi, NR, precision)
    mu = batch.run(NC, use_graph=use_graph)
    if avg_output:
        return mu.mean(axis=1).real.copy()
    return mu
