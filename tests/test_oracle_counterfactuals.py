"""Planted mistakes in the oracle must fail its pins (the reference's
mutation-harness idea, test/mutation/run_mutations.jl:1-18, applied to the
C restatement on the CPU).  Each mutant is a textual edit of
oracle/kpm_oracle.c compiled into a temporary library; the golden-fixture
checks of test_oracle_pins.py must reject every one."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "kpm_oracle.c")
GOLD = os.path.join(ROOT, "tests", "golden")

MUTANTS = {
    "dot_without_conj": ("for (int64_t i = 0; i < n; ++i) acc += conj(a[i]) * b[i];",
                         "for (int64_t i = 0; i < n; ++i) acc += a[i] * b[i];"),
    "doubling_wrong_base": ("colwise_dot(mu, NR, 2 * n - 1, views[ip], views[ip], NH, 2.0, 1, -1.0);",
                            "colwise_dot(mu, NR, 2 * n - 1, views[ip], views[ip], NH, 2.0, 2, -1.0);"),
    "doubling_factor": ("colwise_dot(mu, NR, 2 * n, views[ip], views[ipp], NH, 2.0, 2, -1.0);",
                        "colwise_dot(mu, NR, 2 * n, views[ip], views[ipp], NH, 1.0, 2, -1.0);"),
    "recurrence_sign": ("block_mul5(H, NH, NR, Vp, Vpp, 2.0, -1.0);",
                        "block_mul5(H, NH, NR, Vp, Vpp, 2.0, 1.0);"),
    "contract_without_conj": ("sr += ar * br + ai * bi;\n                        si += ar * bi - ai * br;",
                              "sr += ar * br - ai * bi;\n                        si += ar * bi + ai * br;"),
    "contract_chunk_skipped": ("for (size_t c0 = 0; c0 < K; c0 += kChunk) {",
                               "for (size_t c0 = kChunk; c0 < K; c0 += kChunk) {"),
    "kpm2d_missing_conj": ("mu[(n - 1) + NC * (m - 1)] += conj(mub[(jn - 1) + K_right * (jm - 1)]);",
                           "mu[(n - 1) + NC * (m - 1)] += mub[(jn - 1) + K_right * (jm - 1)];"),
    "kpm2d_transposed": ("mu[(n - 1) + NC * (m - 1)] += conj(mub[(jn - 1) + K_right * (jm - 1)]);",
                         "mu[(m - 1) + NC * (n - 1)] += conj(mub[(jn - 1) + K_right * (jm - 1)]);"),
    "ring_slot": ("const int ri = (int)((n - 1) % 3), rip = (int)((n + 1) % 3), ripp = (int)(n % 3);",
                  "const int ri = (int)((n - 1) % 3), rip = (int)(n % 3), ripp = (int)((n + 1) % 3);"),
    "current_no_adjoint": ("acc += conj(nz[k]) * x[rowval[k]];", "acc += nz[k] * x[rowval[k]];"),
    "action_skip_seed": ("    ACCUM(slots[0], 0);\n", "\n"),
    "action_conj_coef": ("for (size_t e = 0; e < blk; ++e) o[e] += (x)[e] * ck;",
                         "for (size_t e = 0; e < blk; ++e) o[e] += (x)[e] * conj(ck);"),
}


def build_mutant(name):
    old, new = MUTANTS[name]
    src = open(SRC).read()
    assert src.count(old) >= 1, f"mutation site for {name} not found"
    src = src.replace(old, new)
    d = tempfile.mkdtemp(prefix="kpm_mut_")
    cpath = os.path.join(d, "m.c")
    so = os.path.join(d, "libm.so")
    open(cpath, "w").write(src)
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-std=c11", "-shared", "-o", so, cpath],
                   check=True)
    return so


def pins_pass(lib_path) -> bool:
    """Golden-fixture pins evaluated with the library at lib_path."""
    saved = oracle._lib
    try:
        oracle._lib = None
        L = C.CDLL(lib_path)
        oracle._lib = None
        old_lib_path = oracle.LIB
        oracle.LIB = lib_path
        try:
            g1 = dict(np.load(os.path.join(GOLD, "kpm1d_dense.npz"), allow_pickle=False))
            g2 = dict(np.load(os.path.join(GOLD, "kpm2d_dense.npz"), allow_pickle=False))
            H = sp.csc_matrix(g1["H"])
            NC = int(g1["NC"])
            ok = np.allclose(oracle.kpm_1d(H, NC, g1["psi"]), g1["mu_eq"], rtol=1e-12, atol=1e-12)
            ok &= np.allclose(oracle.kpm_1d_current(H, sp.csc_matrix(g1["J"]), NC, g1["psi"]),
                              g1["mu_cur"], rtol=1e-12, atol=1e-12)
            args = [sp.csc_matrix(g2[k]) for k in ("H", "Ja", "Jb")]
            NC2 = int(g2["NC"])
            ok &= np.allclose(oracle.kpm_2d(*args, NC2, g2["psi_l"], g2["psi_r"], 3, 2), g2["mu"],
                              rtol=1e-12, atol=1e-12)
            g4 = dict(np.load(os.path.join(GOLD, "action_dense.npz"), allow_pickle=False))
            try:
                ok &= np.allclose(oracle.cheb_action(sp.csc_matrix(g4["H"]), g4["V"], g4["C"]),
                                  g4["out"], rtol=1e-12, atol=1e-12)
            except oracle.OracleUnstable:  # a mutant can make the recurrence grow
                ok = False
            return bool(ok)
        finally:
            oracle.LIB = old_lib_path
            del L
    finally:
        oracle._lib = saved


def test_unmutated_oracle_passes_pins():
    assert pins_pass(oracle.LIB if os.path.exists(oracle.LIB) else oracle.build())


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_caught(name):
    assert not pins_pass(build_mutant(name)), f"mutant {name} survived the pins"
