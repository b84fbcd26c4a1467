"""Disorder-ensemble pipeline for conductivity moments
(scripts/anderson_seebeck_ensemble.jl, scripts/anderson_seebeck_manifest.toml).

The reference runs one process per task. A task is one antithetic disorder
pair at one lattice size. Each process builds H(+eps) and H(-eps) =
-Gamma H(+eps) Gamma with Gamma-paired probes, computes both conductivity
moment tables at the manifest's fixed rescaling, and writes them to one
file per task. Averaging and reconstruction happen offline.

Here ``run_ensemble`` executes any set of tasks in one process on the
device engine: one context, operators resident per task, and every table
a device kpm_2d (MFMA contraction). It writes the same per-task record as
the reference: ``params``, ``mu_plus``, ``mu_minus`` and ``meta``, with
matching-parameter skips and a refusal to overwrite a file whose
parameters differ.

The reference stores the record as JLD2, which is HDF5; no HDF5 library is
available here, so the record is an ``.npz`` archive. The archive holds the
same keys and the same NC x NC ComplexF64 tables (Julia column-major
mu[n, m]), with ``params``/``meta`` as JSON strings; ``load_task`` reads it
without unpickling. File names follow the reference's pattern with the
``.npz`` extension.

The disorder and probe streams come from the build's generator
(kpmhip.rng.Xoshiro256pp seeded with the task seed), not Julia's Xoshiro,
so tables are statistically, not bitwise, the reference's.
"""
from __future__ import annotations

import json
import os
import socket
import time

import numpy as np
import scipy.sparse as sp

from ._lib import ArgumentError
from .frontend import RescaledHamiltonian, cond_moments, random_phase_vectors
from .rng import Xoshiro256pp

SCHEMA = 1


def load_manifest(path: str) -> dict:
    """The TOML manifest (scripts/anderson_seebeck_manifest.toml): [model] W,
    t, a, b; [kpm] NR, arr_size; [ensemble] seed_base, sizes, NC,
    pairs_per_size."""
    try:
        import tomllib as _toml
    except ImportError:  # Python < 3.11
        import tomli as _toml
    with open(path, "rb") as f:
        mf = _toml.load(f)
    return validate_manifest(mf)


def validate_manifest(mf: dict) -> dict:
    for sec, keys in (("model", ("W", "t", "a", "b")), ("kpm", ("NR", "arr_size")),
                      ("ensemble", ("seed_base", "sizes", "NC", "pairs_per_size"))):
        if sec not in mf:
            raise ArgumentError(f"manifest: missing [{sec}]")
        for k in keys:
            if k not in mf[sec]:
                raise ArgumentError(f"manifest: missing {sec}.{k}")
    if len(mf["ensemble"]["NC"]) != len(mf["ensemble"]["sizes"]):
        raise ArgumentError("manifest: NC list must match sizes list")
    return mf


def n_tasks(mf: dict) -> int:
    return len(mf["ensemble"]["sizes"]) * int(mf["ensemble"]["pairs_per_size"])


def task_point(mf: dict, task: int):
    """Task -> (size_index, pair_index, L, NC, seed), 1-based as in the
    reference: task = (size_index - 1) * pairs_per_size + pair_index and
    seed = seed_base + 10000 size_index + pair_index."""
    ens = mf["ensemble"]
    P = int(ens["pairs_per_size"])
    if not (1 <= task <= n_tasks(mf)):
        raise ArgumentError(f"task {task} outside 1:{n_tasks(mf)}")
    size_index = (task - 1) // P + 1
    pair_index = (task - 1) % P + 1
    L = int(ens["sizes"][size_index - 1])
    NC = int(ens["NC"][size_index - 1])
    seed = int(ens["seed_base"]) + 10_000 * size_index + pair_index
    return size_index, pair_index, L, NC, seed


def cubic_hamiltonian_and_current(L: int, disorder, t: float = 1.0):
    """Periodic L^3 cubic lattice with on-site disorder and hopping -t; the
    x-current J_ij = H_ij dx_ij (minimum image) on bonds with dx != 0; the
    sublattice sign gamma_i = (-1)^(ix+iy+iz) (even L). Site i = ix + L (iy
    + L iz). Returns (H, Jx, gamma), CSC complex."""
    if L % 2:
        raise ArgumentError("antithetic particle-hole pairing requires even L")
    NH = L ** 3
    disorder = np.asarray(disorder, dtype=np.float64)
    if disorder.shape != (NH,):
        raise ArgumentError(f"expected {NH} disorder values")
    iz, iy, ix = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    i = ix + L * (iy + L * iz)
    gamma = np.where((ix + iy + iz) % 2 == 0, 1.0, -1.0)
    rows, cols, vals = [i], [i], [disorder[i].astype(np.complex128)]
    jr, jc, jv = [], [], []
    for dx, dy, dz in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        j = (ix + dx) % L + L * (((iy + dy) % L) + L * ((iz + dz) % L))
        for a, b in ((i, j), (j, i)):  # add_hopping!(i, j) then (j, i)
            rows.append(a)
            cols.append(b)
            vals.append(np.full(a.size, -t + 0j))
            xa, xb = a % L, b % L
            d = (xa - xb + L // 2) % L - L // 2  # minimum_image
            nz = d != 0
            jr.append(a[nz])
            jc.append(b[nz])
            jv.append((-t + 0j) * d[nz])
    H = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(NH, NH))
    Jx = sp.csc_matrix((np.concatenate(jv), (np.concatenate(jr), np.concatenate(jc))), shape=(NH, NH))
    Jx.eliminate_zeros()
    return H, Jx, gamma


def task_filename(L: int, NC: int, pair_index: int) -> str:
    return f"anderson_seebeck_L{L:02d}_NC{NC:04d}_pair{pair_index:02d}.npz"


def _params(mf, L, NC, seed, pair_index) -> dict:
    m, k = mf["model"], mf["kpm"]
    return {"W": float(m["W"]), "t": float(m["t"]), "a": float(m["a"]), "b": float(m["b"]), "L": L,
            "NC": NC, "NR": int(k["NR"]), "arr_size": int(k["arr_size"]), "seed": seed,
            "pair_index": pair_index, "schema": SCHEMA}


def save_task(path: str, params: dict, mu_plus, mu_minus, meta: dict):
    """Write one task record (write to a temporary name, then rename)."""
    tmp = path + ".part.npz"
    np.savez(tmp, params=np.array(json.dumps(params, sort_keys=True)), mu_plus=np.asarray(mu_plus),
             mu_minus=np.asarray(mu_minus), meta=np.array(json.dumps(meta, sort_keys=True)))
    os.replace(tmp, path)


def load_task(path: str) -> dict:
    """Read one task record: params and meta as dicts, the two tables as
    complex NC x NC arrays (no pickled objects are accepted)."""
    with np.load(path, allow_pickle=False) as z:
        return {"params": json.loads(str(z["params"])), "mu_plus": z["mu_plus"], "mu_minus": z["mu_minus"],
                "meta": json.loads(str(z["meta"]))}


def antithetic_residual(mu_plus, mu_minus) -> float:
    """max |mu_-[n,m] - (-1)^(n+m) mu_+[n,m]| / max |mu_+| (0-based n, m:
    the same parity as the reference's (n-1)+(m-1))."""
    mp, mm = np.asarray(mu_plus), np.asarray(mu_minus)
    sn = 1.0 - 2.0 * (np.arange(mp.shape[0]) % 2)
    sm = 1.0 - 2.0 * (np.arange(mp.shape[1]) % 2)
    return float(np.abs(mm - (sn[:, None] * mp) * sm[None, :]).max() / np.abs(mp).max())


def run_task(mf: dict, task: int, outdir: str, *, gemm: str = "f64", precision: str = "f64",
             log=print) -> str:
    """One task of the reference's main() (anderson_seebeck_ensemble.jl:
    119-276): returns the path written (or found with matching params)."""
    size_index, pair_index, L, NC, seed = task_point(mf, task)
    params = _params(mf, L, NC, seed, pair_index)
    out = os.path.join(outdir, task_filename(L, NC, pair_index))
    if os.path.isfile(out):
        try:
            existing = load_task(out)["params"]
        except Exception:
            existing = None
        if existing == params:
            log(f"SKIP: {out} exists with matching parameters")
            return out
        raise ArgumentError(f"{out} exists with DIFFERENT parameters — refusing to overwrite")
    os.makedirs(outdir, exist_ok=True)
    NH, NR = L ** 3, params["NR"]
    log(f"task {task}/{n_tasks(mf)}: L={L} NH={NH} NC={NC} NR={NR} pair={pair_index} seed={seed}")
    rng = Xoshiro256pp(seed)
    disorder = params["W"] * (rng.rand(NH) - 0.5)
    psi_plus = random_phase_vectors(rng, NH, NR)
    t0 = time.perf_counter()
    H_plus, Jx, gamma = cubic_hamiltonian_and_current(L, disorder, params["t"])
    H_minus, Jx_minus, _ = cubic_hamiltonian_and_current(L, -disorder, params["t"])
    t_build = time.perf_counter() - t0
    if (Jx != Jx_minus).nnz:
        raise AssertionError("the current operator must not depend on the disorder sign")
    G = sp.diags(gamma)
    if abs(H_minus + G @ H_plus @ G).max() / abs(H_minus).max() >= 1e-13:
        raise AssertionError("antithetic identity H(-eps) = -Gamma H(+eps) Gamma violated")
    psi_minus = np.asfortranarray(gamma[:, None] * psi_plus)
    a, b = params["a"], params["b"]
    h_plus = RescaledHamiltonian((H_plus / a).tocsc(), a, b)
    h_minus = RescaledHamiltonian((H_minus / a).tocsc(), a, b)
    # the manifest's arr_size is recorded in params; the tables come from the
    # engine's slot plan, and any blocking gives the same tables to rounding
    # (test/kpm2d_blocking_test.jl:97-117)
    kw = dict(NC=NC, precision=precision, gemm=gemm)
    t0 = time.perf_counter()
    m_plus = cond_moments(h_plus, Jx, Jx, psi_in=psi_plus, **kw)
    t_plus = time.perf_counter() - t0
    t0 = time.perf_counter()
    m_minus = cond_moments(h_minus, Jx, Jx, psi_in=psi_minus, **kw)
    t_minus = time.perf_counter() - t0
    res = antithetic_residual(m_plus.mu, m_minus.mu)
    log(f"antithetic pair identity residual = {res:.3e}")
    meta = {"hostname": socket.gethostname(), "device": "gpu", "engine": "libkpm_hip",
            "slot_plan": "engine", "gemm": gemm, "precision": precision, "pair_residual": res,
            "t_build": t_build,
            "t_moments_plus": t_plus, "t_moments_minus": t_minus}
    save_task(out, params, m_plus.mu, m_minus.mu, meta)
    log(f"build {t_build:.2f}s, moments +eps {t_plus:.2f}s, -eps {t_minus:.2f}s")
    log(f"WROTE: {out}")
    return out


def run_ensemble(mf: dict, outdir: str, tasks=None, **kw) -> list:
    """Run the given tasks (default: all) in this process on the device;
    returns the written paths in task order."""
    tasks = range(1, n_tasks(mf) + 1) if tasks is None else tasks
    return [run_task(mf, int(t), outdir, **kw) for t in tasks]


def average_pairs(mf: dict, datadir: str, size_index: int, *, allow_partial: bool = False):
    """Average the tables of all pairs of one size over both partners (the
    first step of scripts/anderson_seebeck_analysis.jl): returns (mu_avg,
    pairs found, the largest antithetic residual)."""
    ens = mf["ensemble"]
    L = int(ens["sizes"][size_index - 1])
    NC = int(ens["NC"][size_index - 1])
    acc, found, worst = None, 0, 0.0
    for p in range(1, int(ens["pairs_per_size"]) + 1):
        path = os.path.join(datadir, task_filename(L, NC, p))
        if not os.path.isfile(path):
            if allow_partial:
                continue
            raise ArgumentError(f"missing task file {path}")
        rec = load_task(path)
        pair = (rec["mu_plus"] + rec["mu_minus"]) / 2
        acc = pair if acc is None else acc + pair
        found += 1
        worst = max(worst, antithetic_residual(rec["mu_plus"], rec["mu_minus"]))
    if found == 0:
        raise ArgumentError(f"no task files for L={L}, NC={NC} in {datadir}")
    return acc / found, found, worst
