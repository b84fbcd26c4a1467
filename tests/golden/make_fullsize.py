"""Writes the full-size fixtures tests/golden/c3_sigma.npz and
tests/golden/c2_fp64_two_probes.npz from the oracle (the C restatement of
src/moment.jl in oracle/, test infrastructure) on this CPU.

c3: Kubo-Bastin sigma_xy of the Haldane model, honeycomb 256 x 256 cells
    (NH = 131,072), t1 = 1, t2 = 0.1 e^{i pi/2}, M = 0, periodic, Jx/Jy the
    bond currents; fixed rescaling a = 3.8, b = 0 (above the Gershgorin bound
    3 t1 + 6 t2 = 3.6); NC = 512, NR = 16 probes = random_phase_vectors(
    Xoshiro256pp(42), NH, 16) used as both psi_l and psi_r (cond_moments,
    src/frontend.jl:270-327); oracle kpm_2d with K_left = 512, K_right = 16
    (src/moment.jl:987-1088); sigma_xy from the oracle's restatement of
    kubo_bastin_cond (src/applications/conductivity.jl:101-157) at Fermi
    energies inside the gap (|E| < 0.52).  Stored: sigma_xy per energy and
    table checksums (row and column sums, the leading 32 x 32 block, four
    sums against seeded weight matrices, the Frobenius norm).
c2: kpm_1d of graphene 1024 x 1024 x 2 (NH = 2,097,152), t = 1, on-site
    disorder U[-0.5, 0.5] from Xoshiro256pp(42), H / 3.6, NC = 2048, on two
    probe columns random_phase_vectors(Xoshiro256pp(2024), NH, 2), FP64
    (src/moment.jl:420-492).  Stored: mu_all (2 x 2048).

Run: python tests/golden/make_fullsize.py [c3|c2]   (about 4 and 2 minutes on
8 cores; c3 holds 18 GB of slots in host memory).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "kpm.jl_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from kpmhip import models  # noqa: E402
from kpmhip.frontend import random_phase_vectors  # noqa: E402
from kpmhip.rng import Xoshiro256pp  # noqa: E402

C3 = dict(L=256, t2=0.1, a=3.8, b=0.0, NC=512, NR=16, seed=42,
          Ef=np.array([-0.3, -0.2, -0.1, 0.0, 0.1, 0.2, 0.3]))
C2 = dict(L=1024, disorder=0.5, hseed=42, a=3.6, NC=2048, NR=2, pseed=2024)


def table_checksums(mu, nweights=4, seed=7):
    """Size-independent summaries of an NC x NC complex table."""
    NC = mu.shape[0]
    rng = Xoshiro256pp(seed)
    W = [rng.rand(NC, NC) - 0.5 for _ in range(nweights)]
    return dict(rowsum=mu.sum(axis=1), colsum=mu.sum(axis=0), lead=np.array(mu[:32, :32]),
                wsum=np.array([np.sum(w * mu) for w in W]), fro=np.linalg.norm(mu))


def make_c3():
    p = C3
    H, Jx, Jy, area = models.haldane_model(p["L"], p["L"], t=1.0, t2=p["t2"])
    NH = H.shape[0]
    Hn = (H / p["a"]).tocsc()  # b = 0
    psi = random_phase_vectors(Xoshiro256pp(p["seed"]), NH, p["NR"])
    t0 = time.time()
    mu = oracle.kpm_2d(Hn, Jx, Jy, p["NC"], psi, psi, p["NC"], 16)
    print(f"c3 oracle kpm_2d: {time.time() - t0:.1f} s", flush=True)
    sig = np.array([oracle.kubo_bastin_cond(mu, p["a"], e, b=p["b"], NH=NH, area=area)
                    for e in p["Ef"]])
    print("c3 sigma_xy:", sig, flush=True)
    cs = table_checksums(mu)
    np.savez_compressed(os.path.join(HERE, "c3_sigma.npz"), sigma=sig, Ef=p["Ef"], L=p["L"],
                        t2=p["t2"], a=p["a"], b=p["b"], NC=p["NC"], NR=p["NR"], seed=p["seed"],
                        area=area, NH=NH, **{f"mu_{k}": v for k, v in cs.items()})


def make_c2():
    p = C2
    H = (models.graphene(p["L"], p["L"], disorder=p["disorder"], seed=p["hseed"]) / p["a"]).tocsc()
    NH = H.shape[0]
    psi = random_phase_vectors(Xoshiro256pp(p["pseed"]), NH, p["NR"])
    t0 = time.time()
    mu = oracle.kpm_1d(H, p["NC"], psi)
    print(f"c2 oracle kpm_1d: {time.time() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_fp64_two_probes.npz"), mu=mu, NH=NH, **p)


if __name__ == "__main__":
    which = sys.argv[1:] or ["c3", "c2"]
    if "c3" in which:
        make_c3()
    if "c2" in which:
        make_c2()
