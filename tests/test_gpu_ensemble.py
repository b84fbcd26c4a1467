"""GPU run of the ensemble pipeline on a small manifest: every task's two
conductivity tables against the oracle's kpm_2d on the same H(+-eps) and
Gamma-paired probes (1e-11), the antithetic identity on the device tables,
the skip-on-rerun rule and pair averaging."""
import os

import numpy as np
import pytest

from kpmhip import ensemble as E
from kpmhip.frontend import random_phase_vectors
from kpmhip.rng import Xoshiro256pp
import oracle as O

pytestmark = pytest.mark.gpu

MANIFEST = """
[model]
W = 12.0
t = 1.0
a = 12.25
b = 0.0

[kpm]
NR = 2
arr_size = 4

[ensemble]
seed_base = 7
sizes = [4, 6]
NC = [16, 24]
pairs_per_size = 2
"""


def test_small_ensemble_vs_oracle(tmp_path):
    p = tmp_path / "m.toml"
    p.write_text(MANIFEST)
    mf = E.load_manifest(str(p))
    out = tmp_path / "data"
    paths = E.run_ensemble(mf, str(out), log=lambda *a: print(*a, flush=True))
    assert len(paths) == 4 and all(os.path.isfile(x) for x in paths)
    for task, path in zip(range(1, 5), paths):
        _, pair, L, NC, seed = E.task_point(mf, task)
        rec = E.load_task(path)
        assert rec["params"]["seed"] == seed and rec["params"]["L"] == L
        NH = L ** 3
        rng = Xoshiro256pp(seed)
        eps = 12.0 * (rng.rand(NH) - 0.5)
        psi = random_phase_vectors(rng, NH, 2)
        Hp, Jx, g = E.cubic_hamiltonian_and_current(L, eps)
        Hm, _, _ = E.cubic_hamiltonian_and_current(L, -eps)
        for H, ps, mu in ((Hp, psi, rec["mu_plus"]), (Hm, g[:, None] * psi, rec["mu_minus"])):
            ref = O.kpm_2d((H / 12.25).tocsc(), Jx, Jx, NC, ps, ps, NC, NC)
            np.testing.assert_allclose(mu, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
        assert E.antithetic_residual(rec["mu_plus"], rec["mu_minus"]) < 1e-10
        assert rec["meta"]["pair_residual"] < 1e-10
    stamp = [os.path.getmtime(x) for x in paths]
    assert E.run_ensemble(mf, str(out), log=lambda *a: None) == paths  # matching params: skipped
    assert [os.path.getmtime(x) for x in paths] == stamp
    avg, found, worst = E.average_pairs(mf, str(out), 2)
    assert found == 2 and avg.shape == (24, 24) and worst < 1e-10
