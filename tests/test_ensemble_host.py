"""CPU pins of the ensemble pipeline (scripts/anderson_seebeck_ensemble.jl):
task enumeration and seeds, the cubic Anderson model with its x current and
the antithetic identity H(-eps) = -Gamma H(+eps) Gamma, the task record
round trip, the skip / refuse-to-overwrite rules and pair averaging."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import kpmhip as K
from kpmhip import ensemble as E

MANIFEST = """
[model]
W = 12.0
t = 1.0
a = 12.25
b = 0.0

[kpm]
NR = 8
arr_size = 16

[ensemble]
seed_base = 99532492
sizes = [16, 20, 24]
NC = [1024, 1024, 1024]
pairs_per_size = 12
"""


def manifest(tmp_path, text=MANIFEST):
    p = tmp_path / "m.toml"
    p.write_text(text)
    return E.load_manifest(str(p))


def test_task_enumeration(tmp_path):
    mf = manifest(tmp_path)
    assert E.n_tasks(mf) == 36
    assert E.task_point(mf, 1) == (1, 1, 16, 1024, 99532492 + 10001)
    assert E.task_point(mf, 12) == (1, 12, 16, 1024, 99532492 + 10012)
    assert E.task_point(mf, 13) == (2, 1, 20, 1024, 99532492 + 20001)
    assert E.task_point(mf, 36) == (3, 12, 24, 1024, 99532492 + 30012)
    with pytest.raises(K.ArgumentError):
        E.task_point(mf, 37)
    with pytest.raises(K.ArgumentError):
        E.validate_manifest({"model": {}, "kpm": {}, "ensemble": {}})


@pytest.mark.parametrize("L", [4, 6])  # L = 2 aliases the +-x bonds (the reference's own assert fails)
def test_cubic_model_and_antithetic_identity(L):
    rng = np.random.default_rng(L)
    NH = L ** 3
    eps = 12.0 * (rng.random(NH) - 0.5)
    Hp, Jx, g = E.cubic_hamiltonian_and_current(L, eps)
    Hm, Jxm, _ = E.cubic_hamiltonian_and_current(L, -eps)
    assert (Hp != Hp.conj().T).nnz == 0
    assert abs(Jx + Jx.conj().T).max() <= 1e-13 * abs(Jx).max()
    assert (Jx != Jxm).nnz == 0
    G = sp.diags(g)
    assert abs(Hm + G @ Hp @ G).max() == 0
    np.testing.assert_array_equal(Hp.diagonal().real, eps)
    off = Hp - sp.diags(Hp.diagonal())
    np.testing.assert_allclose(np.asarray(abs(off).sum(axis=1)).ravel(), 6.0)
    # Jx = H_ij dx_ij with the minimum image, on x bonds only
    Jd, Hd = Jx.toarray(), Hp.toarray()
    for i in range(0, NH, max(1, NH // 7)):
        for j in range(NH):
            dx = ((i % L) - (j % L) + L // 2) % L - L // 2
            assert Jd[i, j] == pytest.approx(Hd[i, j] * dx if i != j else 0.0)
    with pytest.raises(K.ArgumentError):
        E.cubic_hamiltonian_and_current(3, np.zeros(27))


def test_record_round_trip_skip_and_refuse(tmp_path):
    mf = manifest(tmp_path)
    _, pair, L, NC, seed = E.task_point(mf, 2)
    params = E._params(mf, L, NC, seed, pair)
    rng = np.random.default_rng(1)
    mp = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    n, m = np.indices(mp.shape)
    mm = np.where((n + m) % 2, -1, 1) * mp
    path = tmp_path / "out" / E.task_filename(L, NC, pair)
    os.makedirs(path.parent)
    E.save_task(str(path), params, mp, mm, {"note": "x"})
    rec = E.load_task(str(path))
    assert rec["params"] == params and rec["meta"] == {"note": "x"}
    np.testing.assert_array_equal(rec["mu_plus"], mp)
    assert E.antithetic_residual(rec["mu_plus"], rec["mu_minus"]) == 0
    # matching parameters: skipped before any device work
    assert E.run_task(mf, 2, str(path.parent), log=lambda *_: None) == str(path)
    # different parameters: refuse to overwrite
    bad = dict(params, W=13.0)
    E.save_task(str(path), bad, mp, mm, {})
    with pytest.raises(K.ArgumentError, match="DIFFERENT"):
        E.run_task(mf, 2, str(path.parent), log=lambda *_: None)
    # the archive holds no pickled objects
    with np.load(str(path), allow_pickle=False) as z:
        assert set(z.files) == {"params", "mu_plus", "mu_minus", "meta"}
        json.loads(str(z["params"]))


def test_average_pairs(tmp_path):
    text = MANIFEST.replace("pairs_per_size = 12", "pairs_per_size = 3")
    mf = manifest(tmp_path, text)
    rng = np.random.default_rng(2)
    L, NC = 16, 1024
    tabs = []
    for p in (1, 2, 3):
        mp = rng.standard_normal((3, 3)) + 0j
        mm = mp.copy()
        tabs.append((mp + mm) / 2)
        E.save_task(str(tmp_path / E.task_filename(L, NC, p)), {"p": p}, mp, mm, {})
    avg, found, worst = E.average_pairs(mf, str(tmp_path), 1)
    assert found == 3
    np.testing.assert_allclose(avg, sum(tabs) / 3)
    os.remove(tmp_path / E.task_filename(L, NC, 3))
    with pytest.raises(K.ArgumentError):
        E.average_pairs(mf, str(tmp_path), 1)
    assert E.average_pairs(mf, str(tmp_path), 1, allow_partial=True)[1] == 2
