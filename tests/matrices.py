"""Test operators shared by the GPU test files: random Hermitian matrices
with ragged rows (empty rows up to 41 nonzeros), sized so each
nonzeros-per-row bucket of the recurrence pass is taken."""
import numpy as np
import scipy.sparse as sp

# band half-width -> the nonzeros-per-row bucket the widest row falls in
BUCKETS = {1: (1, 4), 3: (5, 8), 5: (9, 12), 7: (13, 16), 20: (17, 41)}


def ragged_hermitian(n, w, seed):
    """Random complex Hermitian with row i coupled to a random subset of rows
    i+1 .. i+w (and their mirror), a random real diagonal, every 11th row and
    column empty, and one full-width row, rescaled into (-1, 1) by its
    Gershgorin bound."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    full = n // 2
    for i in range(n):
        for d in range(1, w + 1):
            j = i + d
            if j >= n:
                break
            if i == full or j == full or rng.random() < 0.6:
                rows.append(i)
                cols.append(j)
                vals.append(complex(rng.normal(), rng.normal()))
    up = sp.coo_matrix((vals, (rows, cols)), shape=(n, n))
    H = (up + up.conj().T + sp.diags(rng.normal(size=n))).tocsr()
    keep = np.ones(n)
    keep[::11] = 0.0
    D = sp.diags(keep)
    H = (D @ H @ D).tocsr()
    H.eliminate_zeros()
    bound = np.max(np.asarray(abs(H).sum(axis=1)).ravel())
    return (H / (1.05 * bound)).tocsc()
