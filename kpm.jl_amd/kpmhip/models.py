"""Lattice model generators for the benchmark and test workloads (vectorised,
CSC ComplexF64, 0-based).  Each restates a model builder of the reference:

- ``square_lattice``   benchmarks/core_bench.jl:35-64 (periodic square, hopping -1,
                       minimum-image currents)
- ``square_anderson``  the square lattice with on-site disorder U[-W/2, W/2]
- ``graphene``         examples/GrapheneModel.jl:33-71 (honeycomb, cell index
                       mod(y) + Ny*mod(x), A = 2c, B = 2c+1), plus on-site disorder
- ``haldane_model``    test/ed_reference.jl:69-147 (with Jx, Jy, area)
- ``cubic_model``      test/ed_reference.jl:396-442 (with Jx, Jy, Jz, volume)
- ``chain``            the periodic 1-D chain of test/dos_test.jl:214-228
- ``haldane_open_model`` test/ed_reference.jl:149-220 (open flake, site
                       coordinates and cell area, for local Chern markers)
- ``ring_model`` / ``flux_ring_model``  test/ed_reference.jl:530-570
- ``square_disordered`` the periodic square lattice with on-site eps_i ~ U[-1, 1]
- ``square_model``     test/ed_reference.jl:653-690 (periodic square with site
                       coordinates and the minimum-image displacement)
- ``ring_geometry``    the coordinates / displacement of ring_model and
                       flux_ring_model (test/ed_reference.jl:535-573)
- ``helix_cluster``    test/cpge_test.jl:10-35 (six-site chiral helix with three
                       bond-current operators, for the three-current response)
                       and Hermitian-tied bond hoppings t_ij ~ 1 + U[-0.1, 0.1]
                       (the trainable-parameter lattice of the gradient workload)

Disorder comes from the build's xoshiro256++ generator (kpmhip.rng), not from
Julia's Xoshiro stream.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from .rng import Xoshiro256pp


def _csc(rows, cols, vals, n):
    return sp.coo_matrix((np.asarray(vals, dtype=np.complex128), (np.asarray(rows), np.asarray(cols))),
                         shape=(n, n)).tocsc()


def square_lattice(Lx: int, Ly: int, *, currents: bool = False, t: float = 1.0):
    """Periodic square lattice, every directed NN entry -t; currents
    Jα[i,j] = H[i,j] (r_i - r_j)_α with minimum-image displacements."""
    NH = Lx * Ly
    x, y = np.meshgrid(np.arange(Lx), np.arange(Ly), indexing="ij")
    x = x.T.ravel()  # y outer, x inner (site = x + Lx*y)
    y = y.T.ravel()
    site = x + Lx * y
    rows, cols, hv, jx, jy = [], [], [], [], []
    mi = lambda d, L: (d + L // 2) % L - L // 2  # noqa: E731
    for xn, yn in (((x + 1) % Lx, y), ((x - 1) % Lx, y), (x, (y + 1) % Ly), (x, (y - 1) % Ly)):
        j = xn + Lx * yn
        rows.append(site)
        cols.append(j)
        hv.append(np.full(NH, -t, dtype=np.complex128))
        jx.append(-t * mi(x - xn, Lx).astype(np.complex128))
        jy.append(-t * mi(y - yn, Ly).astype(np.complex128))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    H = _csc(rows, cols, np.concatenate(hv), NH)
    if not currents:
        return H
    return H, _csc(rows, cols, np.concatenate(jx), NH), _csc(rows, cols, np.concatenate(jy), NH)


def square_anderson(L: int, W: float = 2.0, *, t: float = 1.0, seed: int = 42, rng=None):
    """L x L periodic square lattice, hopping -t, on-site ε_i ~ U[-W/2, W/2]."""
    H = square_lattice(L, L, t=t)
    rng = rng if rng is not None else Xoshiro256pp(seed)
    eps = W * (rng.rand(L * L) - 0.5)
    return (H + sp.diags(eps.astype(np.complex128), format="csc")).tocsc()


def graphene(Nx: int, Ny: int, *, t: float = 1.0, disorder: float = 0.0, delta: float = 0.0,
             seed: int = 42, rng=None):
    """Honeycomb Nx x Ny cells x 2 orbitals, NN hopping -t, staggered Δ/2, and
    on-site disorder U[-disorder, disorder] (disorder = half-width)."""
    N = 2 * Nx * Ny
    xs, ys = np.meshgrid(np.arange(Nx), np.arange(Ny), indexing="ij")
    xs = xs.ravel()
    ys = ys.ravel()
    cell = lambda x, y: (y % Ny) + Ny * (x % Nx)  # noqa: E731  0-based cell index
    cur = cell(xs, ys)
    pbx = cell(xs + 1, ys)
    pby = cell(xs, ys + 1)
    A = 2 * cur
    B = 2 * cur + 1
    rows = np.concatenate([A, A, B, B, B, 2 * pbx, B, 2 * pby])
    cols = np.concatenate([A, B, A, B, 2 * pbx, B, 2 * pby, B])
    nc = len(cur)
    vals = np.concatenate([np.full(nc, delta / 2), np.full(nc, -t), np.full(nc, -t),
                           np.full(nc, -delta / 2), np.full(nc, -t), np.full(nc, -t),
                           np.full(nc, -t), np.full(nc, -t)]).astype(np.complex128)
    if disorder != 0.0:
        rng = rng if rng is not None else Xoshiro256pp(seed)
        eps = disorder * (2.0 * rng.rand(N) - 1.0)
        rows = np.concatenate([rows, np.arange(N)])
        cols = np.concatenate([cols, np.arange(N)])
        vals = np.concatenate([vals, eps.astype(np.complex128)])
    return _csc(rows, cols, vals, N)


def haldane_model(Lx: int, Ly: int, *, t: float = 1.0, t2: float = 0.2, phi: float = math.pi / 2,
                  m: float = 0.0):
    """Haldane model on an Lx x Ly honeycomb torus: returns (H, Jx, Jy, area)."""
    a1 = np.array([math.sqrt(3), 0.0])
    a2 = np.array([math.sqrt(3) / 2, 1.5])
    dAB = np.array([math.sqrt(3) / 2, 0.5])
    N = 2 * Lx * Ly
    X, Y = np.meshgrid(np.arange(1, Lx + 1), np.arange(1, Ly + 1), indexing="ij")
    X = X.ravel()
    Y = Y.ravel()
    cell = lambda x, y: ((y - 1) % Ly) + Ly * ((x - 1) % Lx)  # noqa: E731  0-based
    sA = lambda x, y: 2 * cell(x, y)  # noqa: E731
    sB = lambda x, y: 2 * cell(x, y) + 1  # noqa: E731
    I, J, V, VX, VY = [], [], [], [], []

    def addhop(i, j, amp, d):
        amp = np.broadcast_to(np.asarray(amp, dtype=np.complex128), i.shape)
        I.extend([i, j])
        J.extend([j, i])
        V.extend([amp, np.conj(amp)])
        VX.extend([amp * d[0], -np.conj(amp) * d[0]])
        VY.extend([amp * d[1], -np.conj(amp) * d[1]])

    addhop(sB(X, Y), sA(X, Y), -t + 0j, dAB)
    addhop(sB(X, Y), sA(X + 1, Y), -t + 0j, dAB - a1)
    addhop(sB(X, Y), sA(X, Y + 1), -t + 0j, dAB - a2)
    for d, (xo, yo) in ((a1, (1, 0)), (a2 - a1, (-1, 1)), (-a2, (0, -1))):
        addhop(sA(X + xo, Y + yo), sA(X, Y), t2 * np.exp(1j * phi), d)
        addhop(sB(X + xo, Y + yo), sB(X, Y), t2 * np.exp(-1j * phi), d)
    z = np.zeros(X.shape, dtype=np.complex128)
    I.extend([sA(X, Y), sB(X, Y)])
    J.extend([sA(X, Y), sB(X, Y)])
    V.extend([np.full(X.shape, m + 0j), np.full(X.shape, -m + 0j)])
    VX.extend([z, z])
    VY.extend([z, z])
    I, J = np.concatenate(I), np.concatenate(J)
    H = _csc(I, J, np.concatenate(V), N)
    Jx = _csc(I, J, np.concatenate(VX), N)
    Jy = _csc(I, J, np.concatenate(VY), N)
    area = Lx * Ly * abs(a1[0] * a2[1] - a1[1] * a2[0])
    return H, Jx, Jy, area


def cubic_model(L: int, *, t: float = 1.0, W: float = 0.0, seed: int = 1, rng=None):
    """Periodic L^3 simple-cubic model, hopping -t, on-site W*(U - 1/2) drawn in
    site order (x fastest); returns (H, Jx, Jy, Jz, volume)."""
    if L < 3:
        from ._lib import ArgumentError
        raise ArgumentError("cubic_model: L must be at least 3")
    if W < 0:
        from ._lib import ArgumentError
        raise ArgumentError("cubic_model: W must be nonnegative")
    N = L ** 3
    z, y, x = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    site = lambda a, b, c: (a % L) + L * (b % L) + L * L * (c % L)  # noqa: E731
    j = site(x, y, z)
    I, J, V, VX, VY, VZ = [], [], [], [], [], []
    for (dx, dy, dz) in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        i = site(x + dx, y + dy, z + dz)
        amp = np.full(N, -t + 0j)
        I.extend([i, j])
        J.extend([j, i])
        V.extend([amp, np.conj(amp)])
        VX.extend([amp * dx, -np.conj(amp) * dx])
        VY.extend([amp * dy, -np.conj(amp) * dy])
        VZ.extend([amp * dz, -np.conj(amp) * dz])
    rng = rng if rng is not None else Xoshiro256pp(seed)
    onsite = W * (rng.rand(N) - 0.5) if W != 0 else np.zeros(N)
    zz = np.zeros(N, dtype=np.complex128)
    I.append(j)
    J.append(j)
    V.append(onsite.astype(np.complex128))
    VX.append(zz)
    VY.append(zz)
    VZ.append(zz)
    I, J = np.concatenate(I), np.concatenate(J)
    mk = lambda vals: _csc(I, J, np.concatenate(vals), N)  # noqa: E731
    return mk(V), mk(VX), mk(VY), mk(VZ), float(N)


def chain(N: int, *, t: float = 1.0, shift: float = 0.0):
    """Periodic 1-D chain with on-site `shift` and hopping -t."""
    i = np.arange(N)
    rows = np.concatenate([i, i, (i + 1) % N])
    cols = np.concatenate([i, (i + 1) % N, i])
    vals = np.concatenate([np.full(N, shift), np.full(N, -t), np.full(N, -t)]).astype(np.complex128)
    return _csc(rows, cols, vals, N)


def haldane_open_model(Lx: int, Ly: int, *, t: float = 1.0, t2: float = 0.2,
                       phi: float = math.pi / 2, m: float = 0.0):
    """Haldane model on an open Lx x Ly honeycomb flake (hops leaving the grid
    are dropped): returns (H, pos, cell_area) with pos the N x 2 site
    coordinates r_A = (x-1) a1 + (y-1) a2, r_B = r_A + dAB.  Sites: cell
    c = (y-1) + Ly (x-1), A = 2c, B = 2c + 1 (0-based)."""
    a1 = np.array([math.sqrt(3), 0.0])
    a2 = np.array([math.sqrt(3) / 2, 1.5])
    dAB = np.array([math.sqrt(3) / 2, 0.5])
    N = 2 * Lx * Ly
    X, Y = np.meshgrid(np.arange(1, Lx + 1), np.arange(1, Ly + 1), indexing="ij")
    X = X.ravel()
    Y = Y.ravel()
    cell = (Y - 1) + Ly * (X - 1)
    pos = np.zeros((N, 2))
    rA = (X - 1)[:, None] * a1[None, :] + (Y - 1)[:, None] * a2[None, :]
    pos[2 * cell] = rA
    pos[2 * cell + 1] = rA + dAB
    I, J, V = [], [], []

    def addhop(i, j, amp):
        amp = np.broadcast_to(np.asarray(amp, dtype=np.complex128), i.shape)
        I.extend([i, j])
        J.extend([j, i])
        V.extend([amp, np.conj(amp)])

    def inb(x, y):
        return (x >= 1) & (x <= Lx) & (y >= 1) & (y <= Ly)

    def sA(x, y):
        return 2 * ((y - 1) + Ly * (x - 1))

    addhop(sA(X, Y) + 1, sA(X, Y), -t + 0j)
    for xo, yo in ((1, 0), (0, 1)):
        k = inb(X + xo, Y + yo)
        addhop(sA(X[k], Y[k]) + 1, sA(X[k] + xo, Y[k] + yo), -t + 0j)
    for xo, yo in ((1, 0), (-1, 1), (0, -1)):  # a1, a2 - a1, -a2
        k = inb(X + xo, Y + yo)
        addhop(sA(X[k] + xo, Y[k] + yo), sA(X[k], Y[k]), t2 * np.exp(1j * phi))
        addhop(sA(X[k] + xo, Y[k] + yo) + 1, sA(X[k], Y[k]) + 1, t2 * np.exp(-1j * phi))
    I.extend([2 * cell, 2 * cell + 1])
    J.extend([2 * cell, 2 * cell + 1])
    V.extend([np.full(cell.shape, m + 0j), np.full(cell.shape, -m + 0j)])
    H = _csc(np.concatenate(I), np.concatenate(J), np.concatenate(V), N)
    return H, pos, abs(a1[0] * a2[1] - a1[1] * a2[0])


def ring_model(N: int, *, t: float = 1.0):
    """N-site periodic ring with hopping -t (real symmetric)."""
    i = np.arange(N)
    rows = np.concatenate([i, i])
    cols = np.concatenate([(i + 1) % N, (i - 1) % N])
    return _csc(rows, cols, np.full(2 * N, -float(t)), N)


def flux_ring_model(N: int, *, t: float = 1.0, phi: float = 0.35):
    """N-site periodic ring with h[i, i+1] = -t e^{i phi} and its conjugate."""
    i = np.arange(N)
    j = (i + 1) % N
    hop = -float(t) * np.exp(1j * phi)
    return _csc(np.concatenate([i, j]), np.concatenate([j, i]),
                np.concatenate([np.full(N, hop), np.full(N, np.conj(hop))]), N)


def square_disordered(L: int, *, seed: int = 42, eps: float = 1.0, dt: float = 0.1):
    """Periodic L x L square lattice, H_ii = eps_i ~ U[-eps, eps] and
    H_ij = H_ji = -t_ij with t_ij ~ 1 + U[-dt, dt] per bond (real)."""
    rng = Xoshiro256pp(seed)
    N = L * L
    x, y = np.meshgrid(np.arange(L), np.arange(L), indexing="ij")
    x = x.T.ravel()
    y = y.T.ravel()
    site = x + L * y
    right = (x + 1) % L + L * y
    up = x + L * ((y + 1) % L)
    tr = 1.0 + dt * (2 * rng.rand(N) - 1)
    tu = 1.0 + dt * (2 * rng.rand(N) - 1)
    e = eps * (2 * rng.rand(N) - 1)
    rows = np.concatenate([site, right, site, up, site])
    cols = np.concatenate([right, site, up, site, site])
    vals = np.concatenate([-tr, -tr, -tu, -tu, e])
    return _csc(rows, cols, vals, N)


def helix_cluster(*, t: float = 1.0, t2: float = 0.45, phi: float = 0.7, w: float = 0.3):
    """Six sites on a helix (angle 2 pi j/3, height 0.4 j), hopping -t,
    second-neighbour t2 e^{i phi}, staggered on-site w (-1)^j (1-based j);
    J_alpha[i, j] = H[i, j] (r_i - r_j)_alpha (test/cpge_test.jl:10-35).
    Returns (H, [Jx, Jy, Jz], positions)."""
    N = 6
    pos = np.array([[math.cos(2 * math.pi * j / 3), math.sin(2 * math.pi * j / 3), 0.4 * j]
                    for j in range(N)])
    H = np.zeros((N, N), dtype=np.complex128)
    for j in range(N - 1):
        H[j + 1, j] += -t
        H[j, j + 1] += -t
    for j in range(N - 2):
        H[j + 2, j] += t2 * np.exp(1j * phi)
        H[j, j + 2] += t2 * np.exp(-1j * phi)
    for j in range(N):
        H[j, j] = w * (-1) ** (j + 1)
    J = [sp.csc_matrix(H * (pos[:, a][:, None] - pos[:, a][None, :])) for a in range(3)]
    return sp.csc_matrix(H), J, pos


def square_model(Lx: int, Ly: int, *, t: float = 1.0):
    """Periodic Lx x Ly square lattice, hopping -t, site i = ix + Lx iy at
    (ix, iy); returns (h, pos, disp) with disp(i, j) the minimum-image
    r_i - r_j (test/ed_reference.jl:653-690)."""
    if not (Lx > 0 and Ly > 0):
        raise ValueError("square_model: Lx and Ly must be positive")
    N = Lx * Ly
    ix, iy = np.meshgrid(np.arange(Lx), np.arange(Ly), indexing="ij")
    ix, iy = ix.T.ravel(), iy.T.ravel()  # site order i = ix + Lx iy
    i = ix + Lx * iy
    pos = np.zeros((N, 2))
    pos[i, 0], pos[i, 1] = ix, iy
    H = sp.lil_matrix((N, N))
    for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        j = (ix + dx) % Lx + Lx * ((iy + dy) % Ly)
        H[i, j] = -float(t)
    h = sp.csc_matrix(H)

    def disp(a, b):
        dx = pos[a, 0] - pos[b, 0]
        dy = pos[a, 1] - pos[b, 1]
        if dx > Lx / 2:
            dx -= Lx
        if dx < -Lx / 2:
            dx += Lx
        if dy > Ly / 2:
            dy -= Ly
        if dy < -Ly / 2:
            dy += Ly
        return np.array([dx, dy])

    return h, pos, disp


def ring_geometry(N: int):
    """Coordinates (k, 0) and the minimum-image displacement of the N-site
    ring (test/ed_reference.jl:535-550)."""
    pos = np.stack([np.arange(N, dtype=np.float64), np.zeros(N)], axis=1)

    def disp(a, b):
        dx = pos[a, 0] - pos[b, 0]
        if dx > N / 2:
            dx -= N
        if dx < -N / 2:
            dx += N
        return np.array([dx, 0.0])

    return pos, disp
