"""Seeded synthetic eigenproblems, shared by the HIP path, the CPU oracle,
the tests and the benchmark.

Randomness is a counter-based SplitMix64 keyed by (seed, stream, index), so a
value never depends on how many values were drawn before it or on the
platform:

    key = mix(seed * G + stream + 1)
    h   = mix(key + (index + 1) * G)          (mod 2**64)
    uniform = (h >> 11) * 2**-53               in [0, 1)
    normal  = Box-Muller on uniforms 2i, 2i+1

Streams: 0 initial block, 1 diagonal potential, 2 stiffness entries, 3 mass
diagonal, 4 grid sides, 5 edge weights, 6 loss weights, 7 finite-difference
coordinates.

Matrices are CSR with int32 indices, natural (lexicographic) row ordering and
sorted columns.  Stencil operators follow the reference's test operators
(tests/openmp/laplace.hh:19-31, tests/openmp/diagonal.hh:15-28) and the
configurations named by the benchmark (2D/3D Laplacians, a structural
stiffness/mass pair, a weighted graph Laplacian).
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def rng_u64(seed: int, stream: int, index: np.ndarray) -> np.ndarray:
    idx = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _G + np.uint64(stream + 1))
        return _mix(key + (idx + np.uint64(1)) * _G)


def uniform(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    h = rng_u64(seed, stream, np.arange(offset, offset + count, dtype=np.uint64))
    return (h >> np.uint64(11)).astype(np.float64) * (2.0**-53)


def normal(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    i = np.arange(offset, offset + count, dtype=np.uint64)
    u1 = (rng_u64(seed, stream, 2 * i) >> np.uint64(11)).astype(np.float64) * (2.0**-53)
    u2 = (rng_u64(seed, stream, 2 * i + np.uint64(1)) >> np.uint64(11)).astype(np.float64) * (2.0**-53)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


@dataclass
class CSR:
    n: int
    rowptr: np.ndarray  # int32 [n+1]
    col: np.ndarray  # int32 [nnz]
    val: np.ndarray  # float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.rowptr.astype("<i4"), self.col.astype("<i4"), self.val.astype("<f8")):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))

    def to_dense(self) -> np.ndarray:
        A = np.zeros((self.n, self.n))
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        A[rows, self.col] = self.val
        return A


def _stencil_csr(shape: tuple[int, ...], offsets: list[tuple[int, ...]], values: list[np.ndarray | float]) -> CSR:
    """CSR of a grid stencil on a lexicographically ordered grid (first axis
    fastest).  offsets must be given in increasing order of linear offset so
    that columns come out sorted."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    coords = np.unravel_index(np.arange(n), shape[::-1])[::-1]  # coords[d] along axis d (axis 0 fastest)
    strides = [int(np.prod(shape[:d])) for d in range(len(shape))]
    cols, vals, valid = [], [], []
    for off, v in zip(offsets, values):
        ok = np.ones(n, dtype=bool)
        lin = np.zeros(n, dtype=np.int64)
        for d, o in enumerate(off):
            c = coords[d] + o
            ok &= (c >= 0) & (c < shape[d])
            lin += o * strides[d]
        cols.append(np.arange(n, dtype=np.int64) + lin)
        vals.append(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)))
        valid.append(ok)
    C = np.stack(cols, axis=1)
    Vv = np.stack(vals, axis=1)
    M = np.stack(valid, axis=1)
    counts = M.sum(axis=1)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return CSR(n, rowptr.astype(np.int32), C[M].astype(np.int32), np.ascontiguousarray(Vv[M]))


def laplacian_1d(n: int) -> CSR:
    """tridiag(-1, 2, -1), Dirichlet (tests/openmp/laplace.hh:19-31)."""
    return _stencil_csr((n,), [(-1,), (0,), (1,)], [-1.0, 2.0, -1.0])


def diagonal(n: int) -> CSR:
    """diag((i+1)^2) (tests/openmp/diagonal.hh:15-20)."""
    return CSR(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), (np.arange(1, n + 1, dtype=np.float64)) ** 2)


def laplacian_2d(nx: int, ny: int, potential: np.ndarray | None = None) -> CSR:
    """5-point Dirichlet Laplacian: diagonal 4 (+ potential), neighbours -1."""
    diag = 4.0 if potential is None else 4.0 + potential
    return _stencil_csr((nx, ny), [(0, -1), (-1, 0), (0, 0), (1, 0), (0, 1)], [-1.0, -1.0, diag, -1.0, -1.0])


def laplacian_3d(nx: int, ny: int, nz: int, potential: np.ndarray | None = None) -> CSR:
    """7-point Dirichlet Laplacian: diagonal 6 (+ potential), neighbours -1."""
    diag = 6.0 if potential is None else 6.0 + potential
    offs = [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    return _stencil_csr((nx, ny, nz), offs, [-1.0, -1.0, -1.0, diag, -1.0, -1.0, -1.0])


def laplacian_2d_eigenvalues(nx: int, ny: int) -> np.ndarray:
    """All eigenvalues of laplacian_2d(nx, ny), ascending:
    4 - 2cos(i pi/(nx+1)) - 2cos(j pi/(ny+1))."""
    i = np.arange(1, nx + 1)
    j = np.arange(1, ny + 1)
    ev = (2 - 2 * np.cos(i * np.pi / (nx + 1)))[:, None] + (2 - 2 * np.cos(j * np.pi / (ny + 1)))[None, :]
    return np.sort(ev.ravel())


def initial_block(n: int, b: int, seed: int = 42) -> np.ndarray:
    """V0 ~ N(0, 1), n x b row-major (stream 0)."""
    return normal(seed, 0, n * b).reshape(n, b)


# ---------------------------------------------------------------------------
# Weighted graph Laplacian + diagonal (the differentiation problem)


@dataclass
class WeightedGrid:
    nx: int
    ny: int
    weights: np.ndarray  # [n_edges] (horizontal edges first, then vertical)
    theta: np.ndarray  # [n]
    edges: np.ndarray  # [n_edges, 2] int64 (i < j)
    csr: CSR = field(default=None)  # type: ignore
    edge_slot: np.ndarray = field(default=None)  # type: ignore  # [n_edges, 2]: CSR slots of (i,j) and (j,i)


def weighted_grid(nx: int, ny: int, seed: int = 42) -> WeightedGrid:
    """A(theta, w) = L_w + diag(theta) on an nx x ny grid (5-point graph):
    (L_w)_ii = sum of incident weights, (L_w)_ij = -w_ij.  w ~ U[0.5, 2]
    (stream 5), theta ~ U[0, 1] (stream 1)."""
    n = nx * ny
    idx = np.arange(n).reshape(ny, nx)  # node (i, j) -> i + nx*j
    h = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], axis=1)
    v = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], axis=1)
    edges = np.concatenate([h, v]).astype(np.int64)
    ne = edges.shape[0]
    w = 0.5 + 1.5 * uniform(seed, 5, ne)
    theta = uniform(seed, 1, n)
    g = WeightedGrid(nx, ny, w, theta, edges)
    rebuild_weighted(g)
    return g


def rebuild_weighted(g: WeightedGrid) -> None:
    n = g.nx * g.ny
    e = g.edges
    ne = e.shape[0]
    deg = np.zeros(n)
    np.add.at(deg, e[:, 0], g.weights)
    np.add.at(deg, e[:, 1], g.weights)
    rows = np.concatenate([np.arange(n), e[:, 0], e[:, 1]])
    cols = np.concatenate([np.arange(n), e[:, 1], e[:, 0]])
    vals = np.concatenate([deg + g.theta, -g.weights, -g.weights])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    g.csr = CSR(n, rowptr.astype(np.int32), cols.astype(np.int32), np.ascontiguousarray(vals))
    inv = np.empty_like(order)
    inv[order] = np.arange(order.size)
    g.edge_slot = np.stack([inv[n : n + ne], inv[n + ne :]], axis=1)


# ---------------------------------------------------------------------------
# Structural stiffness/mass pair (the generalized problem)


def stiffness_mass(nx: int = 16, ny: int = 16, nz: int = 10, dof: int = 6, seed: int = 42) -> tuple[CSR, np.ndarray]:
    """K x = lambda M x shaped like a structural pair.

    Sparsity rule: nodes on an nx x ny x nz grid, `dof` DOFs per node,
    DOF (p, d) at index p*dof + d.  Each node has a dense dof x dof self
    block; the 26 neighbours of the 27-point node stencil form 13 offset
    pairs +-o_s (s = 0..12 in lexicographic order of the positive offsets),
    and DOF d of node p couples to DOF d of nodes p +- o_s iff
    (s + d) mod 13 < 6.  Off-diagonal values are N(0, 1) (stream 2, indexed
    by the upper-triangle entry), mirrored; the diagonal is the absolute row
    sum + 1, so K is strictly diagonally dominant and SPD.  M is a lumped
    mass diag(exp(0.5 z)), z ~ N(0, 1) (stream 3).
    """
    nn = nx * ny * nz
    n = nn * dof
    pos = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    lin = lambda o: o[0] + nx * (o[1] + ny * o[2])  # noqa: E731
    positive = sorted([o for o in pos if lin(o) > 0], key=lin)
    assert len(positive) == 13
    node = np.arange(nn)
    cx, cy, cz = node % nx, (node // nx) % ny, node // (nx * ny)
    ri, rj = [], []
    # self blocks: upper triangle d < d'
    for d in range(dof):
        for d2 in range(d + 1, dof):
            ri.append(node * dof + d)
            rj.append(node * dof + d2)
    for s, o in enumerate(positive):
        ok = (cx + o[0] >= 0) & (cx + o[0] < nx) & (cy + o[1] >= 0) & (cy + o[1] < ny) & (cz + o[2] >= 0) & (cz + o[2] < nz)
        p = node[ok]
        q = p + lin(o)
        for d in range(dof):
            if (s + d) % 13 < 6:
                ri.append(p * dof + d)
                rj.append(q * dof + d)
    I = np.concatenate(ri)
    J = np.concatenate(rj)
    order = np.lexsort((J, I))
    I, J = I[order], J[order]
    vals = normal(seed, 2, I.size)
    rows = np.concatenate([I, J])
    cols = np.concatenate([J, I])
    v = np.concatenate([vals, vals])
    absrow = np.zeros(n)
    np.add.at(absrow, rows, np.abs(v))
    rows = np.concatenate([rows, np.arange(n)])
    cols = np.concatenate([cols, np.arange(n)])
    v = np.concatenate([v, absrow + 1.0])
    o2 = np.lexsort((cols, rows))
    rows, cols, v = rows[o2], cols[o2], v[o2]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    K = CSR(n, rowptr.astype(np.int32), cols.astype(np.int32), np.ascontiguousarray(v))
    mass = np.exp(0.5 * normal(seed, 3, n))
    return K, mass


def fe_laplacian_2d(nx: int, ny: int) -> tuple[CSR, CSR]:
    """Bilinear finite elements for -Laplace u = lambda u on the unit square,
    Dirichlet, nx x ny interior nodes (first axis fastest): stiffness
    K = K1x (x) M1y + M1x (x) K1y and consistent mass M = M1x (x) M1y, with
    the 1D factors K1 = tridiag(-1, 2, -1) / h and M1 = h tridiag(1, 4, 1) / 6,
    h = 1 / (nodes + 1).  Both 9-point, SPD; M is a sparse (not diagonal) B."""
    if nx < 3 or ny < 3:
        raise ValueError("fe_laplacian_2d needs at least 3 nodes per side")

    def one(m):
        h = 1.0 / (m + 1)
        k = (np.diag(np.full(m, 2.0)) - np.diag(np.ones(m - 1), 1) - np.diag(np.ones(m - 1), -1)) / h
        mm = (np.diag(np.full(m, 4.0)) + np.diag(np.ones(m - 1), 1) + np.diag(np.ones(m - 1), -1)) * h / 6.0
        return k, mm

    kx, mx = one(nx)
    ky, my = one(ny)
    offs = [(dx, dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]

    def kron_stencil(ax, ay):
        # entry of offset (dx, dy) at node (i, j) = ax[i, i+dx] ay[j, j+dy]: the
        # 1D factors are Toeplitz, so it depends on the offset only (row 1 of
        # each factor holds all three values)
        return [float(ax[1, 1 + dx]) * float(ay[1, 1 + dy]) for dx, dy in offs]

    K = _stencil_csr((nx, ny), offs, [a + b for a, b in zip(kron_stencil(kx, my), kron_stencil(mx, ky))])
    M = _stencil_csr((nx, ny), offs, kron_stencil(mx, my))
    return K, M


def fe_laplacian_2d_eigenvalues(nx: int, ny: int) -> np.ndarray:
    """All eigenvalues of K x = lambda M x for fe_laplacian_2d(nx, ny),
    ascending: lambda_k + lambda_l with the 1D values
    lambda_k = (6 / h^2) (1 - cos t_k) / (2 + cos t_k), t_k = k pi / (m + 1)
    (the sine vectors diagonalise both 1D factors)."""
    def one(m):
        h = 1.0 / (m + 1)
        t = np.arange(1, m + 1) * np.pi / (m + 1)
        return 6.0 / h**2 * (1 - np.cos(t)) / (2 + np.cos(t))

    return np.sort((one(nx)[:, None] + one(ny)[None, :]).ravel())


def stiffness_sparse_mass(nx: int = 8, ny: int = 8, nz: int = 6, dof: int = 3, seed: int = 42, coupling: float = 0.5) -> tuple[CSR, CSR]:
    """The structural pair of stiffness_mass with a sparse SPD mass B on the
    stiffness pattern: off-diagonal B_ij = coupling sqrt(m_i m_j) s_ij /
    (row length), s_ij ~ U[-1, 1] (stream 8, by upper-triangle entry,
    mirrored), diagonal m_i + sum_j |B_ij| (strictly diagonally dominant)."""
    K, m = stiffness_mass(nx, ny, nz, dof, seed)
    n = K.n
    rows = np.repeat(np.arange(n), np.diff(K.rowptr))
    cols = K.col.astype(np.int64)
    upper = cols > rows
    iu, ju = rows[upper], cols[upper]
    s = 2.0 * uniform(seed, 8, iu.size) - 1.0
    rl = np.diff(K.rowptr).astype(np.float64)
    v = coupling * np.sqrt(m[iu] * m[ju]) * s / np.maximum(rl[iu], rl[ju])
    Bv = np.zeros(K.nnz)
    # positions of the upper entries and of their mirrors in K's pattern
    key = rows * n + cols
    pos_u = np.flatnonzero(upper)
    pos_l = np.searchsorted(key, ju * n + iu)
    Bv[pos_u] = v
    Bv[pos_l] = v
    absrow = np.zeros(n)
    np.add.at(absrow, rows, np.abs(Bv))
    diag = rows == cols
    Bv[diag] = m + absrow
    return K, CSR(n, K.rowptr.copy(), K.col.copy(), Bv)


# ---------------------------------------------------------------------------
# Named configurations


@dataclass
class Config:
    name: str
    csr: CSR
    b: int
    nev: int
    ncv: int
    tol: float = 1e-8
    order: int = 1  # 0 largest magnitude (reference default), 1 ascending, 2 descending
    mass: np.ndarray | None = None
    exact: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def config_c1() -> Config:
    """2D 5-point Laplacian 128 x 128, b=4, 8 smallest, ncv=256."""
    A = laplacian_2d(128, 128)
    return Config("laplace2d_128", A, b=4, nev=8, ncv=256, exact=laplacian_2d_eigenvalues(128, 128)[:8])


def config_c2() -> Config:
    """K x = lambda M x, n = 15360, b = 6, 20 smallest, ncv = 240 (40 blocks:
    13 restart cycles instead of 54 with 20 blocks, the GPU solve ~1.9x
    faster; both the CPU oracle and the GPU use it)."""
    K, mass = stiffness_mass()
    return Config("stiffness_mass_15360", K, b=6, nev=20, ncv=240, mass=mass)


def config_c3(side: int = 160, seed: int = 42) -> Config:
    """3D 7-point Laplacian side^3 + diag U[0,1], b=8, 32 smallest, ncv=256."""
    n = side**3
    pot = uniform(seed, 1, n)
    A = laplacian_3d(side, side, side, pot)
    return Config(f"laplace3d_{side}", A, b=8, nev=32, ncv=256)


def config_c4_sides(count: int = 64, seed: int = 42) -> np.ndarray:
    return 48 + (uniform(seed, 4, count) * (256 - 48 + 1)).astype(np.int64)


def config_c4(count: int = 64, seed: int = 42) -> list[Config]:
    """count 2D Laplacians, side ~ U{48..256}, potential U[0, 2], b=4, 10
    smallest, ncv = 128 (32 blocks; the batch then needs 33 restart cycles,
    against 311 with 12 blocks, and solves the 64 problems ~1.6x faster)."""
    out = []
    off = 0
    for p, s in enumerate(config_c4_sides(count, seed)):
        s = int(s)
        pot = 2.0 * uniform(seed, 1, s * s, offset=off)
        off += s * s
        out.append(Config(f"batch{p}_side{s}", laplacian_2d(s, s, pot), b=4, nev=10, ncv=128))
    return out
