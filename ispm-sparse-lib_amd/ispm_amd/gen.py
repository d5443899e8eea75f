"""Seeded synthetic workloads (host, numpy).

Random numbers come from a SplitMix64 stream: value i (1-based) of stream `seed` is
mix(seed + i * 0x9E3779B97F4A7C15); uniform doubles take the top 53 bits.  Every generator
is deterministic in its seed.

Workloads (row counts and nnz are exact):
  poisson7(32)            7-point Dirichlet Poisson, diagonal 6 + 1e-3, off-diagonals -1,
                          off-grid neighbours dropped: n = 32,768, nnz = 223,232.
  fv7_lognormal(128)      7-point finite-volume matrix on a 128^3 hex mesh, face
                          conductances lognormal(0, 1.5), Dirichlet boundary faces add their
                          conductance to the diagonal; rows randomly permuted and then
                          RCM-reordered: n = 2,097,152, nnz = 14,581,760.
  stencil27_contrast(192) 27-point stencil, coefficient kappa in {1, 1e4} per random 8^3
                          block, a_ij = -(kappa_i + kappa_j)/2, Dirichlet neighbours add
                          kappa_i to the diagonal: n = 7,077,888, nnz = 189,119,224.  Every
                          value is a multiple of 1/2 below 2^19, so FP32 storage is exact.
  rgg_laplacian(n, k)     graph Laplacian of a 3-D random geometric k-nearest-neighbour
                          graph plus 1e-2 I.
"""
from __future__ import annotations

import numpy as np

from . import CSR

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def uniform(seed: int, n: int, lo=-1.0, hi=1.0) -> np.ndarray:
    return lo + (hi - lo) * uniform01(seed, n)


def normal(seed: int, n: int) -> np.ndarray:
    """Box-Muller on consecutive pairs of the stream."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:n]


def _stencil_csr(nx, ny, nz, offsets, val_fn, dtype=np.float64, diag=None, chunk=1 << 15):
    """CSR of a structured-grid stencil.  offsets: list of (dx,dy,dz); val_fn(mask, rows, off)
    -> values of one offset for the given rows (one per row; only the valid entries, `mask`,
    are kept).  Offsets are visited in increasing linear offset, so rows come out sorted by
    column.  diag: an array that val_fn accumulates into for its rows; when given, it becomes
    the (0,0,0) entries once every offset of those rows has been visited.  Rows are built in
    independent chunks on several threads; each row's offsets are still visited in order, so
    the result does not depend on the chunking."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    n = nx * ny * nz
    lin = [dz * nx * ny + dy * nx + dx for dx, dy, dz in offsets]
    order = np.argsort(lin, kind="stable")
    offs = [offsets[o] for o in order]
    lins = [lin[o] for o in order]
    K = len(offs)
    center = offs.index((0, 0, 0)) if diag is not None else None

    def masks(r0, r1):
        rows = np.arange(r0, r1, dtype=np.int64)
        x, y, z = rows % nx, (rows // nx) % ny, rows // (nx * ny)
        mask = np.empty((r1 - r0, K), bool)
        for j, (dx, dy, dz) in enumerate(offs):
            mask[:, j] = (x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny) & (z + dz >= 0) & (z + dz < nz)
        return rows, mask

    bounds = [(r0, min(n, r0 + chunk)) for r0 in range(0, n, chunk)]
    row_ptr = np.zeros(n + 1, np.int64)
    threads = max(1, min(16, os.cpu_count() or 1))

    def count(b):
        _, mask = masks(*b)
        row_ptr[b[0] + 1:b[1] + 1] = mask.sum(1)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(count, bounds))
    np.cumsum(row_ptr, out=row_ptr)
    c = np.empty(int(row_ptr[-1]), np.int32)
    v = np.empty(int(row_ptr[-1]), dtype)

    def fill(b):
        rows, mask = masks(*b)
        cols = np.empty(mask.shape, np.int32)
        vals = np.empty(mask.shape, dtype)
        for j, (o, ln) in enumerate(zip(offs, lins)):
            cols[:, j] = rows + ln
            vals[:, j] = val_fn(mask[:, j], rows, o)
        if center is not None:
            vals[:, center] = diag[rows]
        c[row_ptr[b[0]]:row_ptr[b[1]]] = cols[mask]
        v[row_ptr[b[0]]:row_ptr[b[1]]] = vals[mask]

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(fill, bounds))
    return CSR(n, n, row_ptr.astype(np.int32), c, v)


def poisson7(m: int = 32, shift: float = 1e-3) -> CSR:
    offs = [(0, 0, -1), (0, -1, 0), (-1, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]

    def val(mask, idx, o):
        return np.where(mask, (6.0 + shift) if o == (0, 0, 0) else -1.0, 0.0)

    return _stencil_csr(m, m, m, offs, val)


def stencil27_contrast(m: int = 192, block: int = 8, contrast: float = 1e4, seed: int = 3001) -> CSR:
    nb = m // block
    kb = np.where(uniform01(seed, nb ** 3) < 0.5, 1.0, contrast)
    idx = np.arange(m ** 3, dtype=np.int64)
    x = idx % m
    y = (idx // m) % m
    z = idx // (m * m)
    kappa = kb[(x // block) + nb * ((y // block) + nb * (z // block))]
    del x, y, z
    offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    diag = np.zeros(m ** 3)

    def val(mask, rows, o):
        if o == (0, 0, 0):
            return np.zeros(rows.size)
        lin = o[2] * m * m + o[1] * m + o[0]
        nbr = np.clip(rows + lin, 0, m ** 3 - 1)
        ki = kappa[rows]
        kn = np.where(mask, kappa[nbr], ki)           # off-grid neighbour: Dirichlet, kappa_i
        c = 0.5 * (ki + kn)
        diag[rows] += c
        return np.where(mask, -c, 0.0)

    return _stencil_csr(m, m, m, offs, val, diag=diag)


def stencil(nx: int, ny: int, nz: int, points: int = 7, anisotropy: float = 1.0,
            contrast: float = 1.0, block: int = 4, seed: int = 0, shift: float = 1e-3) -> CSR:
    """SPD structured-grid operator with a 7-, 19- or 27-point stencil: a weighted graph
    Laplacian, edge weight = offset class factor (face 1, edge 1/2, corner 1/4) times z
    anisotropy^|dz| times the harmonic mean of the two cells' coefficients, the coefficients
    log-uniform in [1, contrast] on random block^3 blocks; diagonal = sum of the row's edge
    weights (Dirichlet boundary: edges leaving the grid count) plus shift."""
    n = nx * ny * nz
    offs = []
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dx) + abs(dy) + abs(dz)
                if k == 0 or (points == 7 and k > 1) or (points == 19 and k > 2):
                    continue
                offs.append((dx, dy, dz))
    offs.append((0, 0, 0))
    bx, by, bz = -(-nx // block), -(-ny // block), -(-nz // block)
    kb = np.exp(np.log(max(contrast, 1.0)) * uniform01(seed, bx * by * bz))
    idx = np.arange(n, dtype=np.int64)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    kappa = kb[(x // block) + bx * ((y // block) + by * (z // block))]
    del x, y, z
    diag = np.full(n, float(shift))
    cls = {1: 1.0, 2: 0.5, 3: 0.25}

    def val(mask, rows, o):
        if o == (0, 0, 0):
            return np.zeros(rows.size)
        lin = o[2] * nx * ny + o[1] * nx + o[0]
        nbr = np.clip(rows + lin, 0, n - 1)
        ki = kappa[rows]
        kn = np.where(mask, kappa[nbr], ki)
        # harmonic mean from the symmetric product and sum, so that w(i, j) == w(j, i) exactly
        hm = 2 * (ki * kn) / (ki + kn)
        w = cls[abs(o[0]) + abs(o[1]) + abs(o[2])] * anisotropy ** abs(o[2]) * hm
        diag[rows] += w
        return np.where(mask, -w, 0.0)

    return _stencil_csr(nx, ny, nz, offs, val, diag=diag)


def fv7_lognormal(m: int = 128, sigma: float = 1.5, seed: int = 2001, permute_seed: int = 2002,
                  rcm: bool = True) -> CSR:
    import scipy.sparse as sp
    n = m ** 3
    # conductance per face: x-faces (m+1)*m*m incl. the two boundary layers, same for y, z
    cx = np.exp(sigma * normal(seed, (m + 1) * m * m)).reshape(m, m, m + 1)      # [z, y, xface]
    cy = np.exp(sigma * normal(seed + 1, (m + 1) * m * m)).reshape(m, m + 1, m)  # [z, yface, x]
    cz = np.exp(sigma * normal(seed + 2, (m + 1) * m * m)).reshape(m + 1, m, m)  # [zface, y, x]
    diag = (cx[:, :, :-1] + cx[:, :, 1:] + cy[:, :-1, :] + cy[:, 1:, :] + cz[:-1] + cz[1:]).ravel()
    idx = np.arange(n).reshape(m, m, m)
    rows, cols, vals = [np.arange(n)], [np.arange(n)], [diag]
    for a, b, c in ((idx[:, :, :-1], idx[:, :, 1:], cx[:, :, 1:-1]),
                    (idx[:, :-1, :], idx[:, 1:, :], cy[:, 1:-1, :]),
                    (idx[:-1], idx[1:], cz[1:-1])):
        a, b, c = a.ravel(), b.ravel(), c.ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [-c, -c]
    M = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    perm = np.argsort(uniform01(permute_seed, n), kind="stable")
    M = M[perm][:, perm].tocsr()
    if rcm:
        from scipy.sparse.csgraph import reverse_cuthill_mckee
        p = reverse_cuthill_mckee(M, symmetric_mode=True)
        M = M[p][:, p].tocsr()
    M.sort_indices()
    return CSR.from_scipy(M)


def ldu_mesh(nx: int, ny: int, nz: int, sigma: float = 1.0, seed: int = 4001, asymmetry: float = 0.0):
    """An OpenFOAM-style lduMatrix of a diffusion operator on an nx*ny*nz hexahedral mesh
    (cell x + nx (y + ny z)): internal faces in upper-triangular order (sorted by owner, then
    neighbour, owner < neighbour), face conductances lognormal(0, sigma), upper = -c, diagonal =
    the cell's face conductances plus 1 per boundary face (Dirichlet, so the matrix is SPD).
    asymmetry > 0 gives lower = -c (1 + asymmetry u), u ~ U[0,1) (no longer symmetric).
    Returns (n, lower_addr, upper_addr, diag, upper, lower or None)."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    own = np.concatenate([idx[:, :, :-1].ravel(), idx[:, :-1, :].ravel(), idx[:-1].ravel()])
    nei = np.concatenate([idx[:, :, 1:].ravel(), idx[:, 1:, :].ravel(), idx[1:].ravel()])
    order = np.lexsort((nei, own))
    own, nei = own[order].astype(np.int32), nei[order].astype(np.int32)
    nf = len(own)
    c = np.exp(sigma * normal(seed, nf))
    deg = np.bincount(own, minlength=n) + np.bincount(nei, minlength=n)
    diag = np.bincount(own, weights=c, minlength=n) + np.bincount(nei, weights=c, minlength=n) + (6 - deg)
    upper = -c
    lower = None if asymmetry <= 0 else -c * (1.0 + asymmetry * uniform01(seed + 1, nf))
    return n, own, nei, diag, upper, lower


def rgg_laplacian(n: int, k: int, seed: int, shift: float = 1e-2, dtype=np.float64) -> CSR:
    """Graph Laplacian of a 3-D random geometric kNN graph (symmetrised) plus shift*I."""
    import scipy.sparse as sp
    from scipy.spatial import cKDTree
    pts = uniform01(seed, 3 * n).reshape(n, 3)
    _, nb = cKDTree(pts).query(pts, k=k + 1)
    r = np.repeat(np.arange(n), k)
    c = nb[:, 1:].ravel()
    W = sp.csr_matrix((np.ones(r.size), (r, c)), shape=(n, n))
    W = ((W + W.T) > 0).astype(np.float64)
    L = sp.diags(np.asarray(W.sum(1)).ravel() + shift) - W
    L = L.tocsr()
    L.sort_indices()
    out = CSR.from_scipy(L)
    if dtype != np.float64:
        out.vals = out.vals.astype(dtype)
    return out


def random_sparse(n: int, seed: int, min_len=1, max_len=40, n_cols=None, integer=False,
                  pair_bias=0.0, spd=False) -> CSR:
    """Irregular test matrix: row lengths uniform in [min_len, max_len]; with pair_bias > 0
    that fraction of entries is emitted as aligned column pairs (2c, 2c+1)."""
    n_cols = n if n_cols is None else n_cols
    u = uniform01(seed, 4 * n * max_len + 4 * n)
    pos = 0
    rows, cols = [], []
    for r in range(n):
        L = min_len + int(u[pos] * (max_len - min_len + 1)); pos += 1
        cs = set()
        while len(cs) < min(L, n_cols):
            if pair_bias and u[pos] < pair_bias:
                c0 = 2 * int(u[pos + 1] * (n_cols // 2))
                cs.add(c0)
                if c0 + 1 < n_cols:
                    cs.add(c0 + 1)
                pos += 2
            else:
                cs.add(int(u[pos] * n_cols)); pos += 1
            if pos + 4 >= u.size:
                u = np.concatenate([u, uniform01(seed + 7919, u.size, offset=u.size)])
        cs = sorted(cs)[:max(L, 1)]
        rows += [r] * len(cs)
        cols += cs
    rows = np.array(rows, np.int64)
    cols = np.array(cols, np.int64)
    vals = uniform(seed + 1, rows.size)
    if integer:
        vals = np.round(vals * 8)
    import scipy.sparse as sp
    M = sp.csr_matrix((vals, (rows, cols)), shape=(n, n_cols))
    if spd:
        M = M + M.T
        M = M + sp.diags(np.asarray(abs(M).sum(1)).ravel() + 1.0)
    M = M.tocsr()
    M.sort_indices()
    return CSR.from_scipy(M)


def banded_pairs(nodes: int, bw: int = 40, per: int = 5, seed: int = 0, far_every: int = 0) -> CSR:
    """Two unknowns per node (rows 2i, 2i+1; a 2-D elasticity-like coupling), each node
    coupled by full 2x2 blocks to itself and to `per` nodes drawn uniformly within `bw` of
    it: every row consists of aligned column pairs (2j, 2j+1), the structure the hblock
    encoder stores as pairs, and a slice's columns lie close to its rows.  far_every > 0:
    every such node also couples to the node nodes/2 away."""
    import scipy.sparse as sp
    u = uniform01(seed, nodes * per)
    i = np.repeat(np.arange(nodes, dtype=np.int64), per)
    j = np.clip(i + np.floor(u * (2 * bw + 1)).astype(np.int64) - bw, 0, nodes - 1)
    ii = [np.arange(nodes, dtype=np.int64), i]
    jj = [np.arange(nodes, dtype=np.int64), j]
    if far_every:
        f = np.arange(0, nodes, far_every, dtype=np.int64)
        ii.append(f)
        jj.append((f + nodes // 2) % nodes)
    ni, nj = np.concatenate(ii), np.concatenate(jj)
    # each node pair once (the blocks of repeated draws merge), then the 2x2 block entries
    key = np.unique(ni * nodes + nj)
    ni, nj = key // nodes, key % nodes
    r = np.concatenate([2 * ni + a for a in (0, 0, 1, 1)])
    c = np.concatenate([2 * nj + b for b in (0, 1, 0, 1)])
    v = uniform(seed + 1, r.size)
    M = sp.csr_matrix((v, (r, c)), shape=(2 * nodes, 2 * nodes))
    M.sort_indices()
    return CSR.from_scipy(M)
