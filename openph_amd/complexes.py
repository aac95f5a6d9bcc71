"""Synthetic filtrations and their Z/2 boundary matrices.

The reference receives its boundary matrices from Javaplex streams
(numerics/utils/stream2cmo.m:1-26, numerics/utils/example_factory.m:17-122)
as 1-based COO triplets in column-major order with rows ascending inside a
column (src/matlab/ph/openph.m:6, src/cuda/openph.cu:44). This module builds
the same format without Java, for the workloads the benchmarks use:

* Vietoris-Rips 2-skeleta of point clouds (all simplices of dimension <= 2,
  full radius). Simplices are ordered by filtration value (diameter), ties
  broken by dimension and then by the lexicographic vertex tuple.
* Lower-star filtrations of 3-D cubical complexes on an s^3 grid: all
  (2s-1)^3 cells, value of a cell = max over its vertices, ties broken by
  cell dimension and then by the cell's linear index.

Every generator returns a ``Boundary`` with 0-based CSC arrays (``colptr``,
``rows``) and helpers that emit the reference's 1-based COO form.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "Boundary",
    "vr_2skeleton",
    "cubical_lower_star",
    "points_cube",
    "points_torus",
    "points_sphere",
    "points_gmm",
    "smoothed_field",
    "vr40",
    "torus100",
    "cubical64",
    "batched_mix",
    "served_request",
    "served_input",
    "random_simplicial",
]


@dataclasses.dataclass
class Boundary:
    """A Z/2 boundary matrix in 0-based CSC form.

    ``colptr`` has m+1 entries; the rows of column j are
    ``rows[colptr[j]:colptr[j+1]]`` in strictly ascending order, and every row
    index is smaller than j (a filtration order). ``dim`` is the geometric
    dimension of every cell and ``value`` its filtration value.
    """

    colptr: np.ndarray
    rows: np.ndarray
    dim: np.ndarray
    value: np.ndarray
    name: str = ""

    @property
    def m(self) -> int:
        return int(self.colptr.shape[0] - 1)

    @property
    def nnz(self) -> int:
        return int(self.rows.shape[0])

    def coo1(self):
        """1-based (rows, cols) int32 COO in column-major order, as the
        reference's ``ph`` mex receives them (src/matlab/ph/ph.cu:34-40)."""
        counts = np.diff(self.colptr)
        cols = np.repeat(np.arange(self.m, dtype=np.int32), counts)
        return (self.rows.astype(np.int32) + 1), (cols + 1).astype(np.int32)

    def column(self, j: int) -> np.ndarray:
        return self.rows[self.colptr[j]:self.colptr[j + 1]]


def _csc_from_lists(nnz_per_col: np.ndarray, rows: np.ndarray) -> np.ndarray:
    colptr = np.zeros(nnz_per_col.shape[0] + 1, dtype=np.int64)
    np.cumsum(nnz_per_col, out=colptr[1:])
    return colptr


# ---------------------------------------------------------------------------
# Point clouds
# ---------------------------------------------------------------------------

def points_cube(rng: np.random.Generator, n: int) -> np.ndarray:
    return rng.uniform(0.0, 1.0, size=(n, 3))


def points_torus(rng: np.random.Generator, n: int, R: float = 1.0,
                 r: float = 0.4, sigma: float = 0.03) -> np.ndarray:
    theta = rng.uniform(0.0, 2.0 * math.pi, size=n)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=n)
    x = (R + r * np.cos(phi)) * np.cos(theta)
    y = (R + r * np.cos(phi)) * np.sin(theta)
    z = r * np.sin(phi)
    pts = np.stack([x, y, z], axis=1)
    return pts + rng.normal(0.0, sigma, size=(n, 3))


def points_sphere(rng: np.random.Generator, n: int, sigma: float = 0.05) -> np.ndarray:
    g = rng.normal(0.0, 1.0, size=(n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    return g + rng.normal(0.0, sigma, size=(n, 3))


def points_gmm(rng: np.random.Generator, n: int, k: int = 5,
               sigma: float = 0.1) -> np.ndarray:
    centers = rng.uniform(0.0, 1.0, size=(k, 3))
    labels = rng.integers(0, k, size=n)
    return centers[labels] + rng.normal(0.0, sigma, size=(n, 3))


# ---------------------------------------------------------------------------
# Vietoris-Rips 2-skeleton
# ---------------------------------------------------------------------------

def vr_2skeleton(points: np.ndarray, name: str = "") -> Boundary:
    """Full-radius VR 2-skeleton of ``points``.

    m = N + C(N,2) + C(N,3) and nnz = 2 C(N,2) + 3 C(N,3).
    """
    pts = np.asarray(points, dtype=np.float64)
    n = pts.shape[0]
    diff = pts[:, None, :] - pts[None, :, :]
    dist = np.sqrt((diff * diff).sum(-1))

    iu, ju = np.triu_indices(n, k=1)
    ne = iu.shape[0]
    # all triangles i<j<k
    tri = _triangles(n)
    nt = tri.shape[0]

    # Values: vertices 0, edges = length, triangles = max edge length.
    ev = dist[iu, ju]
    if nt:
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        tv = np.maximum(np.maximum(dist[a, b], dist[a, c]), dist[b, c])
    else:
        tv = np.zeros(0)

    m = n + ne + nt
    value = np.concatenate([np.zeros(n), ev, tv])
    dim = np.concatenate([np.zeros(n, np.int64), np.ones(ne, np.int64),
                          np.full(nt, 2, np.int64)])
    # lexicographic key columns (vertex tuples padded with -1 so shorter
    # tuples never compare; dims differ anyway when lengths differ)
    k0 = np.concatenate([np.arange(n), iu, tri[:, 0] if nt else np.zeros(0, np.int64)])
    k1 = np.concatenate([np.full(n, -1), ju, tri[:, 1] if nt else np.zeros(0, np.int64)])
    k2 = np.concatenate([np.full(n, -1), np.full(ne, -1), tri[:, 2] if nt else np.zeros(0, np.int64)])
    order = np.lexsort((k2, k1, k0, dim, value))
    pos = np.empty(m, dtype=np.int64)
    pos[order] = np.arange(m)

    # edge id of (i,j), i<j, within the edge block
    def edge_id(i, j):
        # index in triu_indices order
        return i * n - i * (i + 1) // 2 + (j - i - 1)

    # boundary rows of each original cell (by position)
    nnzc = np.concatenate([np.zeros(n, np.int64), np.full(ne, 2), np.full(nt, 3)])
    rows_e = np.stack([pos[iu], pos[ju]], axis=1)
    if nt:
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        rows_t = np.stack([pos[n + edge_id(b, c)], pos[n + edge_id(a, c)],
                           pos[n + edge_id(a, b)]], axis=1)
    # reorder columns by filtration position
    nnz_sorted = nnzc[order]
    colptr = _csc_from_lists(nnz_sorted, None)
    rows = np.empty(int(colptr[-1]), dtype=np.int32)
    # edges
    e_cols = pos[n:n + ne]
    re = np.sort(rows_e, axis=1)
    base = colptr[e_cols]
    rows[base] = re[:, 0]
    rows[base + 1] = re[:, 1]
    if nt:
        t_cols = pos[n + ne:]
        rt = np.sort(rows_t, axis=1)
        base = colptr[t_cols]
        rows[base] = rt[:, 0]
        rows[base + 1] = rt[:, 1]
        rows[base + 2] = rt[:, 2]
    return Boundary(colptr=colptr, rows=rows, dim=dim[order].astype(np.int32),
                    value=value[order], name=name or f"vr{n}")


def _triangles(n: int):
    # vectorised enumeration of i<j<k
    if n < 3:
        return np.zeros((0, 3), dtype=np.int64)
    out = []
    for i in range(n - 2):
        j, k = np.triu_indices(n - i - 1, k=1)
        blk = np.empty((j.shape[0], 3), dtype=np.int64)
        blk[:, 0] = i
        blk[:, 1] = j + i + 1
        blk[:, 2] = k + i + 1
        out.append(blk)
    return np.concatenate(out, axis=0)


# ---------------------------------------------------------------------------
# Cubical lower-star filtration
# ---------------------------------------------------------------------------

def smoothed_field(rng: np.random.Generator, side: int, sigma: float = 2.0) -> np.ndarray:
    """i.i.d. N(0,1) float32 grid smoothed by a Gaussian of ``sigma`` voxels
    (scipy.ndimage.gaussian_filter, mode 'reflect', truncate 4.0)."""
    from scipy.ndimage import gaussian_filter

    f = rng.standard_normal((side, side, side), dtype=np.float32)
    return gaussian_filter(f, sigma=sigma, mode="reflect", truncate=4.0).astype(np.float32)


def cubical_lower_star(field: np.ndarray, name: str = "") -> Boundary:
    """All cells of the cubical complex on a 3-D grid of vertex values.

    A cell is addressed by its coordinates (c0,c1,c2) in [0, 2s-2]^3 (even
    coordinate = vertex position, odd = spans an edge along that axis); its
    linear index is (c0*(2s-1)+c1)*(2s-1)+c2 and its dimension the number of
    odd coordinates. Value = max of the field over the cell's vertices.
    """
    f = np.asarray(field)
    s0, s1, s2 = f.shape
    n0, n1, n2 = 2 * s0 - 1, 2 * s1 - 1, 2 * s2 - 1
    # dilate by max along each axis
    a = np.empty((n0, s1, s2), dtype=f.dtype)
    a[0::2] = f
    a[1::2] = np.maximum(f[:-1], f[1:])
    b = np.empty((n0, n1, s2), dtype=f.dtype)
    b[:, 0::2] = a
    b[:, 1::2] = np.maximum(a[:, :-1], a[:, 1:])
    g = np.empty((n0, n1, n2), dtype=f.dtype)
    g[:, :, 0::2] = b
    g[:, :, 1::2] = np.maximum(b[:, :, :-1], b[:, :, 1:])

    c0 = (np.arange(n0) % 2).astype(np.int8)
    c1 = (np.arange(n1) % 2).astype(np.int8)
    c2 = (np.arange(n2) % 2).astype(np.int8)
    dimg = (c0[:, None, None] + c1[None, :, None] + c2[None, None, :]).astype(np.int64)

    m = n0 * n1 * n2
    value = g.reshape(-1)
    dim = dimg.reshape(-1)
    index = np.arange(m, dtype=np.int64)
    order = np.lexsort((index, dim, value))
    pos = np.empty(m, dtype=np.int64)
    pos[order] = index

    nnz_per = 2 * dim[order]
    colptr = _csc_from_lists(nnz_per, None)
    X, Y, Z = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    X = X.reshape(-1); Y = Y.reshape(-1); Z = Z.reshape(-1)
    strides = (n1 * n2, n2, 1)
    coords = (X, Y, Z)
    # each odd coordinate contributes two facets
    facet_lists = []
    for ax in range(3):
        odd = (coords[ax] % 2) == 1
        cells = index[odd]
        st = strides[ax]
        facet_lists.append((cells, cells - st))
        facet_lists.append((cells, cells + st))
    cells_all = np.concatenate([c for c, _ in facet_lists])
    facets_all = np.concatenate([f_ for _, f_ in facet_lists])
    colpos = pos[cells_all]
    rowpos = pos[facets_all]
    # sort by (column, row) and scatter
    o = np.lexsort((rowpos, colpos))
    rows = rowpos[o].astype(np.int32)
    assert rows.shape[0] == colptr[-1]
    return Boundary(colptr=colptr, rows=rows, dim=dim[order].astype(np.int32),
                    value=value[order].astype(np.float64), name=name or f"cubical{s0}")


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------

def vr40(seed: int = 0) -> Boundary:
    """40 points i.i.d. U[0,1]^3, full VR 2-skeleton (m = 10,700)."""
    rng = np.random.default_rng(seed)
    return vr_2skeleton(points_cube(rng, 40), name="vr40")


def torus100(seed: int = 0) -> Boundary:
    """100 points on a torus (R=1, r=0.4) with N(0, 0.03^2) noise, full VR
    2-skeleton (m = 166,750)."""
    rng = np.random.default_rng(seed)
    return vr_2skeleton(points_torus(rng, 100), name="torus100")


def cubical64(seed: int = 0, side: int = 64) -> Boundary:
    """Lower-star filtration of a sigma=2 smoothed N(0,1) field on side^3
    (m = (2 side - 1)^3)."""
    rng = np.random.default_rng(seed)
    return cubical_lower_star(smoothed_field(rng, side), name=f"cubical{side}")


def batched_mix(count: int = 512, seed: int = 0, nmin: int = 20, nmax: int = 80):
    """``count`` VR 2-skeleta; N ~ U{nmin..nmax}; shape chosen uniformly from
    {U[0,1]^3, S^2 + N(0,0.05^2), 5-component Gaussian mixture}. Child seeds
    come from ``SeedSequence(seed).spawn(count)``."""
    out = []
    for i, child in enumerate(np.random.SeedSequence(seed).spawn(count)):
        rng = np.random.default_rng(child)
        n = int(rng.integers(nmin, nmax + 1))
        shape = int(rng.integers(0, 3))
        if shape == 0:
            pts = points_cube(rng, n)
        elif shape == 1:
            pts = points_sphere(rng, n, 0.05)
        else:
            pts = points_gmm(rng, n)
        out.append(vr_2skeleton(pts, name=f"mix{i}_n{n}_s{shape}"))
    return out


def served_input(rng: np.random.Generator, vr_range=(30, 200), side_range=(16, 96),
                 p_vr: float = 0.7):
    """The raw input of one request of the serving mix: ("points", N x 3)
    for a VR 2-skeleton of N ~ U{30..200} uniform points (70%), or ("field",
    side^3 float32) for a cubical lower-star grid of side ~ U{16..96}."""
    if rng.uniform() < p_vr:
        n = int(rng.integers(vr_range[0], vr_range[1] + 1))
        return "points", points_cube(rng, n)
    side = int(rng.integers(side_range[0], side_range[1] + 1))
    return "field", smoothed_field(rng, side)


def served_request(rng: np.random.Generator, vr_range=(30, 200), side_range=(16, 96),
                   p_vr: float = 0.7) -> Boundary:
    """One request of the serving mix as a boundary matrix (served_input,
    built on the host)."""
    kind, x = served_input(rng, vr_range, side_range, p_vr)
    if kind == "points":
        return vr_2skeleton(x, name=f"vr{x.shape[0]}")
    return cubical_lower_star(x, name=f"cubical{x.shape[0]}")


def random_simplicial(rng: np.random.Generator, n_vertices: int, p_edge: float = 0.5,
                      p_tri: float = 0.5, p_tet: float = 0.5) -> Boundary:
    """A random simplicial complex (flag-free: each candidate simplex whose
    faces are all present is kept with a probability) with a random
    filtration order compatible with faces. Dimensions up to 3."""
    verts = [(i,) for i in range(n_vertices)]
    present = set(verts)
    simplices = list(verts)
    import itertools
    for k, p in ((2, p_edge), (3, p_tri), (4, p_tet)):
        for combo in itertools.combinations(range(n_vertices), k):
            faces = [tuple(x for x in combo if x != v) for v in combo]
            if all(f in present for f in faces) and rng.uniform() < p:
                present.add(combo)
                simplices.append(combo)
    # random filtration value = max(face values) + positive random increment
    val = {}
    for s in simplices:
        if len(s) == 1:
            val[s] = float(rng.uniform())
        else:
            faces = [tuple(x for x in s if x != v) for v in s]
            val[s] = max(val[f] for f in faces) + float(rng.uniform(0.0, 0.5)) * (rng.uniform() < 0.8)
    order = sorted(simplices, key=lambda s: (val[s], len(s), s))
    pos = {s: i for i, s in enumerate(order)}
    counts = np.array([0 if len(s) == 1 else len(s) for s in order], dtype=np.int64)
    colptr = _csc_from_lists(counts, None)
    rows = np.empty(int(colptr[-1]), dtype=np.int32)
    for j, s in enumerate(order):
        if len(s) > 1:
            faces = sorted(pos[tuple(x for x in s if x != v)] for v in s)
            rows[colptr[j]:colptr[j + 1]] = faces
    dim = np.array([len(s) - 1 for s in order], dtype=np.int32)
    value = np.array([val[s] for s in order])
    return Boundary(colptr=colptr, rows=rows, dim=dim, value=value, name="random")


def random_upper(rng: np.random.Generator, m: int, kmax: int = 5) -> Boundary:
    """A random strictly upper-triangular Z/2 matrix with 0..kmax entries per
    column. It is not a boundary matrix (its square is not zero), so only the
    algorithms that do not clear columns apply to it: standard and the row
    algorithms. Columns with one low then often have different lengths, which
    tells ph_row (neighbours of its pivot's pseudo-dimension only) from
    ph_row_parallel."""
    counts = np.array([rng.integers(0, min(j, kmax) + 1) for j in range(m)], dtype=np.int64)
    colptr = _csc_from_lists(counts, None)
    rows = np.empty(int(colptr[-1]), dtype=np.int32)
    for j in range(m):
        if counts[j]:
            rows[colptr[j]:colptr[j + 1]] = np.sort(rng.choice(j, size=int(counts[j]),
                                                               replace=False))
    return Boundary(colptr=colptr, rows=rows, dim=np.maximum(counts - 1, 0).astype(np.int32),
                    value=np.arange(m, dtype=np.float64), name=f"upper{m}")
