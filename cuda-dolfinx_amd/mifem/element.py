"""Reference elements and quadrature (the Basix stand-in).

Elements
--------
* ``Lagrange(tetrahedron, k)``, k = 1..3, equispaced nodes.  Local order is
  the Basix entity order: 4 vertices, then edges (e0=(2,3), e1=(1,3),
  e2=(1,2), e3=(0,3), e4=(0,2), e5=(0,1); nodes ordered from the first to
  the second vertex), then faces (face i opposite vertex i), then interior.
* ``DG(tetrahedron, k)``: same nodes, no inter-cell sharing.
* ``Lagrange(hexahedron, k)``, k = 1..3, Gauss-Lobatto-Legendre (GLL) nodes
  in lexicographic tensor order (index a + (k+1) b + (k+1)^2 c).
* ``Lagrange(triangle, k)`` / ``DG(triangle, k)``, k = 1..3, equispaced
  nodes in the Basix order: 3 vertices, edges e0=(1,2), e1=(0,2), e2=(0,1)
  (nodes from the first to the second vertex), then interior.  Tables carry
  two reference derivatives.

Each node also carries an *integer* position: barycentric multi-indices
(tetrahedra, sum = k) or tensor indices (hexahedra).  The function-space
layer turns those into global lattice keys, so shared nodes are identified
exactly.

The basis is built by inverting the Vandermonde matrix of the monomials
(P_k or Q_k) at the nodes; tables are evaluated in float64.

Quadrature
----------
* tetrahedra: centroid (degree 1), Hammer-Stroud 4 points (degree 2),
  otherwise the collapsed (Stroud conical) Gauss-Jacobi product with
  m = ceil((d+1)/2) points per direction;
* triangles (tetrahedron facets): centroid, 3-point degree-2 rule,
  otherwise collapsed Gauss-Jacobi;
* hexahedra / quadrilaterals: Gauss-Legendre tensor products.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy.special import roots_jacobi, roots_legendre

TET_VERTICES = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.float64)
TRI_VERTICES = np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float64)
# Triangle facets (edges): facet i is opposite vertex i.
TRI_EDGES = [(1, 2), (0, 2), (0, 1)]
TET_EDGES = [(2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1)]
TET_FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]
# Outward unit normals of the reference tetrahedron facets (facet i opposite vertex i).
TET_FACET_NORMALS = np.array([[1, 1, 1] / np.sqrt(3.0), [-1, 0, 0], [0, -1, 0], [0, 0, -1]],
                             dtype=np.float64)


def gll_points(k: int) -> np.ndarray:
    """Gauss-Lobatto-Legendre points of degree k on [0, 1] (k + 1 points)."""
    if k == 1:
        return np.array([0.0, 1.0])
    # interior GLL points are the roots of P'_k, i.e. Jacobi(k-1, 1, 1)
    r, _ = roots_jacobi(k - 1, 1.0, 1.0)
    return np.concatenate([[0.0], np.sort((r + 1.0) / 2.0), [1.0]])


@dataclass(frozen=True)
class Element:
    family: str          # "P" (continuous Lagrange) or "DG"
    cell: str            # "tetrahedron" or "hexahedron"
    degree: int

    @property
    def discontinuous(self) -> bool:
        return self.family == "DG"

    @property
    def ndofs(self) -> int:
        k = self.degree
        if self.cell == "tetrahedron":
            return (k + 1) * (k + 2) * (k + 3) // 6
        if self.cell == "triangle":
            return (k + 1) * (k + 2) // 2
        return (k + 1) ** 3

    @property
    def tdim(self) -> int:
        return 2 if self.cell == "triangle" else 3

    # ------------------------------------------------------------------ nodes
    @property
    def node_indices(self) -> np.ndarray:
        """Integer node positions.

        tetrahedron / triangle: (nd, 4) / (nd, 3) barycentric multi-indices m
        (sum k); node X = sum_v m_v / k * vertex_v.
        hexahedron: (nd, 3) tensor indices (a, b, c) in 0..k.
        """
        return _node_indices(self.cell, self.degree)

    @property
    def nodes(self) -> np.ndarray:
        """(nd, tdim) reference coordinates of the nodes."""
        idx = self.node_indices
        k = self.degree
        if self.cell in ("tetrahedron", "triangle"):
            verts = TET_VERTICES if self.cell == "tetrahedron" else TRI_VERTICES
            if k == 0:
                return verts.mean(axis=0, keepdims=True)
            return (idx / k) @ verts
        p = gll_points(k)
        return p[idx]

    # ------------------------------------------------------------------ basis
    def _monomials(self):
        k = self.degree
        if self.cell == "tetrahedron":
            return [e for e in itertools.product(range(k + 1), repeat=3) if sum(e) <= k]
        if self.cell == "triangle":
            return [e for e in itertools.product(range(k + 1), repeat=2) if sum(e) <= k]
        return list(itertools.product(range(k + 1), repeat=3))

    @property
    def _coeffs(self) -> np.ndarray:
        return _basis_coeffs(self.cell, self.degree)

    def tabulate(self, points: np.ndarray, nderiv: int = 1) -> np.ndarray:
        """Values and first derivatives at points.

        Returns (1 + tdim*nderiv, npts, nd): [phi, dphi/dX, dphi/dY(, dphi/dZ)].
        """
        pts = np.atleast_2d(np.asarray(points, dtype=np.float64))
        if self.cell == "hexahedron":
            return self._tabulate_tensor(pts, nderiv)
        mons = self._monomials()
        dim = len(mons[0])

        def mono(e):
            v = np.ones(pts.shape[0])
            for d in range(dim):
                v = v * pts[:, d] ** e[d]
            return v

        out = []
        V = np.empty((pts.shape[0], len(mons)))
        for j, e in enumerate(mons):
            V[:, j] = mono(e)
        out.append(V @ self._coeffs)
        if nderiv:
            for d in range(dim):
                D = np.zeros_like(V)
                for j, e in enumerate(mons):
                    if e[d] == 0:
                        continue
                    ee = list(e)
                    ee[d] -= 1
                    D[:, j] = e[d] * mono(ee)
                out.append(D @ self._coeffs)
        return np.stack(out)


    def _tabulate_tensor(self, pts: np.ndarray, nderiv: int) -> np.ndarray:
        """Q_k as a tensor product of 1D Lagrange polynomials on the GLL nodes."""
        k = self.degree
        C1 = np.linalg.inv(np.vander(gll_points(k), k + 1, increasing=True))   # (k+1, k+1)
        val, der = [], []
        for d in range(3):
            t = pts[:, d]
            V = np.vander(t, k + 1, increasing=True)
            D = np.zeros_like(V)
            for j in range(1, k + 1):
                D[:, j] = j * t ** (j - 1)
            val.append(V @ C1)
            der.append(D @ C1)
        idx = self.node_indices
        a, b, c = idx[:, 0], idx[:, 1], idx[:, 2]
        out = [val[0][:, a] * val[1][:, b] * val[2][:, c]]
        if nderiv:
            out.append(der[0][:, a] * val[1][:, b] * val[2][:, c])
            out.append(val[0][:, a] * der[1][:, b] * val[2][:, c])
            out.append(val[0][:, a] * val[1][:, b] * der[2][:, c])
        return np.stack(out)


@lru_cache(maxsize=None)
def _node_indices(cell: str, k: int) -> np.ndarray:
    if cell == "triangle":
        if k == 0:
            return np.array([[1, 1, 1]], dtype=np.int64)
        nodes = []
        for v in range(3):
            m = [0, 0, 0]
            m[v] = k
            nodes.append(m)
        for a, b in TRI_EDGES:
            for s_ in range(1, k):
                m = [0, 0, 0]
                m[a], m[b] = k - s_, s_
                nodes.append(m)
        for t in range(1, k):
            for s_ in range(1, k - t):
                nodes.append([k - s_ - t, s_, t])
        return np.array(nodes, dtype=np.int64)
    if cell == "hexahedron":
        return np.array([(a, b, c) for c in range(k + 1) for b in range(k + 1)
                         for a in range(k + 1)], dtype=np.int64)
    if k == 0:
        return np.array([[1, 1, 1, 1]], dtype=np.int64)   # centroid, scaled by 1/4 below
    nodes = []
    for v in range(4):
        m = [0, 0, 0, 0]
        m[v] = k
        nodes.append(m)
    for a, b in TET_EDGES:
        for s in range(1, k):
            m = [0, 0, 0, 0]
            m[a], m[b] = k - s, s
            nodes.append(m)
    for f in TET_FACES:
        # interior nodes of the face, lexicographic in the face's own (s, t)
        for t in range(1, k):
            for s in range(1, k - t):
                m = [0, 0, 0, 0]
                m[f[0]], m[f[1]], m[f[2]] = k - s - t, s, t
                nodes.append(m)
    for t3 in range(1, k):
        for t2 in range(1, k - t3):
            for t1 in range(1, k - t2 - t3):
                nodes.append([k - t1 - t2 - t3, t1, t2, t3])
    return np.array(nodes, dtype=np.int64)


@lru_cache(maxsize=None)
def _basis_coeffs(cell: str, k: int) -> np.ndarray:
    el = Element("P", cell, k)
    X = el.nodes
    mons = el._monomials()
    V = np.ones((X.shape[0], len(mons)))
    for j, e in enumerate(mons):
        for d in range(len(e)):
            V[:, j] *= X[:, d] ** e[d]
    return np.linalg.inv(V)


# ---------------------------------------------------------------- quadrature
@dataclass(frozen=True)
class Quadrature:
    points: np.ndarray   # (nq, dim)
    weights: np.ndarray  # (nq,)


def _gauss_jacobi01(m: int, alpha: float) -> tuple[np.ndarray, np.ndarray]:
    """m-point Gauss-Jacobi rule for weight (1-t)^alpha on [0, 1]."""
    x, w = roots_jacobi(m, alpha, 0.0)
    return (x + 1.0) / 2.0, w / 2.0 ** (alpha + 1.0)


def _gauss_legendre01(m: int) -> tuple[np.ndarray, np.ndarray]:
    x, w = roots_legendre(m)
    return (x + 1.0) / 2.0, w / 2.0


@lru_cache(maxsize=None)
def tet_collapsed(m: int) -> Quadrature:
    """Stroud conical product: exact for degree 2m - 1 on the unit tetrahedron."""
    a, wa = _gauss_jacobi01(m, 2.0)
    b, wb = _gauss_jacobi01(m, 1.0)
    c, wc = _gauss_legendre01(m)
    pts, wts = [], []
    for i in range(m):
        for j in range(m):
            for l in range(m):
                x = a[i]
                y = b[j] * (1.0 - a[i])
                z = c[l] * (1.0 - a[i]) * (1.0 - b[j])
                pts.append((x, y, z))
                wts.append(wa[i] * wb[j] * wc[l])
    return Quadrature(np.array(pts), np.array(wts))


@lru_cache(maxsize=None)
def tri_collapsed(m: int) -> Quadrature:
    a, wa = _gauss_jacobi01(m, 1.0)
    b, wb = _gauss_legendre01(m)
    pts, wts = [], []
    for i in range(m):
        for j in range(m):
            pts.append((a[i], b[j] * (1.0 - a[i])))
            wts.append(wa[i] * wb[j])
    return Quadrature(np.array(pts), np.array(wts))


@lru_cache(maxsize=None)
def tet_rule(degree: int) -> Quadrature:
    if degree <= 1:
        return Quadrature(np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0]))
    if degree == 2:
        a = (5.0 - math.sqrt(5.0)) / 20.0
        b = (5.0 + 3.0 * math.sqrt(5.0)) / 20.0
        pts = np.array([[a, a, a], [b, a, a], [a, b, a], [a, a, b]])
        return Quadrature(pts, np.full(4, 1.0 / 24.0))
    return tet_collapsed((degree + 2) // 2)


@lru_cache(maxsize=None)
def tri_rule(degree: int) -> Quadrature:
    if degree <= 1:
        return Quadrature(np.array([[1.0 / 3.0, 1.0 / 3.0]]), np.array([0.5]))
    if degree == 2:
        pts = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
        return Quadrature(pts, np.full(3, 1.0 / 6.0))
    return tri_collapsed((degree + 2) // 2)


@lru_cache(maxsize=None)
def hex_gauss(m: int) -> Quadrature:
    """m^3-point Gauss-Legendre rule, points ordered x fastest."""
    x, w = _gauss_legendre01(m)
    pts = np.array([(x[a], x[b], x[c]) for c in range(m) for b in range(m) for a in range(m)])
    wts = np.array([w[a] * w[b] * w[c] for c in range(m) for b in range(m) for a in range(m)])
    return Quadrature(pts, wts)


@lru_cache(maxsize=None)
def quad_gauss(m: int) -> Quadrature:
    x, w = _gauss_legendre01(m)
    pts = np.array([(x[a], x[b]) for b in range(m) for a in range(m)])
    wts = np.array([w[a] * w[b] for b in range(m) for a in range(m)])
    return Quadrature(pts, wts)


# Facet quadrature permutations of a triangle facet: FACET_PERMS[q][k] is the
# rank, among the facet's three global vertex numbers, of the facet's k-th
# vertex as the cell lists it (TET_FACES order).  q = 0 is the identity
# (vertices stored in ascending global order).
FACET_PERMS = list(itertools.permutations(range(3)))


def tet_facet_points(lf: int, pts2d: np.ndarray, perm: int = 0) -> np.ndarray:
    """Map facet reference points (s, t) onto facet lf of the reference tetrahedron.

    (s, t) are coordinates on the facet's vertices in ascending global order
    (barycentric weights (1 - s - t, s, t)); `perm` (FACET_PERMS) says where
    the cell's own facet vertices fall in that order, so both cells of an
    interior facet evaluate at the same physical points."""
    f = TET_FACES[lf]
    lam = np.stack([1.0 - pts2d[:, 0] - pts2d[:, 1], pts2d[:, 0], pts2d[:, 1]], axis=1)
    r = FACET_PERMS[perm]
    return sum(lam[:, r[k]:r[k] + 1] * TET_VERTICES[f[k]][None] for k in range(3))


@lru_cache(maxsize=None)
def line_gauss(m: int) -> Quadrature:
    """m-point Gauss-Legendre rule on [0, 1] (triangle facets), points (m, 1)."""
    x, w = _gauss_legendre01(m)
    return Quadrature(x[:, None], w)


def tri_facet_points(lf: int, pts1d: np.ndarray, perm: int = 0) -> np.ndarray:
    """Map edge reference points s onto facet (edge) lf of the reference triangle.

    s runs from the edge's lower to its higher global vertex; perm 1 means
    the cell lists the edge's vertices the other way round (perm 0: in
    ascending global order), so both cells of an interior edge evaluate at
    the same physical points."""
    a, b = TRI_EDGES[lf]
    s = np.asarray(pts1d, dtype=np.float64).reshape(-1)
    if perm:
        s = 1.0 - s
    return (1.0 - s)[:, None] * TRI_VERTICES[a][None] + s[:, None] * TRI_VERTICES[b][None]


def tri_facet_tables_permuted(el, pts1d: np.ndarray):
    """Basis values (3, 2, nq, nd) and reference gradients (3, 2, nq, nd, 2) at
    the edge points of every local edge under both orientations (the tables
    of interior-facet bodies on triangles, indexed [lf][perm][iq])."""
    nq = len(np.asarray(pts1d).reshape(-1))
    phi = np.empty((3, 2, nq, el.ndofs))
    dphi = np.empty(phi.shape + (2,))
    for lf in range(3):
        for q in range(2):
            t = el.tabulate(tri_facet_points(lf, pts1d, q))
            phi[lf, q] = t[0]
            dphi[lf, q] = np.stack([t[1], t[2]], axis=-1)
    return phi, dphi


def facet_tables_permuted(el, pts2d: np.ndarray):
    """Basis values (4, 6, nq, nd) and reference gradients (4, 6, nq, nd, 3) at
    the facet points of every local facet under every facet permutation (the
    tables of interior-facet bodies, indexed [lf][perm][iq])."""
    phi = np.empty((4, len(FACET_PERMS), len(pts2d), el.ndofs))
    dphi = np.empty(phi.shape + (3,))
    for lf in range(4):
        for q in range(len(FACET_PERMS)):
            t = el.tabulate(tet_facet_points(lf, pts2d, q))
            phi[lf, q] = t[0]
            dphi[lf, q] = np.stack([t[1], t[2], t[3]], axis=-1)
    return phi, dphi
