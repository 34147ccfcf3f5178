"""Structured unit-cube meshes (the mesh side of the DOLFINx stand-in).

`create_unit_cube` mirrors `dolfinx.mesh.create_unit_cube` /
`create_box` for the two cell types the assembly path is measured on:

* tetrahedra: each lattice cube is split into the 6 Kuhn tetrahedra that
  share the cube diagonal v0-v7 (the split DOLFINx's `create_box` uses);
* hexahedra: one cell per cube, vertices in lexicographic (x fastest)
  order, the Basix reference-hexahedron vertex order.

Cells are ordered cube-by-cube along a Morton (Z-order) curve and vertices
by the Morton code of their lattice index, so neighbouring cells and their
vertices sit close together in memory.  Tetrahedron vertices are stored in
ascending global index order (the "ordered simplex" convention), which makes
every shared edge/face orientation agree between the cells that share it.

The integer lattice coordinates of every vertex are kept (`Mesh.lattice`):
the function-space layer keys dofs on them, so dof identity never depends
on floating-point coordinates (jittered meshes included).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from functools import cached_property

import numpy as np

from . import rng

# Kuhn split of a cube into 6 tetrahedra sharing the diagonal v0-v7; cube-local
# vertex id is a + 2 b + 4 c for the corner (a, b, c).
KUHN_TETS = np.array([(0, 1, 3, 7), (0, 1, 7, 5), (0, 5, 7, 4),
                      (0, 3, 2, 7), (0, 6, 4, 7), (0, 2, 6, 7)], dtype=np.int64)
CUBE_CORNERS = np.array([(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)],
                        dtype=np.int64)

# Facet -> local vertices.  Tetrahedron facet i is opposite vertex i.
TET_FACETS = np.array([(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)], dtype=np.int64)
# Triangle facet (edge) i is opposite vertex i.
TRI_FACETS = np.array([(1, 2), (0, 2), (0, 1)], dtype=np.int64)
# Basix reference hexahedron facets: z=0, y=0, x=0, x=1, y=1, z=1.
HEX_FACETS = np.array([(0, 1, 2, 3), (0, 1, 4, 5), (0, 2, 4, 6),
                       (1, 3, 5, 7), (2, 3, 6, 7), (4, 5, 6, 7)], dtype=np.int64)


def morton3(i, j, k, bits: int = 12) -> np.ndarray:
    """Interleave the low `bits` bits of three non-negative integer arrays."""
    i = np.asarray(i, dtype=np.uint64)
    j = np.asarray(j, dtype=np.uint64)
    k = np.asarray(k, dtype=np.uint64)
    code = np.zeros(np.broadcast(i, j, k).shape, dtype=np.uint64)
    one = np.uint64(1)
    for b in range(bits):
        bb = np.uint64(b)
        code |= ((i >> bb) & one) << np.uint64(3 * b)
        code |= ((j >> bb) & one) << np.uint64(3 * b + 1)
        code |= ((k >> bb) & one) << np.uint64(3 * b + 2)
    return code


@dataclass(eq=False)
class Mesh:
    """Unit-cube mesh: geometry, topology and lattice bookkeeping.

    Attributes
    ----------
    x : (Nv, 3) float64 vertex coordinates (always 3 components).
    cells : (Nc, nv) int32 cell -> vertex map (the geometry dofmap).
    cell_type : "tetrahedron" or "hexahedron".
    n : cubes per direction; h = 1/n.
    lattice : (Nv, 3) int64 integer lattice index of every vertex.
    cell_cube : (Nc, 3) int64 lattice cube that contains each cell.
    cell_sub : (Nc,) int64 Kuhn sub-tetrahedron id (0 for hexahedra).
    """

    x: np.ndarray
    cells: np.ndarray
    cell_type: str
    n: int
    lattice: np.ndarray
    cell_cube: np.ndarray
    cell_sub: np.ndarray
    jitter: float = 0.0
    seed: int = 0
    _cache: dict = field(default_factory=dict, repr=False)

    @property
    def h(self) -> float:
        return 1.0 / self.n

    @property
    def num_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def num_vertices(self) -> int:
        return int(self.x.shape[0])

    @property
    def nv_per_cell(self) -> int:
        return int(self.cells.shape[1])

    @property
    def facet_table(self) -> np.ndarray:
        return {"tetrahedron": TET_FACETS, "triangle": TRI_FACETS}.get(self.cell_type, HEX_FACETS)

    @property
    def tdim(self) -> int:
        return 2 if self.cell_type == "triangle" else 3

    # ------------------------------------------------------------------ facets
    @cached_property
    def _facets(self):
        """Unique facets: (keys, cell_facet -> facet id, facet -> cells)."""
        ft = self.facet_table
        nc, nf = self.num_cells, ft.shape[0]
        fv = self.cells[:, ft].astype(np.int64)          # (Nc, nf, nvf)
        fv = np.sort(fv, axis=2).reshape(nc * nf, -1)
        _, first, inv = np.unique(fv, axis=0, return_index=True, return_inverse=True)
        inv = inv.reshape(-1)
        counts = np.bincount(inv)
        return fv[first], inv.reshape(nc, nf), counts

    @cached_property
    def exterior_facets(self) -> np.ndarray:
        """(n, 2) int32 [cell, local facet] pairs of boundary facets, cell-major."""
        _, cf, counts = self._facets
        mask = counts[cf] == 1
        c, lf = np.nonzero(mask)
        return np.stack([c, lf], axis=1).astype(np.int32)

    @cached_property
    def interior_facets(self) -> np.ndarray:
        """(n, 4) int32 [c0, lf0, c1, lf1] with c0 < c1, ordered by facet id."""
        _, cf, counts = self._facets
        flat = cf.reshape(-1)
        order = np.argsort(flat, kind="stable")
        fid = flat[order]
        cells = order // cf.shape[1]
        lfs = order % cf.shape[1]
        inner = counts[fid] == 2
        fid, cells, lfs = fid[inner], cells[inner], lfs[inner]
        c0, c1 = cells[0::2], cells[1::2]
        l0, l1 = lfs[0::2], lfs[1::2]
        return np.stack([c0, l0, c1, l1], axis=1).astype(np.int32)

    def facet_midpoints(self, facets: np.ndarray) -> np.ndarray:
        ft = self.facet_table
        v = self.cells[facets[:, 0][:, None], ft[facets[:, 1]]]
        return self.x[v].mean(axis=1)

    def locate_boundary_facets(self, marker) -> np.ndarray:
        """Exterior [cell, local facet] pairs whose vertices all satisfy `marker`.

        Same rule as `dolfinx.mesh.locate_entities_boundary`: the marker is
        evaluated at the facet vertices, x of shape (3, npoints).
        """
        ext = self.exterior_facets
        ft = self.facet_table
        v = self.cells[ext[:, 0][:, None], ft[ext[:, 1]]]       # (n, nvf)
        ok = marker(self.x[v.reshape(-1)].T).reshape(v.shape)
        return ext[np.all(ok, axis=1)]

    def facet_permutations(self) -> np.ndarray:
        """(Nc, nf) uint8 quadrature permutation of every cell facet.

        Tetrahedra: index into element.FACET_PERMS of the ranks, in ascending
        global order, of the facet's vertices as the cell lists them; 0 when a
        cell stores its vertices in ascending order (the default meshes).
        Generated interior-facet bodies evaluate their facet points through it,
        so both cells of a facet use the same physical points.  Hexahedra keep
        the lexicographic order of the structured lattice: all zeros.
        """
        from .element import FACET_PERMS
        nf = self.facet_table.shape[0]
        if self.cell_type == "triangle":
            # an edge: 1 when the cell lists its two vertices in descending order
            fv = self.cells[:, self.facet_table]                  # (Nc, 3, 2)
            return (fv[..., 0] > fv[..., 1]).astype(np.uint8)
        if self.cell_type != "tetrahedron":
            return np.zeros((self.num_cells, nf), dtype=np.uint8)
        fv = self.cells[:, self.facet_table]                      # (Nc, 4, 3)
        ranks = np.argsort(np.argsort(fv, axis=2), axis=2)       # rank of each listed vertex
        code = ranks[..., 0] * 9 + ranks[..., 1] * 3 + ranks[..., 2]
        lut = np.zeros(27, dtype=np.uint8)
        for q, r in enumerate(FACET_PERMS):
            lut[r[0] * 9 + r[1] * 3 + r[2]] = q
        return lut[code]


def _cube_order(n: int) -> np.ndarray:
    """(n^3, 3) cube lattice indices in Morton order."""
    g = np.arange(n, dtype=np.int64)
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    cubes = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    order = np.argsort(morton3(cubes[:, 0], cubes[:, 1], cubes[:, 2]), kind="stable")
    return cubes[order]


def create_unit_cube(n: int, cell_type: str = "tetrahedron", jitter: float = 0.0,
                     seed: int = 0, shuffle_vertices: int | None = None) -> Mesh:
    """Unit cube split into n^3 cubes (6 Kuhn tetrahedra or 1 hexahedron each).

    jitter > 0 moves every *interior* vertex (i, j, k) by
    h * jitter * (2 U - 1) per component, U = uniform(seed, i, j, k, component):
    a hash of the lattice index, so the perturbation does not depend on how
    the vertices are numbered.  Boundary vertices stay on the cube surface,
    so the mesh volume stays exactly 1.

    shuffle_vertices (tetrahedra): a seed; each cell then lists its four
    vertices in a random order instead of ascending global order, so the two
    cells of an interior facet see its vertices in different orders (non-zero
    facet quadrature permutations).
    """
    if cell_type not in ("tetrahedron", "hexahedron"):
        raise ValueError(f"unsupported cell type {cell_type!r}")
    n = int(n)
    if n < 1:
        raise ValueError("n must be >= 1")
    m = n + 1
    g = np.arange(m, dtype=np.int64)
    vx, vy, vz = np.meshgrid(g, g, g, indexing="ij")
    lat = np.stack([vx.ravel(), vy.ravel(), vz.ravel()], axis=1)
    vorder = np.argsort(morton3(lat[:, 0], lat[:, 1], lat[:, 2]), kind="stable")
    lat = lat[vorder]
    # lexicographic lattice id -> Morton vertex number
    lex_to_vertex = np.empty(m ** 3, dtype=np.int64)
    lex_to_vertex[lat[:, 0] + m * (lat[:, 1] + m * lat[:, 2])] = np.arange(m ** 3)

    x = lat.astype(np.float64) / n
    if jitter:
        interior = np.all((lat > 0) & (lat < n), axis=1)
        li = lat[interior]
        for comp in range(3):
            u = rng.uniform(seed, li[:, 0], li[:, 1], li[:, 2], comp)
            x[interior, comp] += (2.0 * u - 1.0) * jitter / n

    cubes = _cube_order(n)
    corner = cubes[:, None, :] + CUBE_CORNERS[None, :, :]              # (nc, 8, 3)
    cv = lex_to_vertex[corner[..., 0] + m * (corner[..., 1] + m * corner[..., 2])]
    if cell_type == "tetrahedron":
        cells = cv[:, KUHN_TETS].reshape(-1, 4)
        cells = np.sort(cells, axis=1)
        if shuffle_vertices is not None:
            keys = np.random.default_rng(shuffle_vertices).random(cells.shape)
            cells = np.take_along_axis(cells, np.argsort(keys, axis=1), axis=1)
        cell_cube = np.repeat(cubes, 6, axis=0)
        cell_sub = np.tile(np.arange(6, dtype=np.int64), len(cubes))
    else:
        cells = cv
        cell_cube = cubes.copy()
        cell_sub = np.zeros(len(cubes), dtype=np.int64)
    return Mesh(x=np.ascontiguousarray(x), cells=np.ascontiguousarray(cells.astype(np.int32)),
                cell_type=cell_type, n=n, lattice=lat, cell_cube=cell_cube,
                cell_sub=cell_sub, jitter=float(jitter), seed=int(seed))


def create_unit_square(nx: int, ny: int, cell_type: str = "triangle",
                       shuffle_vertices: int | None = None) -> Mesh:
    """`dolfinx.mesh.create_unit_square` look-alike: nx x ny squares, each split
    into two triangles along the diagonal from (i, j) to (i + 1, j + 1)
    (DOLFINx's default "right" diagonal).  Squares in Morton order, vertices
    by the Morton code of their lattice index (z = 0), cell vertices in
    ascending global order (or shuffled with the seed shuffle_vertices, so
    interior edges get non-zero facet permutations)."""
    if cell_type != "triangle":
        raise ValueError(f"unsupported cell type {cell_type!r}")
    nx, ny = int(nx), int(ny)
    if nx < 1 or ny < 1:
        raise ValueError("nx, ny must be >= 1")
    gx, gy = np.arange(nx + 1, dtype=np.int64), np.arange(ny + 1, dtype=np.int64)
    vx, vy = np.meshgrid(gx, gy, indexing="ij")
    lat = np.stack([vx.ravel(), vy.ravel(), np.zeros(vx.size, dtype=np.int64)], axis=1)
    lat = lat[np.argsort(morton3(lat[:, 0], lat[:, 1], lat[:, 2]), kind="stable")]
    lex_to_vertex = np.empty((nx + 1) * (ny + 1), dtype=np.int64)
    lex_to_vertex[lat[:, 0] + (nx + 1) * lat[:, 1]] = np.arange(lat.shape[0])
    x = np.stack([lat[:, 0] / nx, lat[:, 1] / ny, np.zeros(lat.shape[0])], axis=1)
    sq = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij"), axis=-1).reshape(-1, 2)
    sq = sq[np.argsort(morton3(sq[:, 0], sq[:, 1], np.zeros(len(sq), dtype=np.int64)), kind="stable")]
    def vid(i, j):
        return lex_to_vertex[i + (nx + 1) * j]
    v0, v1 = vid(sq[:, 0], sq[:, 1]), vid(sq[:, 0] + 1, sq[:, 1])
    v2, v3 = vid(sq[:, 0], sq[:, 1] + 1), vid(sq[:, 0] + 1, sq[:, 1] + 1)
    cells = np.stack([np.stack([v0, v1, v3], axis=1), np.stack([v0, v2, v3], axis=1)],
                     axis=1).reshape(-1, 3)
    cells = np.sort(cells, axis=1)
    if shuffle_vertices is not None:
        keys = np.random.default_rng(shuffle_vertices).random(cells.shape)
        cells = np.take_along_axis(cells, np.argsort(keys, axis=1), axis=1)
    cube = np.concatenate([np.repeat(sq, 2, axis=0), np.zeros((2 * len(sq), 1), dtype=np.int64)], axis=1)
    return Mesh(x=np.ascontiguousarray(x), cells=np.ascontiguousarray(cells.astype(np.int32)),
                cell_type="triangle", n=max(nx, ny), lattice=lat, cell_cube=cube,
                cell_sub=np.tile(np.arange(2, dtype=np.int64), len(sq)))


def cell_random_uniform(mesh: Mesh, seed: int, field_id: int, lo: float, hi: float) -> np.ndarray:
    """Per-cell U[lo, hi] keyed on (seed, cube lattice index, sub-cell, field)."""
    cc = mesh.cell_cube
    u = rng.uniform(seed, cc[:, 0], cc[:, 1], cc[:, 2], mesh.cell_sub, field_id)
    return lo + (hi - lo) * u


def cell_random_lognormal(mesh: Mesh, seed: int, field_id: int, mu: float = 0.0,
                          sigma: float = 1.0) -> np.ndarray:
    """Per-cell lognormal(mu, sigma) keyed like `cell_random_uniform`."""
    cc = mesh.cell_cube
    z = rng.normal(seed, cc[:, 0], cc[:, 1], cc[:, 2], mesh.cell_sub, field_id)
    return np.exp(mu + sigma * z)
