"""Many independent meshes assembled as one block-diagonal system.

Instances that share an element (same cell type, family, degree, block
size) are merged into one mesh and one function space: vertex, cell and dof
numbers of instance k are offset by the totals of instances 0..k-1, so the
sparsity of the merged form is block diagonal and its CSR is exactly the
concatenation of the per-instance CSR matrices (rows and columns offset).
One assembly launch then covers every instance of the group, and each
instance's matrix is the row block `row_ptr[r0:r1+1] - row_ptr[r0]` with
columns `cols - c0` (the "block-separate CSR" view, `block_csr`).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import forms
from .fem import Constant, Function, FunctionSpace, dirichletbc, functionspace, locate_dofs_topological
from .mesh import Mesh, cell_random_lognormal, create_unit_cube


def merge_meshes(meshes: List[Mesh]) -> Mesh:
    if len({m.cell_type for m in meshes}) != 1:
        raise ValueError("instances of one group must share the cell type")
    voff = np.cumsum([0] + [m.num_vertices for m in meshes])
    cells = np.concatenate([m.cells.astype(np.int64) + voff[k] for k, m in enumerate(meshes)])
    return Mesh(x=np.ascontiguousarray(np.concatenate([m.x for m in meshes])),
                cells=np.ascontiguousarray(cells.astype(np.int32)),
                cell_type=meshes[0].cell_type, n=max(m.n for m in meshes),
                lattice=np.concatenate([m.lattice for m in meshes]),
                cell_cube=np.concatenate([m.cell_cube for m in meshes]),
                cell_sub=np.concatenate([m.cell_sub for m in meshes]))


def merge_spaces(spaces: List[FunctionSpace], mesh: Mesh) -> FunctionSpace:
    el, bs = spaces[0].element, spaces[0].bs
    if any(V.element != el or V.bs != bs for V in spaces):
        raise ValueError("instances of one group must share the element")
    V = object.__new__(FunctionSpace)
    V.mesh, V.element, V.shape, V.bs = mesh, el, spaces[0].shape, bs
    noff = np.cumsum([0] + [S.num_nodes for S in spaces])
    V._dofmap = np.ascontiguousarray(
        np.concatenate([S.dofmap.astype(np.int64) + noff[k] for k, S in enumerate(spaces)]).astype(np.int32))
    V._num_nodes = int(noff[-1])
    return V


@dataclass(eq=False)
class Group:
    """One element group: merged mesh/space/forms plus per-instance offsets."""
    degree: int
    instances: List[int]                 # indices into the batch
    mesh: Mesh
    V: FunctionSpace
    a: forms.Form
    L: forms.Form
    bcs: list
    cell_offsets: np.ndarray             # (k + 1,)
    dof_offsets: np.ndarray              # (k + 1,) scalar dofs
    sizes: List[int] = field(default_factory=list)

    def block_csr(self, row_ptr, cols, values, j: int):
        """CSR (row_ptr, cols, values) of instance j of the group (local numbering)."""
        r0, r1 = self.dof_offsets[j], self.dof_offsets[j + 1]
        p0, p1 = row_ptr[r0], row_ptr[r1]
        return row_ptr[r0:r1 + 1] - p0, cols[p0:p1] - r0, values[p0:p1]


@dataclass(eq=False)
class Batch:
    sizes: np.ndarray                    # cubes per direction of each instance
    degrees: np.ndarray
    groups: List[Group]

    @property
    def num_instances(self) -> int:
        return int(self.sizes.size)


def poisson_instance(n: int, degree: int, kappa_seed: int = 4):
    """-div(kappa grad u) = 1, u = 0 on the boundary, P_degree on n^3 x 6 tets."""
    mesh = create_unit_cube(int(n))
    V = functionspace(mesh, ("P", int(degree)))
    return mesh, V, cell_random_lognormal(mesh, kappa_seed, 0, 0.0, 1.0)


def many_poisson(count: int = 512, seed: int = 3, nmin: int = 4, nmax: int = 24,
                 degrees=(1, 2, 3), kappa_seed: int = 4) -> Batch:
    """`count` independent Poisson problems, n ~ U{nmin..nmax}, degree ~ U(degrees).

    The draws use numpy's PCG64 generator seeded with `seed`; kappa ~
    lognormal(0, 1) per cell keyed on the cell's lattice position.
    """
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    degs = rng.choice(np.asarray(degrees), size=count)
    groups = []
    for k in sorted(set(int(d) for d in degs)):
        idx = [i for i in range(count) if degs[i] == k]
        parts = [poisson_instance(sizes[i], k, kappa_seed) for i in idx]
        mesh = merge_meshes([p[0] for p in parts])
        V = merge_spaces([p[1] for p in parts], mesh)
        D = functionspace(mesh, ("DG", 0))
        kap = Function(D, "kappa")
        kap.set_cellwise(np.concatenate([p[2] for p in parts]))
        a = forms.stiffness(V, kap)
        L = forms.source(V, Constant([1.0]))
        bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
        coff = np.cumsum([0] + [p[0].num_cells for p in parts])
        doff = np.cumsum([0] + [p[1].size for p in parts])
        groups.append(Group(k, idx, mesh, V, a, L, [bc], coff, doff,
                            [int(sizes[i]) for i in idx]))
    return Batch(sizes, degs, groups)
