"""Function spaces, functions, constants and Dirichlet conditions.

This is the DOLFINx-side stand-in the assembly path consumes: it produces
the arrays `dolfinx.fem.FunctionSpace`, `Function`, `Constant` and
`DirichletBC` hand to cuDOLFINx (cell -> dof map, block size, coefficient
vectors, BC markers / indices / dense values).

Dof numbering: every node gets an integer key on the degree-k refined vertex
lattice; unique keys are numbered in Morton order of that lattice, which
keeps the dofs of neighbouring cells close (the role DOLFINx's dofmap
reordering plays).  DG spaces number dofs cell by cell.  A vector space of
block size bs unrolls node n to scalar dofs bs*n + j (DOLFINx convention).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from functools import cached_property
from typing import Callable, Optional, Sequence

import numpy as np

from .element import Element
from .mesh import Mesh, morton3


class FunctionSpace:
    def __init__(self, mesh: Mesh, element: Element, shape: Optional[tuple] = None):
        if element.cell != mesh.cell_type:
            raise ValueError("element cell type does not match the mesh")
        self.mesh = mesh
        self.element = element
        self.shape = tuple(shape) if shape else ()
        self.bs = int(np.prod(self.shape)) if self.shape else 1
        self._build_dofmap()

    # DOLFINx-like accessors
    @property
    def dofmap(self) -> np.ndarray:
        """(Nc, nd) int32 blocked cell -> node map."""
        return self._dofmap

    @property
    def num_nodes(self) -> int:
        return self._num_nodes

    @property
    def size(self) -> int:
        """Number of scalar dofs (nodes * block size)."""
        return self._num_nodes * self.bs

    def _build_dofmap(self) -> None:
        mesh, el = self.mesh, self.element
        nc, nd = mesh.num_cells, el.ndofs
        if el.discontinuous:
            self._dofmap = np.arange(nc * nd, dtype=np.int32).reshape(nc, nd)
            self._num_nodes = nc * nd
            return
        k = el.degree
        idx = el.node_indices
        lat = mesh.lattice[mesh.cells]                          # (Nc, nv, 3)
        if mesh.cell_type in ("tetrahedron", "triangle"):
            keys = np.einsum("dv,cvx->cdx", idx, lat)           # (Nc, nd, 3), refined by k
            code = morton3(keys[..., 0], keys[..., 1], keys[..., 2])
        else:
            # hexahedra: nodes grouped by owning cube (cube Morton order), each
            # cube owning its k x k x k block of the refined lattice (the last
            # layer of cubes also owns the upper boundary nodes)
            keys = k * lat[:, 0, None, :] + idx[None, :, :]
            cube = np.minimum(keys // k, mesh.n - 1)
            loc = keys - k * cube
            k1 = np.uint64(k + 1)
            code = (morton3(cube[..., 0], cube[..., 1], cube[..., 2]) * k1 ** np.uint64(3)
                    + (loc[..., 2].astype(np.uint64) * k1 + loc[..., 1].astype(np.uint64)) * k1
                    + loc[..., 0].astype(np.uint64))
        uniq, inv = np.unique(code.reshape(-1), return_inverse=True)
        self._dofmap = inv.reshape(nc, nd).astype(np.int32)
        self._num_nodes = int(uniq.shape[0])

    @cached_property
    def node_coordinates(self) -> np.ndarray:
        """(num_nodes, 3) physical coordinates of the nodes."""
        mesh, el = self.mesh, self.element
        X = el.nodes
        xc = mesh.x[mesh.cells]                                  # (Nc, nv, 3)
        if mesh.cell_type in ("tetrahedron", "triangle"):
            lam = np.concatenate([1.0 - X.sum(axis=1, keepdims=True), X], axis=1)  # (nd, nv)
            xp = np.einsum("dv,cvx->cdx", lam, xc)
        else:
            a, b, c = X[:, 0], X[:, 1], X[:, 2]
            w = []
            for cc in (0, 1):
                for bb in (0, 1):
                    for aa in (0, 1):
                        w.append((a if aa else 1 - a) * (b if bb else 1 - b) * (c if cc else 1 - c))
            w = np.stack(w, axis=1)                              # (nd, 8)
            xp = np.einsum("dv,cvx->cdx", w, xc)
        out = np.empty((self._num_nodes, 3))
        out[self._dofmap.reshape(-1)] = xp.reshape(-1, 3)
        return out

    def tabulate_dof_coordinates(self) -> np.ndarray:
        return self.node_coordinates

    def unrolled_dofmap(self) -> np.ndarray:
        """(Nc, nd*bs) scalar dofs bs*node + j (cuDOLFINx `CUDADofMap` layout)."""
        d = self._dofmap.astype(np.int64)[:, :, None] * self.bs + np.arange(self.bs)[None, None, :]
        return d.reshape(d.shape[0], -1).astype(np.int32)

    def nodes_on_facets(self, facets: np.ndarray) -> np.ndarray:
        """Sorted unique nodes that lie on the given [cell, local facet] pairs."""
        el = self.element
        idx = el.node_indices
        k = el.degree
        sel = []
        for lf in range(self.mesh.facet_table.shape[0]):
            cells = facets[facets[:, 1] == lf, 0]
            if cells.size == 0:
                continue
            if self.mesh.cell_type in ("tetrahedron", "triangle"):
                on = idx[:, lf] == 0 if k > 0 else np.zeros(len(idx), bool)
            else:
                axis, val = {0: (2, 0), 1: (1, 0), 2: (0, 0), 3: (0, k), 4: (1, k), 5: (2, k)}[lf]
                on = idx[:, axis] == val
            sel.append(self._dofmap[cells][:, on].reshape(-1))
        if not sel:
            return np.zeros(0, dtype=np.int32)
        return np.unique(np.concatenate(sel)).astype(np.int32)


def functionspace(mesh: Mesh, element, shape: Optional[Sequence[int]] = None) -> FunctionSpace:
    """`dolfinx.fem.functionspace` look-alike: element = ("P"|"DG", degree)."""
    if isinstance(element, Element):
        el = element
    else:
        family, degree = element[0], int(element[1])
        family = {"Lagrange": "P", "CG": "P", "Q": "P", "Discontinuous Lagrange": "DG"}.get(family, family)
        el = Element(family, mesh.cell_type, degree)
    return FunctionSpace(mesh, el, shape)


_WRITING_FUNCS = {np.copyto, np.put, np.place, np.putmask, np.fill_diagonal}


class _TrackedArray(np.ndarray):
    """The values of a Function: a float64 array that bumps its Function's
    `version` whenever it, or a view of it, is written through numpy
    (item assignment, in-place arithmetic, ufuncs with out=, fill, copyto,
    put, place, putmask).  The device copy of a coefficient is refreshed
    when the version changed, so host writes such as `f.x[:] = v` reach the
    next assembly.  Writes through raw pointers (ctypes, buffers) are not
    seen: bump `version` by hand after those."""

    _owner = None

    def __array_finalize__(self, obj):
        self._owner = getattr(obj, "_owner", None)

    def _touch(self):
        if self._owner is not None:
            self._owner.version += 1

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        self._touch()

    def fill(self, value):
        super().fill(value)
        self._touch()

    def __array_ufunc__(self, ufunc, method, *inputs, out=None, **kwargs):
        args = [x.view(np.ndarray) if isinstance(x, _TrackedArray) else x for x in inputs]
        touched = []
        if out is not None:
            touched = [o for o in out if isinstance(o, _TrackedArray)]
            kwargs["out"] = tuple(o.view(np.ndarray) if isinstance(o, _TrackedArray) else o
                                  for o in out)
        res = getattr(ufunc, method)(*args, **kwargs)
        for o in touched:
            o._touch()
        if out is not None and len(out) == 1:
            return out[0]
        return res

    def __array_function__(self, func, types, args, kwargs):
        res = super().__array_function__(func, types, args, kwargs)
        if func in _WRITING_FUNCS and args and isinstance(args[0], _TrackedArray):
            args[0]._touch()
        return res


class Function:
    """Coefficient vector on a function space (scalar dofs, bs-unrolled)."""

    def __init__(self, V: FunctionSpace, name: str = "f"):
        self.function_space = V
        self.name = name
        self.version = 0            # bumped on every host-side change of x
        x = np.zeros(V.size, dtype=np.float64).view(_TrackedArray)
        x._owner = self
        self._x = x

    @property
    def x(self) -> np.ndarray:
        return self._x

    @x.setter
    def x(self, values) -> None:
        self._x[:] = values

    @property
    def array(self) -> np.ndarray:
        return self.x

    def interpolate(self, f: Callable) -> None:
        """Nodal interpolation: f(x) with x of shape (3, npts) -> (npts,) or (bs, npts)."""
        V = self.function_space
        vals = np.asarray(f(V.node_coordinates.T), dtype=np.float64)
        if V.bs == 1:
            self.x[:] = np.broadcast_to(vals.reshape(-1), (V.num_nodes,))
        else:
            vals = np.broadcast_to(vals.reshape(V.bs, -1), (V.bs, V.num_nodes))
            self.x[:] = vals.T.reshape(-1)
        self.version += 1

    def set_cellwise(self, values: np.ndarray) -> None:
        """Set a DG0 function from one value per cell."""
        V = self.function_space
        if V.element.degree != 0 or not V.element.discontinuous:
            raise ValueError("set_cellwise needs a DG0 space")
        self.x[V.dofmap[:, 0]] = values
        self.version += 1


class Constant:
    def __init__(self, value):
        self.value = np.atleast_1d(np.asarray(value, dtype=np.float64)).copy()


@dataclass(eq=False)
class DirichletBC:
    """Dirichlet condition u = g on a set of (scalar, bs-unrolled) dofs."""
    function_space: FunctionSpace
    dofs: np.ndarray            # sorted scalar dofs
    values: np.ndarray          # dense (V.size,) array, g at constrained dofs

    def dof_indices(self):
        """(dofs, number owned) as DOLFINx `DirichletBC::dof_indices`; one rank owns all."""
        return self.dofs, len(self.dofs)

    def mark_dofs(self, markers: np.ndarray) -> None:
        markers[self.dofs] = 1


def locate_dofs_geometrical(V: FunctionSpace, marker: Callable) -> np.ndarray:
    """Nodes whose coordinates satisfy marker(x), x of shape (3, n)."""
    ok = marker(V.node_coordinates.T)
    return np.nonzero(ok)[0].astype(np.int32)


def locate_dofs_topological(V: FunctionSpace, entity_dim: int, entities: np.ndarray) -> np.ndarray:
    """Nodes on the given boundary facets ([cell, local facet] pairs)."""
    if entity_dim != 2:
        raise NotImplementedError("only facets (entity_dim=2) are supported")
    return V.nodes_on_facets(np.asarray(entities))


def dirichletbc(value, dofs: np.ndarray, V: FunctionSpace) -> DirichletBC:
    """u = value on the given *blocked* nodes (all components for vector spaces).

    value: scalar, sequence of bs values, or a Function on V.
    """
    nodes = np.asarray(dofs, dtype=np.int64)
    bs = V.bs
    sdofs = (nodes[:, None] * bs + np.arange(bs)[None, :]).reshape(-1)
    sdofs = np.sort(sdofs).astype(np.int32)
    g = np.zeros(V.size, dtype=np.float64)
    if isinstance(value, Function):
        g[sdofs] = value.x[sdofs]
    else:
        v = np.broadcast_to(np.atleast_1d(np.asarray(value, dtype=np.float64)), (bs,))
        g[sdofs] = v[sdofs % bs]
    return DirichletBC(V, sdofs, g)
