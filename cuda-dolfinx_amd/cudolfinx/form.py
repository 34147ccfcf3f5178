"""Device forms (python/cudolfinx/form.py).

`form(f)` wraps a `mifem.forms.Form` (the DOLFINx/FFCx stand-in) in a
`CUDAForm`: the mesh and dofmaps go to the device once (cached on the
Python objects, like `_create_mesh_on_device`'s functools.cache,
form.py:252-262), integrals are registered with their UFCx tabulate_tensor
sources, and the form is compiled with hipRTC.
"""
from __future__ import annotations

import collections.abc
import ctypes as C
import os
import typing
from pathlib import Path

import numpy as np

from ._lib import MF_CELL, MF_EXTERIOR_FACET, MF_INTERIOR_FACET, check, lib
from .context import get_cuda_context

DEFAULT_CUDA_JIT_ARGS = {
    "cachedir": str(Path(os.getenv("XDG_CACHE_HOME", Path.home() / ".cache")) / "mifem"),
    "verbose": False,
}

_ITYPE = {"cell": MF_CELL, "exterior_facet": MF_EXTERIOR_FACET, "interior_facet": MF_INTERIOR_FACET}


class _DeviceMesh:
    def __init__(self, ctx, mesh):
        self._ctx = ctx
        x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        cells = np.ascontiguousarray(mesh.cells, dtype=np.int32)
        fp = np.ascontiguousarray(mesh.facet_permutations(), dtype=np.uint8)
        h = C.c_void_p()
        check(lib().mf_mesh_create(ctx.handle, x.ctypes.data_as(C.POINTER(C.c_double)), x.shape[0],
                                   cells.ctypes.data_as(C.POINTER(C.c_int32)), cells.shape[0],
                                   cells.shape[1], fp.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   fp.shape[1], C.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().mf_mesh_destroy(self.handle)
            except Exception:
                pass


class _DeviceDofMap:
    def __init__(self, ctx, V):
        self._ctx = ctx
        dm = np.ascontiguousarray(V.dofmap, dtype=np.int32)
        h = C.c_void_p()
        check(lib().mf_dofmap_create(ctx.handle, dm.ctypes.data_as(C.POINTER(C.c_int32)),
                                     dm.shape[0], dm.shape[1], V.bs, V.num_nodes, C.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().mf_dofmap_destroy(self.handle)
            except Exception:
                pass


def _create_mesh_on_device(mesh):
    ctx = get_cuda_context()
    d = mesh._cache.get("cuda_mesh")
    if d is None:
        d = _DeviceMesh(ctx, mesh)
        mesh._cache["cuda_mesh"] = d
    return d


def _create_dofmap_on_device(V):
    d = getattr(V, "_cuda_dofmap", None)
    if d is None:
        d = _DeviceDofMap(get_cuda_context(), V)
        V._cuda_dofmap = d
    return d


class CUDAForm:
    """Device wrapper of a form (CUDAForm.h:29-268)."""

    def __init__(self, form, jit_args: typing.Optional[dict] = None, options=None,
                 ffcx_source=None):
        """ffcx_source: FFCx-generated C for this form (a .c path, the C text,
        or a DOLFINx-style form whose module's .c file holds it); its
        tabulate_tensor bodies then replace the form's own, matched to the
        form's integrals by the file's form_integrals order (the reference's
        names / sources / integral_tensor_indices, form.py:38-39)."""
        self._ctx = get_cuda_context()
        self._form = form
        self._cuda_mesh = _create_mesh_on_device(form.mesh)
        self._dofmaps = [_create_dofmap_on_device(V) for V in form.function_spaces]
        args = dict(DEFAULT_CUDA_JIT_ARGS)
        args.update(jit_args or {})
        h = C.c_void_p()
        dm0 = self._dofmaps[0].handle if form.rank >= 1 else None
        dm1 = self._dofmaps[1].handle if form.rank == 2 else None
        check(lib().mf_form_create(self._ctx.handle, self._cuda_mesh.handle, form.rank,
                                   dm0, dm1, C.byref(h)))
        self._handle = h
        self._entities = []
        sources = [(I.name, I.source) for I in form.integrals]
        if ffcx_source is not None:
            from .jit import get_wrapped_tabulate_tensors
            wrapped, order = get_wrapped_tabulate_tensors(ffcx_source)
            if len(order) != len(form.integrals):
                raise ValueError(f"the FFCx source has {len(order)} integrals, the form "
                                 f"{len(form.integrals)}")
            sources = [wrapped[i] for i in order]
        for I, (tname, tsrc) in zip(form.integrals, sources):
            ents = np.ascontiguousarray(I.entities, dtype=np.int32)
            self._entities.append(ents)
            staged = I.staged.source().encode() if I.staged is not None else None
            check(lib().mf_form_add_integral(h, _ITYPE[I.type], I.id, tname.encode(),
                                             tsrc.encode(),
                                             ents.ctypes.data_as(C.POINTER(C.c_int32)),
                                             I.num_entities, staged))
        self._coef_dofmaps = []
        self._coef_versions = []
        self._consts = None
        for f in form.coefficients:
            d = _create_dofmap_on_device(f.function_space)
            self._coef_dofmaps.append(d)
            idx = C.c_int()
            check(lib().mf_form_add_coefficient(h, d.handle, f.function_space.size, C.byref(idx)))
            self._coef_versions.append(-1)
        for key, val in (options or {}).items():
            self.set_option(key, val)
        os.environ.setdefault("MIFEM_CACHE_DIR", args["cachedir"])
        check(lib().mf_form_compile(h, int(bool(args.get("verbose")))))
        self.to_device()

    # ------------------------------------------------------------ properties
    @property
    def handle(self):
        return self._handle

    @property
    def cuda_form(self):
        return self

    @property
    def cuda_mesh(self):
        return self._cuda_mesh

    @property
    def dolfinx_form(self):
        return self._form

    @property
    def function_spaces(self):
        return self._form.function_spaces

    @property
    def rank(self):
        return self._form.rank

    def set_restriction(self, axis: int, node_map, offset: int) -> None:
        """Number argument `axis` in a block space: node r -> scalar dofs
        offset + bs * node_map[r] + j, node_map[r] == -1 drops the node
        (node_map None: identity).  CUDAForm.set_restriction, CUDAForm.h:201-231."""
        V = self.function_spaces[axis]
        if node_map is None:
            check(lib().mf_form_set_restriction(self._handle, axis, None, V.num_nodes, int(offset)))
        else:
            m = np.ascontiguousarray(node_map, dtype=np.int32)
            check(lib().mf_form_set_restriction(self._handle, axis, m.ctypes.data_as(C.c_void_p),
                                                m.size, int(offset)))

    def set_option(self, key: str, value: int) -> None:
        check(lib().mf_form_set_option(self._handle, key.encode(), int(value)))

    def source(self) -> str:
        n = C.c_int64(0)
        check(lib().mf_form_source(self._handle, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().mf_form_source(self._handle, buf, C.byref(n)))
        return buf.value.decode()

    # ------------------------------------------------------------ data movement
    def to_device(self, force: bool = True) -> None:
        """Copy host coefficient values and constants to the device (CUDAForm::to_device).

        force=False skips coefficients whose host values did not change since
        the last upload (Function.version), keeping them device-resident.
        """
        for j, f in enumerate(self._form.coefficients):
            if not force and self._coef_versions[j] is None:
                continue                      # bound to caller device memory
            if force or f.version != self._coef_versions[j]:
                x = np.ascontiguousarray(f.x, dtype=np.float64)
                check(lib().mf_form_set_coefficient(self._handle, j, x.ctypes.data_as(C.c_void_p),
                                                    x.size, 0))
                self._coef_versions[j] = f.version
        c = np.ascontiguousarray(self._form.constant_values(), dtype=np.float64)
        if force or self._consts is None or not np.array_equal(c, self._consts):
            check(lib().mf_form_set_constants(self._handle,
                                              c.ctypes.data_as(C.POINTER(C.c_double)), c.size))
            self._consts = c.copy()

    def set_coefficient_device(self, j: int, address: int, n: int) -> None:
        """Point coefficient j at caller-owned device memory (no copy)."""
        check(lib().mf_form_set_coefficient(self._handle, j, C.c_void_p(address), n, 1))
        self._coef_versions[j] = None

    def pack_coefficients(self, indices=None, stream=None) -> None:
        if indices is None:
            check(lib().mf_form_pack_coefficients(self._handle, None, 0, stream))
        else:
            arr = (C.c_int * len(indices))(*indices)
            check(lib().mf_form_pack_coefficients(self._handle, arr, len(indices), stream))

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_form_destroy(self._handle)
            except Exception:
                pass
            self._handle = None


class BlockCUDAForm:
    """Several forms assembled as one block system (form.py:107-225).

    forms: a list of linear forms (block vector) or a list of lists of
    bilinear forms (block matrix; None marks an empty block).  restrictions:
    for a vector, one array per block of the kept nodes of its space (the
    position in the array is the node's new number); for a matrix, a pair
    (row restrictions, column restrictions); None keeps every node.  Block i
    starts at the sum of the kept sizes (nodes x block size) of blocks < i.
    """

    def __init__(self, forms, restrictions=None):
        self._forms = forms
        self._restrictions = restrictions
        if not len(forms):
            raise ValueError("Must provide at least one form!")
        if isinstance(forms[0], CUDAForm):
            self._init_vector()
        else:
            self._init_matrix()

    @staticmethod
    def _maps_offsets(spaces, restrictions):
        maps, offsets = [], [0]
        for i, V in enumerate(spaces):
            r = None if restrictions is None else restrictions[i]
            if r is None:
                maps.append(None)
                kept = V.num_nodes
            else:
                r = np.asarray(r, dtype=np.int64)
                if r.size != np.unique(r).size or (r.size and (r.min() < 0 or r.max() >= V.num_nodes)):
                    raise ValueError(f"restriction {i}: node indices must be unique and in range")
                m = np.full(V.num_nodes, -1, dtype=np.int32)
                m[r] = np.arange(r.size, dtype=np.int32)
                maps.append(m)
                kept = r.size
            offsets.append(offsets[-1] + kept * V.bs)
        return maps, offsets

    def _init_vector(self):
        spaces = [f.function_spaces[0] for f in self._forms]
        maps, self._offsets = self._maps_offsets(spaces, self._restrictions)
        for f, m, off in zip(self._forms, maps, self._offsets):
            f.set_restriction(0, m, off)
        self._col_offsets = None
        self._global_size = self._offsets[-1]

    def _init_matrix(self):
        nr, nc = len(self._forms), len(self._forms[0])
        if any(len(row) != nc for row in self._forms):
            raise ValueError("block matrix rows must have equal length")

        def space(i, axis):
            cells = self._forms[i] if axis == 0 else [row[i] for row in self._forms]
            for f in cells:
                if f is not None:
                    return f.function_spaces[axis]
            raise ValueError(f"{'row' if axis == 0 else 'column'} {i} has no form")

        rr = self._restrictions[0] if self._restrictions is not None else None
        cr = self._restrictions[1] if self._restrictions is not None else None
        rmaps, self._offsets = self._maps_offsets([space(i, 0) for i in range(nr)], rr)
        cmaps, self._col_offsets = self._maps_offsets([space(j, 1) for j in range(nc)], cr)
        for i, row in enumerate(self._forms):
            for j, f in enumerate(row):
                if f is None:
                    continue
                f.set_restriction(0, rmaps[i], self._offsets[i])
                f.set_restriction(1, cmaps[j], self._col_offsets[j])
        self._global_size = self._offsets[-1]

    @property
    def forms(self):
        return self._forms

    @property
    def dolfinx_forms(self):
        return [f.dolfinx_form if isinstance(f, CUDAForm) else [g.dolfinx_form if g else None for g in f]
                for f in self._forms]

    @property
    def offsets(self):
        return self._offsets

    @property
    def col_offsets(self):
        return self._col_offsets

    @property
    def local_size(self):
        return self._offsets[-1]

    @property
    def global_size(self):
        return self._global_size


def form(form, restriction=None, cuda_jit_args: typing.Optional[dict] = None,
         options: typing.Optional[dict] = None, ffcx_source=None, **kwargs):
    """CUDAForm of a mifem form; a list (of lists) of forms gives a BlockCUDAForm
    with the given restrictions (form.py:227-249).  ffcx_source (single forms):
    FFCx-generated C whose tabulate_tensor bodies the form uses (CUDAForm)."""
    def _create(f):
        if f is None:
            return None
        if isinstance(f, collections.abc.Iterable) and not hasattr(f, "integrals"):
            return [_create(g) for g in f]
        return CUDAForm(f, jit_args=cuda_jit_args, options=options, ffcx_source=ffcx_source)

    cuda_form = _create(form)
    if isinstance(cuda_form, list):
        return BlockCUDAForm(cuda_form, restriction)
    return cuda_form
