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

    def __init__(self, form, jit_args: typing.Optional[dict] = None, options=None):
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
        for I in form.integrals:
            ents = np.ascontiguousarray(I.entities, dtype=np.int32)
            self._entities.append(ents)
            staged = I.staged.source().encode() if I.staged is not None else None
            check(lib().mf_form_add_integral(h, _ITYPE[I.type], I.id, I.name.encode(),
                                             I.source.encode(),
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


def form(form, cuda_jit_args: typing.Optional[dict] = None, options: typing.Optional[dict] = None,
         **kwargs):
    """Create a CUDAForm (or nested lists of them) from mifem forms (form.py:227-249)."""
    if isinstance(form, collections.abc.Iterable) and not hasattr(form, "integrals"):
        return [globals()["form"](f, cuda_jit_args, options, **kwargs) for f in form]
    return CUDAForm(form, jit_args=cuda_jit_args, options=options)
