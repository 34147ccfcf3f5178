"""Device Dirichlet conditions (python/cudolfinx/bcs.py)."""
from __future__ import annotations

import ctypes as C
import typing

import numpy as np

from ._lib import check, lib


class _DeviceBC:
    """mf_bc handle for one function space (CUDADirichletBC.h:56-194)."""

    def __init__(self, ctx, V, bcs):
        self._ctx = ctx
        self.V = V
        n = V.size
        dofs = [np.asarray(bc.dofs, dtype=np.int32) for bc in bcs]
        d = np.unique(np.concatenate(dofs)) if dofs else np.zeros(0, dtype=np.int32)
        self.dofs = np.ascontiguousarray(d.astype(np.int32))
        self.g = self._dense(bcs, n)
        h = C.c_void_p()
        check(lib().mf_bc_create(ctx.handle, n, self.dofs.ctypes.data_as(C.POINTER(C.c_int32)),
                                 self.dofs.size, self.dofs.size,
                                 self.g.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)))
        self.handle = h

    @staticmethod
    def _dense(bcs, n):
        g = np.zeros(n)
        for bc in bcs:            # DirichletBC::set into a dense array, later bcs win
            g[bc.dofs] = bc.values[bc.dofs]
        return np.ascontiguousarray(g)

    def update(self, bcs):
        self.g = self._dense(bcs, self.V.size)
        check(lib().mf_bc_update(self.handle, self.g.ctypes.data_as(C.POINTER(C.c_double))))

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().mf_bc_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class CUDADirichletBC:
    """A collection of boundary conditions, grouped by function space."""

    def __init__(self, ctx, bcs: typing.List):
        self.bcs = list(bcs)
        self._ctx = ctx
        self._function_spaces = []
        self._bc_lists = []
        for bc in self.bcs:
            V = bc.function_space
            for i, W in enumerate(self._function_spaces):
                if W is V:
                    self._bc_lists[i].append(bc)
                    break
            else:
                self._function_spaces.append(V)
                self._bc_lists.append([bc])
        self._device_bcs = [_DeviceBC(ctx, V, l) for V, l in zip(self._function_spaces, self._bc_lists)]
        self._empty = {}

    def _get_cpp_bcs(self, V):
        for W, d in zip(self._function_spaces, self._device_bcs):
            if W is V:
                return d
        key = id(V)
        if key not in self._empty:
            self._empty[key] = _DeviceBC(self._ctx, V, [])
        return self._empty[key]

    def update(self, bcs: typing.Optional[typing.List] = None):
        """Re-upload values (time-dependent BCs), CUDADirichletBC.update (bcs.py:66-79)."""
        if bcs is None:
            bcs = self.bcs
        for W, d, l in zip(self._function_spaces, self._device_bcs, self._bc_lists):
            sel = [bc for bc in bcs if bc.function_space is W]
            if sel:
                # keep bcs not passed at their current values
                rest = [bc for bc in l if all(bc is not s for s in sel)]
                d.update(rest + sel)
