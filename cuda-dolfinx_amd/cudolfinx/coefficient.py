"""Device coefficients and same-map interpolation (python/cudolfinx/coefficient.py).

`Coefficient(f)` keeps a device copy of a Function's dof vector.
`interpolate(coeff0)` replaces it with the interpolant of another Coefficient
on the same mesh and reference map, computed on the device.  Reference:
CUDACoefficient::interpolate (cpp/cudolfinx/fem/CUDACoefficient.h:50-66)
and interpolate_same_map (CUDAInterpolate.cu:48-74).

The element interpolation operator (what DOLFINx's
create_interpolation_operator returns: for Lagrange elements, the source
basis evaluated at the target's reference nodes) and the target nodes'
owning cells are uploaded once per (target, source) space pair.  Every call
is then a single kernel launch with no allocation, so it can sit inside a
captured HIP graph.  A target node shared by several cells takes its value
from the highest-numbered one, as the last write of a sequential cell loop
would.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .context import DeviceArray, get_cuda_context
from .form import _create_dofmap_on_device


def interpolation_operator(el_to, el_from) -> np.ndarray:
    """(n_to, n_from): basis functions of el_from at the reference nodes of el_to."""
    if el_to.cell != el_from.cell:
        raise ValueError("elements on different reference cells")
    return np.ascontiguousarray(el_from.tabulate(el_to.nodes, 0)[0], dtype=np.float64)


class _Interpolator:
    def __init__(self, ctx, V1, V0):
        self._ctx = ctx
        self._dm = (_create_dofmap_on_device(V1), _create_dofmap_on_device(V0))
        im = interpolation_operator(V1.element, V0.element)
        h = C.c_void_p()
        check(lib().mf_interpolator_create(ctx.handle, self._dm[0].handle, self._dm[1].handle,
                                           im.ctypes.data_as(C.c_void_p), C.byref(h)))
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            try:
                lib().mf_interpolator_destroy(self.handle)
            except Exception:
                pass
            self.handle = None


class Coefficient:
    """Device wrapper of a Function (CUDACoefficient.h:19-46)."""

    def __init__(self, f):
        self._ctx = get_cuda_context()
        self._f = f
        self._values = DeviceArray.from_numpy(self._ctx, np.asarray(f.x, dtype=np.float64))
        self._interp = {}

    @property
    def function(self):
        return self._f

    @property
    def device_values(self) -> DeviceArray:
        return self._values

    def copy_host_values_to_device(self) -> None:
        self._values.upload(np.asarray(self._f.x, dtype=np.float64))

    def copy_device_values_to_host(self) -> None:
        """Write the device values into the Function's host vector."""
        self._f.x[:] = self.values()
        self._f.version += 1

    def interpolate(self, coeff0: "Coefficient", stream=None) -> None:
        """Interpolate from coeff0 (same mesh, same reference map) on the device."""
        if not isinstance(coeff0, Coefficient):
            raise TypeError(f"Expected Coefficient, got '{type(coeff0)}'")
        V1, V0 = self._f.function_space, coeff0._f.function_space
        if V1.mesh is not V0.mesh:
            raise ValueError("same-map interpolation needs both functions on one mesh")
        if V1.bs != V0.bs:
            raise ValueError("block sizes differ")
        if coeff0 is self:
            raise ValueError("interpolation in place is not supported")
        ip = self._interp.get(id(V0))
        if ip is None or ip[0] is not V0:
            ip = (V0, _Interpolator(self._ctx, V1, V0))
            self._interp[id(V0)] = ip
        s = None if stream is None else C.c_void_p(getattr(stream, "cuda_stream", stream))
        check(lib().mf_interpolate(ip[1].handle, self._values.ptr, coeff0._values.ptr, s))

    def values(self) -> np.ndarray:
        """A host copy of the device dof vector (synchronizes)."""
        self._ctx.synchronize()
        return self._values.to_numpy()
