"""A Newton step's forms assembled in one launch.

`CUDASystem(J, F, E)` binds a Jacobian J (bilinear), its residual F (linear)
and optionally a functional E (energy or merit function) that share their
integrals, entities and coefficients (the state u), on a discontinuous space
numbered cell by cell.  `assemble` runs one persistent kernel over the cell,
exterior-facet and interior-facet integrals of all three forms: the
reference assembles each form, and each integral type, in separate launches
(python/cudolfinx/assemble.py:52-226).  The call allocates nothing after
the first use of a matrix and does not synchronize, so it can be captured in
a `CUDAGraph`; E is summed in a fixed order and repeats exactly.
"""
from __future__ import annotations

import ctypes as C
import typing

from ._lib import check, lib
from .form import CUDAForm


class CUDASystem:
    def __init__(self, a: CUDAForm, L: CUDAForm, M: typing.Optional[CUDAForm] = None,
                 block: int = 0):
        for f, name in ((a, "a"), (L, "L")):
            if not isinstance(f, CUDAForm):
                raise TypeError(f"Expected CUDAForm for {name}, got '{type(f)}'")
        if M is not None and not isinstance(M, CUDAForm):
            raise TypeError(f"Expected CUDAForm or None for M, got '{type(M)}'")
        self._forms = (a, L, M)
        h = C.c_void_p()
        check(lib().mf_system_create(a.handle, L.handle, M.handle if M is not None else None,
                                     int(block), C.byref(h)))
        self._handle = h
        self.block = int(block) if block else 64

    @property
    def handle(self):
        return self._handle

    @property
    def forms(self):
        return self._forms

    def source(self) -> str:
        n = C.c_int64(0)
        check(lib().mf_system_source(self._handle, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().mf_system_source(self._handle, buf, C.byref(n)))
        return buf.value.decode()

    def assemble(self, asm, A, b, value=None, bcs=(), diagonal: float = 1.0, zero: bool = True,
                 accumulate_value: bool = False, pack: bool = True):
        """A, b (and value[0], a device float64 buffer) for the current state.

        pack=False skips re-packing the coefficients (they are read from the
        Jacobian form's packed array, e.g. when the caller packed already)."""
        a, L, M = self._forms
        if M is not None and value is None:
            raise ValueError("the system has a functional: pass a device buffer for its value")
        if pack:
            a.to_device(force=False)
            asm.pack_coefficients(a)
            if M is not None:
                M.to_device(force=False)
            L.to_device(force=False)
        coll = asm._bc_collection(list(bcs) if isinstance(bcs, tuple) else bcs)
        V = a.function_spaces[0]
        b0 = coll._get_cpp_bcs(V)
        vptr = None
        if value is not None:
            vptr = C.c_void_p(value.data_ptr() if hasattr(value, "data_ptr") else int(value))
        flags = (1 if zero else 0) | (2 if accumulate_value else 0)
        check(lib().mf_system_assemble(self._handle, A.handle, b0.handle, b0.handle, float(diagonal),
                                       b.ptr, b.size, vptr, flags, asm._stream))
        return A, b

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_system_destroy(self._handle)
            except Exception:
                pass
            self._handle = None
