"""Several element groups' matrices assembled by one launch.

`CUDABatch([a0, a1, ...])` takes compiled bilinear forms, one element group
each (for example the degree groups of many independent meshes merged per
degree), each with one cell integral on the owner-computes path (affine
simplices).  `assemble(asm, [A0, A1, ...], [bcs0, bcs1, ...])` writes into
every group's matrix what `CUDAAssembler.assemble_matrix` writes, from one
per-cell record pass per group and ONE persistent kernel that walks every
group's row chunks.  The reference has no counterpart: it launches one kernel
per form and integral type (python/cudolfinx/assemble.py:52-112).
"""
from __future__ import annotations

import ctypes as C
import typing

from ._lib import check, lib
from .form import CUDAForm


class CUDABatch:
    def __init__(self, forms: typing.Sequence[CUDAForm]):
        forms = list(forms)
        if not forms:
            raise ValueError("a batch needs at least one form")
        for f in forms:
            if not isinstance(f, CUDAForm):
                raise TypeError(f"Expected CUDAForm, got '{type(f)}'")
        self._forms = forms
        arr = (C.c_void_p * len(forms))(*[f.handle.value for f in forms])
        h = C.c_void_p()
        check(lib().mf_batch_create(arr, len(forms), C.byref(h)))
        self._handle = h

    @property
    def handle(self):
        return self._handle

    @property
    def forms(self):
        return self._forms

    def source(self) -> str:
        n = C.c_int64(0)
        check(lib().mf_batch_source(self._handle, None, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().mf_batch_source(self._handle, buf, C.byref(n)))
        return buf.value.decode()

    def assemble(self, asm, mats, bcs=None, diagonal: float = 1.0, zero: bool = True,
                 pack: bool = True):
        """Assemble form g into mats[g] with the Dirichlet rows and columns of
        bcs[g] (a CUDADirichletBC or a list of DirichletBC; None for none)."""
        if len(mats) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} matrices, got {len(mats)}")
        if bcs is not None and len(bcs) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} bc sets, got {len(bcs)}")
        hs = []
        for g, a in enumerate(self._forms):
            if pack:
                a.to_device(force=False)
                asm.pack_coefficients(a)
            if bcs is None or bcs[g] is None:
                hs.append(None)
                continue
            coll = asm._bc_collection(list(bcs[g]) if isinstance(bcs[g], tuple) else bcs[g])
            hs.append(coll._get_cpp_bcs(a.function_spaces[0]).handle.value)
        n = len(self._forms)
        A = (C.c_void_p * n)(*[m.handle.value for m in mats])
        B = (C.c_void_p * n)(*hs)
        check(lib().mf_batch_assemble(self._handle, A, B, float(diagonal), int(bool(zero)),
                                      asm._stream))
        return mats

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_batch_destroy(self._handle)
            except Exception:
                pass
            self._handle = None
