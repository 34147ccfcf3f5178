"""Several element groups' matrices, or right-hand sides, assembled by one launch.

`CUDABatch([a0, a1, ...])` takes compiled bilinear forms, one element group
each (for example the degree groups of many independent meshes merged per
degree), each with one cell integral on the owner-computes path (affine
simplices).  `assemble(asm, [A0, A1, ...], [bcs0, bcs1, ...])` writes into
every group's matrix what `CUDAAssembler.assemble_matrix` writes, from one
per-cell record pass per group and ONE persistent kernel that walks every
group's row chunks.

`CUDABatch([L0, L1, ...])` with linear forms (each with a cell integral on
the owner vector path) is the right-hand-side counterpart:
`assemble_vectors(asm, [b0, b1, ...])` writes what
`CUDAAssembler.assemble_vector` writes, with every form's cell integral in
ONE launch, and `set_bc(asm, [b0, ...], [bcs0, ...])` sets the Dirichlet
values of all vectors in one launch.  The reference has no counterpart: it
launches one kernel per form and integral type
(python/cudolfinx/assemble.py:52-112, 136-173).
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

    @property
    def rank(self) -> int:
        return self._forms[0].rank

    def assemble_vectors(self, asm, vecs, zero: bool = True, pack: bool = True):
        """Assemble linear form g into vecs[g] (CUDAVector); one launch for all
        cell integrals, then each form's other integrals."""
        if self.rank != 1:
            raise ValueError("assemble_vectors needs a batch of linear forms")
        if len(vecs) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} vectors, got {len(vecs)}")
        if pack:
            for L in self._forms:
                L.to_device(force=False)
                asm.pack_coefficients(L)
        n = len(vecs)
        ptrs = (C.c_void_p * n)(*[v.ptr.value for v in vecs])
        sizes = (C.c_int64 * n)(*[v.size for v in vecs])
        check(lib().mf_batch_assemble_vector(self._handle, ptrs, sizes, int(bool(zero)),
                                             asm._stream))
        return vecs

    def set_bc(self, asm, vecs, bcs, scale: float = 1.0):
        """CUDAAssembler.set_bc for every vector in one launch: vecs[g] takes
        the values of bcs[g] (a CUDADirichletBC, a list of DirichletBC, or None)
        on the space of linear form g."""
        if len(vecs) != len(self._forms) or len(bcs) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} vectors and bc sets")
        hs, colls = [], []          # colls: a list of DirichletBC packs a new collection,
        for g, L in enumerate(self._forms):       # which must outlive the call
            if bcs[g] is None:
                hs.append(None)
                continue
            coll = asm._bc_collection(bcs[g])
            colls.append(coll)
            hs.append(coll._get_cpp_bcs(L.function_spaces[0]).handle.value)
        n = len(vecs)
        ptrs = (C.c_void_p * n)(*[v.ptr.value for v in vecs])
        sizes = (C.c_int64 * n)(*[v.size for v in vecs])
        B = (C.c_void_p * n)(*hs)
        check(lib().mf_set_bc_batch(asm._ctx.handle, n, ptrs, sizes, B, float(scale), asm._stream))
        del colls
        return vecs

    def assemble(self, asm, mats, bcs=None, diagonal: float = 1.0, zero: bool = True,
                 pack: bool = True):
        """Assemble form g into mats[g] with the Dirichlet rows and columns of
        bcs[g] (a CUDADirichletBC or a list of DirichletBC; None for none)."""
        if len(mats) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} matrices, got {len(mats)}")
        if bcs is not None and len(bcs) != len(self._forms):
            raise ValueError(f"expected {len(self._forms)} bc sets, got {len(bcs)}")
        hs, colls = [], []          # the collections must outlive the call (see set_bc)
        for g, a in enumerate(self._forms):
            if pack:
                a.to_device(force=False)
                asm.pack_coefficients(a)
            if bcs is None or bcs[g] is None:
                hs.append(None)
                continue
            coll = asm._bc_collection(list(bcs[g]) if isinstance(bcs[g], tuple) else bcs[g])
            colls.append(coll)
            hs.append(coll._get_cpp_bcs(a.function_spaces[0]).handle.value)
        n = len(self._forms)
        A = (C.c_void_p * n)(*[m.handle.value for m in mats])
        B = (C.c_void_p * n)(*hs)
        check(lib().mf_batch_assemble(self._handle, A, B, float(diagonal), int(bool(zero)),
                                      asm._stream))
        del colls
        return mats

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_batch_destroy(self._handle)
            except Exception:
                pass
            self._handle = None
