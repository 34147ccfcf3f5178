"""GPU assembly front end (python/cudolfinx/assemble.py).

Same call sequence and argument meaning as the reference's CUDAAssembler:
zero -> pack -> assemble -> diagonal for matrices; zero -> pack -> assemble
for vectors; apply_lifting and set_bc on device vectors.  Differences by
design (documented quirks of the reference): `diagonal` is honoured (Q1),
coefficients are re-uploaded only when their host values changed (Q2),
lifting covers every integral type (Q3), `scale` stays float64 (Q10).
"""
from __future__ import annotations

import ctypes as C
import typing

from ._lib import check, lib
from .bcs import CUDADirichletBC
from .context import get_cuda_context
from .form import BlockCUDAForm, CUDAForm
from .la import CUDAMatrix, CUDAVector


def _stream(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return C.c_void_p(stream.cuda_stream)
    return C.c_void_p(int(stream))


class CUDAAssembler:
    """Assembly on the GPU."""

    def __init__(self, stream=None):
        self._ctx = get_cuda_context()
        self._stream = _stream(stream)
        self._no_bcs = None

    # ------------------------------------------------------------ creation
    def create_matrix(self, a: CUDAForm) -> CUDAMatrix:
        if not isinstance(a, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got type '{type(a)}'.")
        return CUDAMatrix(self._ctx, a)

    def create_vector(self, b: CUDAForm) -> CUDAVector:
        if not isinstance(b, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got type '{type(b)}'.")
        return CUDAVector(self._ctx, b.function_spaces[0].size)

    def create_matrix_block(self, a: BlockCUDAForm) -> CUDAMatrix:
        """Monolithic block matrix (assemble.py:236-246)."""
        if not isinstance(a, BlockCUDAForm):
            raise TypeError(f"Expected BlockCUDAForm, got type '{type(a)}'")
        return CUDAMatrix.block(self._ctx, a)

    def create_vector_block(self, b: BlockCUDAForm) -> CUDAVector:
        """Block vector of size b.local_size (assemble.py:255-263)."""
        if not isinstance(b, BlockCUDAForm):
            raise TypeError(f"Expected BlockCUDAForm, got type '{type(b)}')")
        return CUDAVector(self._ctx, b.local_size)

    def pack_bcs(self, bcs) -> CUDADirichletBC:
        return CUDADirichletBC(self._ctx, bcs)

    def _bc_collection(self, bcs):
        if isinstance(bcs, CUDADirichletBC):
            return bcs
        if isinstance(bcs, (list, tuple)):
            if not bcs:
                # one reusable empty collection: no device allocation per call
                if self._no_bcs is None:
                    self._no_bcs = self.pack_bcs([])
                return self._no_bcs
            return self.pack_bcs(list(bcs))
        raise TypeError(f"Expected either a list of DirichletBC's or a CUDADirichletBC, got '{type(bcs)}'")

    def pack_coefficients(self, a: CUDAForm, coefficients=None):
        if not isinstance(a, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got type '{type(a)}'.")
        if coefficients is None:
            a.pack_coefficients(None, self._stream)
        elif len(coefficients):
            idx = []
            for c in coefficients:
                for j, f in enumerate(a.dolfinx_form.coefficients):
                    if f is c:
                        idx.append(j)
            a.pack_coefficients(idx, self._stream)

    # ------------------------------------------------------------ matrices
    def assemble_matrix(self, a: CUDAForm, mat: typing.Optional[CUDAMatrix] = None, bcs=(),
                        diagonal: float = 1.0, constants=None, coeffs=None, zero: bool = True):
        if not isinstance(a, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got '{type(a)}'")
        if mat is None:
            mat = self.create_matrix(a)
        mat._bind(a)                       # adopted caller matrices: handle on first use
        bc = self._bc_collection(list(bcs) if isinstance(bcs, tuple) else bcs)
        V0, V1 = a.function_spaces[0], a.function_spaces[1]
        b0, b1 = bc._get_cpp_bcs(V0), bc._get_cpp_bcs(V1)
        a.to_device(force=False)
        self.pack_coefficients(a, coeffs)
        check(lib().mf_assemble_matrix(a.handle, mat.handle, b0.handle, b1.handle, float(diagonal),
                                       int(bool(zero)), self._stream))
        return mat

    def assemble_matrix_block(self, a: BlockCUDAForm, mat: typing.Optional[CUDAMatrix] = None,
                              bcs=None, coeffs=None, constants=None, zero: bool = True):
        """Every block into one matrix (assemble.py:112-134): one zeroing, then
        assemble_matrix(form_ij, mat, bcs[i][j], zero=False) per block."""
        if not isinstance(a, BlockCUDAForm):
            raise TypeError(f"Expected BlockCUDAForm, got '{type(a)}'")
        if mat is None:
            mat = self.create_matrix_block(a)
        if zero:
            _, _, vals = mat.device_arrays()
            check(lib().mf_zero(self._ctx.handle, C.c_void_p(vals), mat.nnz, self._stream))
        for i, row in enumerate(a.forms):
            for j, form in enumerate(row):
                if form is None:
                    continue
                _bcs = bcs[i][j] if bcs is not None else []
                _coeffs = coeffs[i][j] if coeffs is not None else None
                self.assemble_matrix(form, mat, bcs=_bcs, coeffs=_coeffs, zero=False)
        return mat

    # ------------------------------------------------------------ vectors
    def assemble_vector(self, b: CUDAForm, vec: typing.Optional[CUDAVector] = None, constants=None,
                        coeffs=None, zero: bool = True):
        if not isinstance(b, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got '{type(b)}'")
        if vec is None:
            vec = self.create_vector(b)
        b.to_device(force=False)
        self.pack_coefficients(b, coeffs)
        check(lib().mf_assemble_vector(b.handle, vec.ptr, vec.size, int(bool(zero)), self._stream))
        return vec

    def assemble_vector_block(self, b: BlockCUDAForm, vec: typing.Optional[CUDAVector] = None,
                              constants=None, coeffs=None, zero: bool = True):
        """Every block form into one vector at its offset (assemble.py:168-202)."""
        if not isinstance(b, BlockCUDAForm):
            raise TypeError(f"Expected BlockCUDAForm, got '{type(b)}'")
        if vec is None:
            vec = self.create_vector_block(b)
        num_forms = len(b.forms)
        _coeffs = [None] * num_forms if coeffs is None else coeffs
        if len(_coeffs) != num_forms:
            raise ValueError(f"Expected coeffs to be None or a list of length: '{num_forms}', "
                             f"got '{len(_coeffs)}' instead!")
        if zero:
            check(lib().mf_zero(self._ctx.handle, vec.ptr, vec.size, self._stream))
        for form, fc in zip(b.forms, _coeffs):
            self.assemble_vector(form, vec=vec, coeffs=fc, zero=False)
        return vec

    # ------------------------------------------------------------ functionals
    def assemble_scalar(self, b: CUDAForm, constants=None, coeffs=None) -> float:
        """Value of a functional (rank-0 form) (assemble.py:204-226).

        Synchronizes the assembler's stream and returns a Python float; the
        per-workgroup partial sums are added in a fixed order, so the value
        repeats exactly from call to call.
        """
        if not isinstance(b, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got '{type(b)}'")
        if b.rank != 0:
            raise ValueError(f"assemble_scalar needs a rank-0 form, got rank {b.rank}")
        b.to_device(force=False)
        self.pack_coefficients(b, coeffs)
        out = C.c_double(0.0)
        check(lib().mf_assemble_scalar(b.handle, C.byref(out), self._stream))
        return out.value

    def assemble_scalar_device(self, b: CUDAForm, out, zero: bool = True, coeffs=None):
        """out[0] = (or +=) the functional's value, on the device, without a host
        synchronization; `out` is a device buffer (an object with `data_ptr()`
        such as a float64 torch tensor, or a raw address)."""
        if not isinstance(b, CUDAForm):
            raise TypeError(f"Expected CUDAForm, got '{type(b)}'")
        if b.rank != 0:
            raise ValueError(f"assemble_scalar needs a rank-0 form, got rank {b.rank}")
        b.to_device(force=False)
        self.pack_coefficients(b, coeffs)
        ptr = out.data_ptr() if hasattr(out, "data_ptr") else int(out)
        check(lib().mf_assemble_scalar_device(b.handle, C.c_void_p(ptr), int(bool(zero)),
                                              self._stream))
        return out

    def apply_lifting(self, b: CUDAVector, a: typing.List[CUDAForm], bcs: typing.List,
                      x0: typing.Optional[typing.List[CUDAVector]] = None, scale: float = 1.0,
                      coeffs=None):
        if len(a) != len(bcs):
            raise ValueError("Lengths of forms and bcs must match!")
        if x0 is not None and len(x0) != len(a):
            raise ValueError("Lengths of forms and x0 must match!")
        if coeffs is None:
            coeffs = [None] * len(a)
        elif not len(coeffs):
            coeffs = [[] for _ in a]
        handles, bch, keep = [], [], []
        for form, bc_list, fc in zip(a, bcs, coeffs):
            form.to_device(force=False)
            self.pack_coefficients(form, fc)
            coll = self._bc_collection(bc_list)
            keep.append(coll)
            handles.append(form.handle)
            bch.append(coll._get_cpp_bcs(form.function_spaces[1]).handle)
        n = len(a)
        fa = (C.c_void_p * n)(*handles)
        ba = (C.c_void_p * n)(*bch)
        xa = (C.c_void_p * n)(*[x.ptr for x in x0]) if x0 is not None else None
        check(lib().mf_apply_lifting(fa, n, b.ptr, b.size, ba, xa, float(scale), self._stream))

    def set_bc(self, b: CUDAVector, bcs, V, x0: typing.Optional[CUDAVector] = None,
               scale: float = 1.0):
        coll = self._bc_collection(bcs)
        d = coll._get_cpp_bcs(V)
        check(lib().mf_set_bc(self._ctx.handle, b.ptr, b.size, d.handle,
                              x0.ptr if x0 is not None else None, float(scale), self._stream))

    def synchronize(self):
        self._ctx.synchronize(self._stream)
