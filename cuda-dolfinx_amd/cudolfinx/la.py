"""Device matrix and vector wrappers (python/cudolfinx/la.py).

Without PETSc-HIP in this environment, `CUDAMatrix.mat` and
`CUDAVector.vector` hand back light objects with the petsc4py calls the
reference tests use (`getValuesCSR`, `array`, `getSize`, `norm`).  With
petsc4py present, `petsc_mat()` wraps the same CSR arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .context import DeviceArray


class DeviceCSRMatrix:
    """The petsc4py-like view of an assembled device CSR matrix."""

    def __init__(self, owner: "CUDAMatrix"):
        self._owner = owner

    def getSize(self):
        return (self._owner.nrows, self._owner.ncols)

    def getValuesCSR(self):
        """(indptr, indices, data) copied to host, as petsc4py Mat.getValuesCSR."""
        return self._owner.csr()

    def assemble(self):
        self._owner.assemble()


class CUDAMatrix:
    """Device CSR matrix (la/CUDAMatrix.cpp:36-225).

    CUDAMatrix(ctx, form) builds the pattern of a bilinear form's integrals.
    CUDAMatrix(ctx, mat) adopts the pattern of a caller's matrix, as the
    reference's CUDAMatrix(ctx, petsc_mat) does (python/cudolfinx/la.py:45-51):

    * a torch sparse CSR tensor on the GPU (float64 values): its values are
      assembled in place (borrowed device memory);
    * a scipy.sparse CSR matrix, or an object with petsc4py's
      getValuesCSR() / setValuesCSR(): values live on the device and are
      written back to the caller's matrix at the handoff (assemble()).

    The device handle is created at the first assembly, against that form's
    spaces (the reference builds its offsets at assembly time too); every
    entry the form touches must be in the pattern (MF_ENOTFOUND otherwise).
    """

    def __init__(self, ctx, form_or_mat):
        self._ctx = ctx
        self._handle = None
        self._foreign = None
        if hasattr(form_or_mat, "handle") and hasattr(form_or_mat, "function_spaces"):
            form = form_or_mat
            self._form = form            # keeps the form (and its dofmaps) alive
            h = C.c_void_p()
            check(lib().mf_matrix_create(form.handle, C.byref(h)))
            self._set_handle(h)
        else:
            self._adopt(form_or_mat)
        self._mat = DeviceCSRMatrix(self) if self._foreign is None else form_or_mat

    def _set_handle(self, h):
        self._handle = h
        nr, nc, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mf_matrix_info(h, C.byref(nr), C.byref(nc), C.byref(nnz)))
        self.nrows, self.ncols, self.nnz = nr.value, nc.value, nnz.value

    def _adopt(self, mat):
        """Record a caller matrix's CSR pattern (and device values if it has them)."""
        dev_vals = None
        if hasattr(mat, "crow_indices") and getattr(mat, "layout", None) is not None:
            import torch
            if mat.layout != torch.sparse_csr or mat.dtype != torch.float64 or not mat.is_cuda:
                raise TypeError("a torch matrix must be a float64 sparse CSR tensor on the GPU")
            rp = mat.crow_indices().cpu().numpy()
            cols = mat.col_indices().cpu().numpy()
            dev_vals = mat.values()
            if not dev_vals.is_contiguous():
                raise ValueError("the CSR values must be contiguous")
            kind, shape = "torch", tuple(mat.shape)
        elif hasattr(mat, "indptr") and hasattr(mat, "indices"):
            if getattr(mat, "format", "csr") != "csr":
                raise TypeError("expected a CSR matrix")
            mat.sort_indices()
            rp, cols, kind, shape = mat.indptr, mat.indices, "scipy", mat.shape
        elif hasattr(mat, "getValuesCSR"):
            rp, cols, _ = mat.getValuesCSR()
            kind, shape = "petsc", tuple(mat.getSize())
        else:
            raise TypeError(f"Expected a CUDAForm or a CSR matrix, got '{type(mat)}'")
        self._foreign = dict(kind=kind, mat=mat, values=dev_vals, shape=shape,
                             rp=np.ascontiguousarray(rp, dtype=np.int32),
                             cols=np.ascontiguousarray(cols, dtype=np.int32))
        self.nrows, self.ncols = int(shape[0]), int(shape[1])
        self.nnz = int(self._foreign["rp"][-1])

    def _bind(self, form):
        """Create the device handle against `form` (first assembly of an adopted matrix)."""
        if self._handle is not None:
            return
        f = self._foreign
        if (self.nrows, self.ncols) != (form.function_spaces[0].size, form.function_spaces[1].size):
            raise ValueError(f"matrix shape {(self.nrows, self.ncols)} does not match the form's spaces")
        self._form = form
        vals = C.c_void_p(f["values"].data_ptr()) if f["values"] is not None else None
        h = C.c_void_p()
        check(lib().mf_matrix_create_from_csr(form.handle, f["rp"].ctypes.data_as(C.POINTER(C.c_int32)),
                                              f["cols"].ctypes.data_as(C.POINTER(C.c_int32)),
                                              self.nrows, self.nnz, vals, C.byref(h)))
        self._set_handle(h)

    @classmethod
    def block(cls, ctx, block_form):
        """Monolithic matrix of a BlockCUDAForm (create_cuda_matrix_block, fem/petsc.h:39-107)."""
        self = cls.__new__(cls)
        self._ctx = ctx
        self._foreign = None
        self._form = block_form
        rows = block_form.forms
        nr, nc = len(rows), len(rows[0])
        arr = (C.c_void_p * (nr * nc))(*[f.handle.value if f is not None else None
                                         for row in rows for f in row])
        h = C.c_void_p()
        check(lib().mf_matrix_create_block(arr, nr, nc, C.byref(h)))
        self._handle = h
        nrw, ncl, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mf_matrix_info(h, C.byref(nrw), C.byref(ncl), C.byref(nnz)))
        self.nrows, self.ncols, self.nnz = nrw.value, ncl.value, nnz.value
        self._mat = DeviceCSRMatrix(self)
        return self

    @property
    def handle(self):
        if self._handle is None:
            raise RuntimeError("an adopted matrix gets its device handle at its first assembly")
        return self._handle

    @property
    def mat(self):
        return self._mat

    def device_arrays(self):
        """(row_ptr, cols, values) device addresses."""
        rp, cols, vals = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().mf_matrix_device_arrays(self._handle, C.byref(rp), C.byref(cols), C.byref(vals)))
        return rp.value, cols.value, vals.value

    def assemble(self):
        """Handoff: make the device values final (CUDAMatrix::to_host + MatAssembly).
        An adopted host matrix (scipy, petsc4py-like) receives the values."""
        self._ctx.synchronize()
        f = self._foreign
        if f is None or self._handle is None or f["kind"] == "torch":
            return
        _, _, vals = self.csr()
        if f["kind"] == "scipy":
            f["mat"].data[:] = vals
        else:
            f["mat"].setValuesCSR(f["rp"], f["cols"], vals)
            f["mat"].assemble()

    @staticmethod
    def _addr(v):
        return C.c_void_p(v.data_ptr() if hasattr(v, "data_ptr") else int(v))

    def mult(self, x, y, stream=None):
        """y = A x on the device (x, y: float64 torch tensors or device addresses),
        the MatMult of a PETSc solve on the handed-off matrix."""
        if hasattr(x, "numel") and (x.numel() != self.ncols or y.numel() != self.nrows):
            raise ValueError("vector sizes do not match the matrix")
        check(lib().mf_spmv(self._handle, self._addr(x), self._addr(y), stream))
        return y

    def diagonal(self, d, stream=None):
        """d = diag(A) on the device (the Jacobi preconditioner)."""
        if hasattr(d, "numel") and d.numel() != self.nrows:
            raise ValueError("vector size does not match the matrix")
        check(lib().mf_matrix_diagonal(self._handle, self._addr(d), stream))
        return d

    def csr(self):
        rp = np.empty(self.nrows + 1, dtype=np.int32)
        cols = np.empty(max(self.nnz, 1), dtype=np.int32)
        vals = np.empty(max(self.nnz, 1), dtype=np.float64)
        check(lib().mf_matrix_copy_csr(self._handle, rp.ctypes.data_as(C.POINTER(C.c_int32)),
                                       cols.ctypes.data_as(C.POINTER(C.c_int32)),
                                       vals.ctypes.data_as(C.POINTER(C.c_double))))
        return rp, cols[:self.nnz], vals[:self.nnz]

    def to_scipy(self):
        import scipy.sparse as sp
        rp, cols, vals = self.csr()
        return sp.csr_matrix((vals, cols, rp), shape=(self.nrows, self.ncols))

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_matrix_destroy(self._handle)
            except Exception:
                pass
            self._handle = None


class DeviceVector:
    """petsc4py-like view of a device vector."""

    def __init__(self, owner: "CUDAVector"):
        self._owner = owner

    @property
    def array(self) -> np.ndarray:
        return self._owner.data.to_numpy()

    def getSize(self):
        return self._owner.size

    def norm(self):
        return float(np.linalg.norm(self.array))


class CUDAVector:
    """Device vector (la/CUDAVector.cpp:35-204)."""

    def __init__(self, ctx, size: int):
        self._ctx = ctx
        self.size = int(size)
        self.data = DeviceArray(ctx, self.size)
        check(lib().mf_zero(ctx.handle, self.data.ptr, self.size, None))
        self._vec = DeviceVector(self)

    @property
    def vector(self):
        return self._vec

    @property
    def ptr(self):
        return self.data.ptr

    def to_host(self):
        return self.data.to_numpy()
