"""Jacobi-preconditioned conjugate gradients on an assembled device matrix.

Stands in for the PETSc KSP (cg + jacobi) the reference's examples hand the
assembled Mat/Vec to (python/examples/poisson.py:116-125) when petsc4py is
absent.  Given a CUDAMatrix, the products and the diagonal run on the
engine's own CSR kernels (mf_spmv / mf_matrix_diagonal) on the matrix's device
arrays; a torch sparse CSR tensor also works.  The iteration runs in FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np


def csr_tensor(A, values=None):
    """torch sparse CSR view of a CUDAMatrix (values: torch tensor or None = copy).

    The index arrays are copied once per matrix and reused by later calls.
    """
    import torch
    dev = torch.device("cuda")
    st = getattr(A, "_torch_structure", None)
    if st is None or values is None:
        rp, cols, vals = A.csr()
        st = (torch.from_numpy(rp.astype(np.int64)).to(dev),
              torch.from_numpy(cols.astype(np.int64)).to(dev))
        A._torch_structure = st
        if values is None:
            values = torch.from_numpy(vals).to(dev)
    return torch.sparse_csr_tensor(st[0], st[1], values, size=(A.nrows, A.ncols))


def pcg(M, b, rtol: float = 1e-10, maxiter: int = 5000):
    """Solve M x = b with Jacobi-PCG.  M: a CUDAMatrix (native CSR SpMV and
    diagonal, mf_spmv / mf_matrix_diagonal, on torch's current stream) or a
    torch sparse CSR tensor; b: a float64 torch vector.

    Returns (x, iterations, final relative residual).
    """
    import torch
    native = hasattr(M, "mult")
    if native:
        stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        d = torch.empty(M.nrows, dtype=torch.float64, device=b.device)
        M.diagonal(d, stream)
        Ap = torch.empty_like(b)

        def matvec(p):
            M.mult(p, Ap, stream)
            return Ap
    else:
        d = torch.zeros(M.shape[0], dtype=torch.float64, device=b.device)
        crow, col, val = M.crow_indices(), M.col_indices(), M.values()
        rows = torch.repeat_interleave(torch.arange(M.shape[0], device=b.device),
                                       crow[1:] - crow[:-1])
        diag_mask = rows == col
        d.index_put_((rows[diag_mask],), val[diag_mask], accumulate=True)

        def matvec(p):
            return torch.mv(M, p)
    dinv = torch.where(d != 0, 1.0 / d, torch.zeros_like(d))
    x = torch.zeros_like(b)
    r = b.clone()
    z = dinv * r
    p = z.clone()
    rz = torch.dot(r, z)
    bn = torch.linalg.norm(b)
    it = 0
    res = 1.0
    for it in range(1, maxiter + 1):
        Ap = matvec(p)
        alpha = rz / torch.dot(p, Ap)
        x += alpha * p
        r -= alpha * Ap
        if it % 10 == 0:
            res = float(torch.linalg.norm(r) / bn)
            if res < rtol:
                break
        z = dinv * r
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    res = float(torch.linalg.norm(r) / bn)
    return x, it, res
