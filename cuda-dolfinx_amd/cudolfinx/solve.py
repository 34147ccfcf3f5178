"""Jacobi-preconditioned conjugate gradients on an assembled device matrix.

Stands in for the PETSc KSP (cg + jacobi) the reference's examples hand the
assembled Mat/Vec to (python/examples/poisson.py:116-125) when petsc4py is
absent.  The CSR arrays are wrapped as a torch sparse CSR tensor (device
memory shared with the matrix for the values); the iteration runs in FP64.
"""
from __future__ import annotations

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
    """Solve M x = b (torch sparse CSR, torch vector) with Jacobi-PCG.

    Returns (x, iterations, final relative residual).
    """
    import torch
    d = torch.zeros(M.shape[0], dtype=torch.float64, device=b.device)
    crow, col, val = M.crow_indices(), M.col_indices(), M.values()
    rows = torch.repeat_interleave(torch.arange(M.shape[0], device=b.device), crow[1:] - crow[:-1])
    diag_mask = rows == col
    d.index_put_((rows[diag_mask],), val[diag_mask], accumulate=True)
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
        Ap = torch.mv(M, p)
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
