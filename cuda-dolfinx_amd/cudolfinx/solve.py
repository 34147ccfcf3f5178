"""Jacobi-preconditioned conjugate gradients on an assembled device matrix.

Stands in for the PETSc KSP (cg + jacobi, or bjacobi) the reference's
examples hand the assembled Mat/Vec to (python/examples/poisson.py:116-125)
when petsc4py is absent.  Given a CUDAMatrix, the products and the diagonal
run on the engine's own CSR kernels (mf_spmv / mf_matrix_diagonal) on the
matrix's device arrays; a torch sparse CSR tensor also works.  With block =
b > 1 the preconditioner inverts the b x b diagonal blocks (consecutive rows;
for a DG space numbered cell by cell, b = dofs per cell gives each cell's
own block), batched on the device.  The iteration runs in FP64.
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


class _DeviceView:
    """A float64 device array at a raw address, for torch.as_tensor."""

    def __init__(self, addr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (addr, False),
                                         "version": 3}


def _block_inverse(M, bsz: int, device):
    """Inverses (n / bsz, bsz, bsz) of M's bsz x bsz diagonal blocks.  The
    positions of the blocks' entries in the CSR value array are found once
    per matrix (host) and reused; the values are gathered on the device."""
    import torch
    native = hasattr(M, "mult")
    key = getattr(M, "_block_pos", None)
    if key is None or key[0] != bsz:
        if native:
            rp, cols, _ = M.csr()
        else:
            rp, cols = M.crow_indices().cpu().numpy(), M.col_indices().cpu().numpy()
        n = rp.size - 1
        if n % bsz:
            raise ValueError(f"{n} rows are not a multiple of the block size {bsz}")
        rows = np.repeat(np.arange(n), np.diff(rp))
        inb = (rows // bsz) == (cols // bsz)
        pos = np.full((n // bsz, bsz, bsz), -1, dtype=np.int64)
        pos[rows[inb] // bsz, rows[inb] % bsz, cols[inb] % bsz] = np.flatnonzero(inb)
        key = (bsz, torch.from_numpy(pos).to(device))
        if native:
            M._block_pos = key
    pos = key[1]
    if native:
        vals = torch.as_tensor(_DeviceView(M.device_arrays()[2], M.nnz), device=device)
    else:
        vals = M.values()
    blocks = torch.where(pos >= 0, vals[pos.clamp(min=0)], torch.zeros((), dtype=vals.dtype,
                                                                        device=device))
    inv, info = torch.linalg.inv_ex(blocks)
    bad = info != 0
    if bool(bad.any()):
        # a singular block (e.g. a row with no entries): plain Jacobi on it
        dg = torch.diagonal(blocks[bad], dim1=1, dim2=2)
        inv[bad] = torch.diag_embed(torch.where(dg != 0, 1.0 / dg, torch.zeros_like(dg)))
    return inv


def pcg(M, b, rtol: float = 1e-10, maxiter: int = 5000, block: int = 1):
    """Solve M x = b with Jacobi-PCG (block = 1) or block-Jacobi-PCG (block =
    b: the b x b diagonal blocks inverted).  M: a CUDAMatrix (native CSR SpMV
    and diagonal, mf_spmv / mf_matrix_diagonal, on torch's current stream) or
    a torch sparse CSR tensor; b: a float64 torch vector.

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
    if block > 1:
        binv = _block_inverse(M, int(block), b.device)

        def precond(r):
            return torch.bmm(binv, r.view(-1, block, 1)).view(-1)
    else:
        dinv = torch.where(d != 0, 1.0 / d, torch.zeros_like(d))

        def precond(r):
            return dinv * r
    x = torch.zeros_like(b)
    r = b.clone()
    z = precond(r)
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
        z = precond(r)
        rz_new = torch.dot(r, z)
        p = z + (rz_new / rz) * p
        rz = rz_new
    res = float(torch.linalg.norm(r) / bn)
    return x, it, res
