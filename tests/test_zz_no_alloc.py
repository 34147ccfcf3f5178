"""The allocation-free lifting entry point (mf_apply_lifting_noalloc), called
through the C ABI, against mf_apply_lifting on the same inputs and against
the oracle: two calls in a row, the second on a smaller problem (a stale
entity count or list in the shared scratch array would show), on the
cell-staged lifting kernel and on the per-entity one."""
import ctypes as C

import numpy as np
import pytest

from oracle import cpu
from mifem import problems
from mifem.fem import Function, dirichletbc, locate_dofs_topological

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)


def _rel(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


def _lift(P, lift_kernel, seed):
    import torch
    import cudolfinx as cufem
    from cudolfinx import _lib
    lib = _lib.lib()
    rng = np.random.default_rng(seed)
    g = Function(P.V)
    g.x[:] = rng.uniform(-1, 1, P.V.size)
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)
    a = cufem.form(P.a, options={"lift_kernel": lift_kernel})
    L = cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    dbc = asm.pack_bcs([bc])
    b0 = asm.assemble_vector(L)
    a.to_device()                        # coefficients of the bilinear form, packed on the device
    asm.pack_coefficients(a)
    asm.synchronize()
    base = torch.from_numpy(b0.to_host()).cuda()
    x0 = torch.from_numpy(rng.uniform(-1, 1, P.V.size)).cuda()
    forms = (C.c_void_p * 1)(a.handle.value)
    bcs = (C.c_void_p * 1)(dbc._get_cpp_bcs(P.V).handle.value)
    x0s = (C.c_void_p * 1)(x0.data_ptr())
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for name in ("mf_apply_lifting", "mf_apply_lifting_noalloc"):
        b = base.clone()
        _lib.check(getattr(lib, name)(forms, 1, C.c_void_p(b.data_ptr()), P.V.size, bcs, x0s, -0.7,
                                      stream))
        torch.cuda.synchronize()
        out[name] = b.cpu().numpy()
    ob = cpu.OracleAssembler(P.L).assemble_vector()
    cpu.OracleAssembler(P.a).apply_lifting(ob, [bc], x0=x0.cpu().numpy(), scale=-0.7)
    return out["mf_apply_lifting_noalloc"], out["mf_apply_lifting"], ob


@pytest.mark.parametrize("lift_kernel", [0, 1])
def test_noalloc_lifting_matches_lifting_two_calls(lift_kernel):
    for P, seed in ((problems.elasticity_p2(4), 1), (problems.poisson_p1(3), 2)):
        _say(f"{P.name} ({P.mesh.num_cells} cells) lift_kernel={lift_kernel}: "
             f"mf_apply_lifting_noalloc vs mf_apply_lifting")
        new, ref, ob = _lift(P, lift_kernel, seed)
        d, e = _rel(new, ref), _rel(new, ob)
        _say(f"  rel diff to mf_apply_lifting {d:.2e}, to the oracle {e:.2e}")
        assert d <= 1e-13 and e <= 1e-12
