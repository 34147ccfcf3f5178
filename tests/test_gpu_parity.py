"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar: CSR row pointers and columns bit-exact; values within 1e-12 relative
(Frobenius norm of the difference over the norm of the oracle matrix, l2
for vectors) on the FP64 path.  Per-entry relative error is not used: P1
rows hold structurally present entries that are numerically ~1e-17.
"""
import sys

import numpy as np
import pytest

from oracle import cpu
from mifem import forms, problems
from mifem.fem import Constant, Function, dirichletbc, functionspace, locate_dofs_topological
from mifem.mesh import create_unit_cube

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def _rel(x, ref):
    return np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300)


def _oracle(P, bcs=None, x0=None, scale=1.0, diagonal=1.0):
    bcs = P.bcs if bcs is None else bcs
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    rp, cols, _, _ = Oa.sparsity()
    vals = Oa.assemble_matrix(rp, cols, bcs, diagonal=diagonal)
    b = OL.assemble_vector()
    Oa.apply_lifting(b, bcs, x0=x0, scale=scale)
    cpu.set_bc(b, bcs, P.V, x0=x0, scale=scale)
    return rp, cols, vals, b


def _gpu(P, options=None, bcs=None, x0=None, scale=1.0, diagonal=1.0, loptions=None):
    import cudolfinx as cufem
    bcs = P.bcs if bcs is None else bcs
    a = cufem.form(P.a, options=options)
    L = cufem.form(P.L, options=loptions)
    asm = cufem.CUDAAssembler()
    dbc = asm.pack_bcs(bcs)
    A = asm.assemble_matrix(a, bcs=dbc, diagonal=diagonal)
    A.assemble()
    b = asm.assemble_vector(L)
    xv = None
    if x0 is not None:
        xv = asm.create_vector(L)
        xv.data.upload(x0)
    asm.apply_lifting(b, [a], [dbc], x0=[xv] if xv is not None else None, scale=scale)
    asm.set_bc(b, dbc, P.V, x0=xv, scale=scale)
    asm.synchronize()
    rp, cols, vals = A.csr()
    return rp, cols, vals, b.to_host(), (a, L, asm, A, dbc)


def _compare(P, gpu, ref):
    rp, cols, vals, b = gpu[:4]
    orp, ocols, ovals, ob = ref
    assert np.array_equal(rp, orp), "row pointers differ"
    assert np.array_equal(cols, ocols), "column indices differ"
    ea, eb = _rel(vals, ovals), _rel(b, ob)
    _say(f"{P.name}: nnz={cols.size} relerr A={ea:.3e} b={eb:.3e}")
    assert ea <= TOL and eb <= TOL, (ea, eb)


@pytest.mark.parametrize("kernel", [0, 1])
def test_poisson_p1(kernel):
    _say(f"poisson_p1 n=8 matrix_kernel={kernel}")
    P = problems.poisson_p1(8)
    _compare(P, _gpu(P, {"matrix_kernel": kernel}), _oracle(P))


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_elasticity_p2(kernel):
    _say(f"elasticity_p2 n=5 matrix_kernel={kernel}")
    P = problems.elasticity_p2(5)
    _compare(P, _gpu(P, {"matrix_kernel": kernel}), _oracle(P))


def test_elasticity_p2_reference_tensor_owner():
    _say("elasticity P2 n=6, owner kernel on the full reference-tensor path (G = |det J| K^T R K)")
    P = problems.elasticity_p2(6)
    P.a = forms.elasticity(P.V, P.extras["mu"], P.extras["lambda"], reference_tensor=True)
    _compare(P, _gpu(P, {"matrix_kernel": 2}), _oracle(P))


@pytest.mark.parametrize("form", ["elasticity", "stiffness_p2"])
@pytest.mark.parametrize("subset", ["mixed", "prefix"])
def test_owner_cell_integral_on_a_subset_of_cells(form, subset):
    """A cell integral over part of the mesh (a subdomain) on the owner path.
    The record pass stages a workgroup's records in LDS when its 256 cells are
    consecutive and stores them one by one otherwise: "mixed" keeps cells
    0..699 (two staged blocks, then a block that is not) and a random third of
    the rest (a partial last block, not staged); "prefix" keeps cells 0..599
    (a partial last block that is staged)."""
    import dataclasses
    _say(f"owner-computes cell integral on a subset of cells: {form}, {subset}")
    if form == "elasticity":
        P = problems.elasticity_p2(6)
    else:
        P = problems.poisson_p1(6)
        P.V = functionspace(P.mesh, ("P", 2))
        kap = Function(functionspace(P.mesh, ("DG", 0)), "kappa")
        kap.x[:] = np.random.default_rng(3).lognormal(size=kap.x.size)
        f = Function(P.V, "f")
        f.interpolate(lambda x: x[0] * x[1] + x[2])
        P.a, P.L = forms.stiffness(P.V, kap), forms.source(P.V, f)
        P.bcs = [dirichletbc(0.0, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)]
        P.name = "stiffness_p2"
    nc = P.mesh.num_cells
    if subset == "mixed":
        rng = np.random.default_rng(5)
        keep = np.concatenate([np.arange(0, 700), rng.choice(np.arange(700, nc), size=(nc - 700) // 3,
                                                             replace=False)])
    else:
        keep = np.arange(0, 600)
    keep = np.sort(keep).astype(np.int32)
    I = P.a.integrals[0]
    P.a = dataclasses.replace(P.a, integrals=[dataclasses.replace(I, entities=keep)])
    _compare(P, _gpu(P, {"matrix_kernel": 2}), _oracle(P))


def test_elasticity_p2_medium():
    _say("elasticity_p2 n=16 owner-computes")
    P = problems.elasticity_p2(16)
    _compare(P, _gpu(P, {"matrix_kernel": 2}), _oracle(P))


def test_poisson_p1_full_size_c1():
    _say("poisson_p1 n=32 (196,608 cells)")
    P = problems.poisson_p1(32)
    _compare(P, _gpu(P), _oracle(P))


@pytest.mark.timeout(900)
def test_elasticity_p2_full_size_c2():
    """Vector P2 elasticity on the 64^3 x 6 mesh (1,572,864 cells, 6,440,067
    dofs, 547,739,145 nonzeros) with the owner-kernel shape the benchmark runs
    (4608-value chunks, 384 pairs, 256 threads): CSR bit-exact and A, b within
    1e-12 of the oracle (its coloured multi-core loops, which equal the serial
    ones to round-off)."""
    import os
    import time
    _say("elasticity_p2 n=64 (1,572,864 cells), owner shape 4608/384/256")
    t0 = time.perf_counter()
    P = problems.elasticity_p2(64)
    shape = {"matrix_kernel": 2, "chunk_values": 4608, "owner_pairs": 384, "owner_block": 256,
             "owner_lds": 73728}
    gpu = _gpu(P, shape)
    _say(f"  GPU A+b done ({time.perf_counter() - t0:.1f} s)")
    gpu = gpu[:4]                       # release the device objects
    threads = max(1, min(32, len(os.sched_getaffinity(0))))
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    orp, ocols, _, _ = Oa.sparsity()
    assert np.array_equal(gpu[0], orp), "row pointers differ"
    assert np.array_equal(gpu[1], ocols), "column indices differ"
    Oa.colouring()
    OL.colouring()
    ovals = Oa.assemble_matrix_mt(orp, ocols, P.bcs, threads=threads)
    ob = OL.assemble_vector_mt(threads=threads)
    Oa.apply_lifting(ob, P.bcs, threads=threads)
    cpu.set_bc(ob, P.bcs, P.V)
    _say(f"  oracle A+b on {threads} threads done ({time.perf_counter() - t0:.1f} s)")
    assert gpu[1].size == 547739145 and gpu[0].size == 6440068
    _compare(P, gpu, (orp, ocols, ovals, ob))


@pytest.mark.parametrize("k", [2, 3])
def test_pk_poisson_dg0_kappa(k):
    _say(f"P{k} Poisson with DG0 kappa, jittered mesh")
    mesh = create_unit_cube(4, jitter=0.2, seed=7)
    V = functionspace(mesh, ("P", k))
    D = functionspace(mesh, ("DG", 0))
    kap = Function(D)
    kap.set_cellwise(np.exp(np.sin(np.arange(mesh.num_cells))))
    f = Function(V)
    f.interpolate(lambda x: 1 + x[0] * x[1])
    a, L = forms.stiffness(V, kap), forms.source(V, f)
    bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
    P = problems.Problem(f"p{k}_kappa", mesh, V, a, L, [bc])
    for kernel in (1, 2):
        _compare(P, _gpu(P, {"matrix_kernel": kernel}), _oracle(P))


def test_lifting_with_values_x0_scale_and_diagonal():
    _say("lifting g != 0, x0, scale, diagonal")
    P = problems.poisson_p1(6)
    rng = np.random.default_rng(11)
    g = Function(P.V)
    g.x[:] = rng.random(P.V.size)
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)
    x0 = rng.random(P.V.size)
    ref = _oracle(P, [bc], x0=x0, scale=-0.6, diagonal=2.5)
    for kernel in (1, 2):
        _compare(P, _gpu(P, {"matrix_kernel": kernel}, [bc], x0=x0, scale=-0.6, diagonal=2.5), ref)


def test_reassembly_after_coefficient_change():
    import cudolfinx as cufem
    _say("reassembly after coefficient change")
    P = problems.elasticity_p2(4)
    gpu = _gpu(P)
    a, L, asm, A, dbc = gpu[4]
    P.extras["mu"].x[:] *= 1.7
    P.extras["mu"].version += 1
    asm.assemble_matrix(a, A, bcs=dbc)
    asm.synchronize()
    rp, cols, vals = A.csr()
    orp, ocols, ovals, _ = _oracle(P)
    assert _rel(vals, ovals) <= TOL


def test_accumulate_without_zero():
    _say("zero=False accumulates")
    P = problems.elasticity_p2(3)
    for kernel in (1, 2):
        gpu = _gpu(P, {"matrix_kernel": kernel})
        a, L, asm, A, dbc = gpu[4]
        asm.assemble_matrix(a, A, bcs=dbc, zero=False)
        asm.synchronize()
        _, _, vals = A.csr()
        orp, ocols, ovals, _ = _oracle(P)
        diag = np.zeros(ovals.size, dtype=bool)
        for r in P.bcs[0].dofs:
            s, e = orp[r], orp[r + 1]
            diag[s + np.nonzero(ocols[s:e] == r)[0]] = True
        exp = 2 * ovals
        exp[diag] = 1.0
        assert _rel(vals, exp) <= TOL


def test_helmholtz_q3_generic_kernel():
    _say("helmholtz Q3 n=2 jittered (per-entity kernel)")
    P = problems.helmholtz_q3(2)
    _compare(P, _gpu(P, {"matrix_kernel": 1}), _oracle(P))


def test_repeatability():
    _say("two assemblies agree")
    P = problems.elasticity_p2(6)
    g1 = _gpu(P)
    a, L, asm, A, dbc = g1[4]
    asm.assemble_matrix(a, A, bcs=dbc)
    asm.synchronize()
    _, _, v2 = A.csr()
    assert _rel(v2, g1[2]) <= 1e-14


@pytest.mark.parametrize("vk,lk", [(1, 1), (2, 0), (2, 1)])
def test_vector_and_lifting_kernels(vk, lk):
    _say(f"vector_kernel={vk} lift_kernel={lk} elasticity + poisson")
    P = problems.elasticity_p2(5)
    rng = np.random.default_rng(5)
    g = Function(P.V)
    g.x[:] = rng.random(P.V.size)
    left = P.mesh.locate_boundary_facets(lambda x: np.isclose(x[0], 0.0))
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, left), P.V)
    ref = _oracle(P, [bc], scale=0.8)
    _compare(P, _gpu(P, {"lift_kernel": lk}, [bc], scale=0.8, loptions={"vector_kernel": vk}), ref)
    Q = problems.poisson_p1(7)
    _compare(Q, _gpu(Q, {"lift_kernel": lk}, loptions={"vector_kernel": vk}), _oracle(Q))


@pytest.mark.parametrize("prec,tol", [(64, 1e-12), (32, 1e-6), (16, 2e-2)])
def test_helmholtz_q3_matrix_core(prec, tol):
    _say(f"helmholtz Q3 n=4 jittered, matrix-core owner kernel precision={prec}")
    P = problems.helmholtz_q3(4)
    g = _gpu(P, {"matrix_kernel": 2, "precision": prec})
    ref = _oracle(P)
    rp, cols, vals, b = g[:4]
    assert np.array_equal(rp, ref[0]) and np.array_equal(cols, ref[1])
    ea, eb = _rel(vals, ref[2]), _rel(b, ref[3])
    _say(f"  precision {prec}: relerr A={ea:.3e} b={eb:.3e}")
    assert ea <= tol and eb <= 1e-12, (ea, eb)


@pytest.mark.parametrize("k,kernel", [(1, 1), (1, 3), (2, 1), (2, 3)])
def test_sipg_plaplace_interior_facets(k, kernel):
    _say(f"SIPG p-Laplacian DG{k} n=3 jittered: cells + Robin facets + interior facets, "
         f"kernel={'cell-owner' if kernel == 3 else 'per-entity'}")
    P = problems.dg_plaplace(3, degree=k)
    rng = np.random.default_rng(11)
    P.extras["u"].x[:] = rng.uniform(-1, 1, P.V.size)
    _compare(P, _gpu(P, {"matrix_kernel": kernel}, loptions={"vector_kernel": kernel}), _oracle(P))


def test_sipg_cell_owner_with_dirichlet_and_accumulate():
    _say("SIPG DG1 cell-owner kernel: Dirichlet rows/cols, diagonal 2.5, accumulate without zero")
    P = problems.dg_plaplace(3)
    rng = np.random.default_rng(12)
    P.extras["u"].x[:] = rng.uniform(-1, 1, P.V.size)
    dofs = np.arange(0, P.V.size, 7, dtype=np.int32)
    g = Function(P.V)
    g.x[:] = rng.random(P.V.size)
    bc = dirichletbc(g, dofs, P.V)
    ref = _oracle(P, [bc], diagonal=2.5)
    got = _gpu(P, {"matrix_kernel": 3}, [bc], diagonal=2.5, loptions={"vector_kernel": 3})
    _compare(P, got, ref)
    a, L, asm, A, dbc = got[4]
    asm.assemble_matrix(a, A, bcs=dbc, diagonal=2.5, zero=False)
    asm.synchronize()
    _, _, v2 = A.csr()
    exp = 2.0 * ref[2]
    rp, cols = ref[0], ref[1]
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    on = np.isin(rows, dofs) & (rows == cols)
    exp[on] = 2.5
    assert _rel(v2, exp) <= TOL


def test_graph_replay_matches_eager_and_oracle():
    _say("HIP graph: capture SIPG J + F (device-bound state u), 3 replays vs eager and oracle")
    import torch
    import cudolfinx as cufem
    P = problems.dg_plaplace(3)
    n = P.V.size
    rng = np.random.default_rng(21)
    s = torch.cuda.Stream()
    J, F = cufem.form(P.a), cufem.form(P.L)
    asm = cufem.CUDAAssembler(stream=s)
    ud = torch.zeros(n, dtype=torch.float64, device="cuda")
    J.set_coefficient_device(0, ud.data_ptr(), n)
    F.set_coefficient_device(0, ud.data_ptr(), n)
    A, b = asm.create_matrix(J), asm.create_vector(F)

    def seq():
        asm.assemble_matrix(J, A)
        asm.assemble_vector(F, b)

    seq()                                   # builds the schedules outside capture
    asm.synchronize()
    g = cufem.CUDAGraph(s)
    with g:
        seq()
    oa, ol = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    rp, cols, _, _ = oa.sparsity()
    for rep in range(3):
        u = rng.uniform(-1, 1, n)
        ud.copy_(torch.from_numpy(u).cuda())
        torch.cuda.synchronize()
        g.launch()
        asm.synchronize()
        vg, bg = A.csr()[2].copy(), b.to_host().copy()
        seq()
        asm.synchronize()
        ve, be = A.csr()[2], b.to_host()
        assert np.array_equal(vg, ve) and np.array_equal(bg, be), f"replay {rep} differs from eager"
        P.extras["u"].x[:] = u
        oa, ol = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
        ea, eb = _rel(vg, oa.assemble_matrix(rp, cols)), _rel(bg, ol.assemble_vector())
        _say(f"  replay {rep}: relerr A={ea:.3e} b={eb:.3e}")
        assert ea <= TOL and eb <= TOL


@pytest.mark.parametrize("k,kv", [(1, 1), (2, 1), (1, 2)])
def test_mixed_functional_three_integral_types(k, kv):
    _say(f"functional exp(u) kappa dx + u kappa ds + avg(v^2) avg(kappa) dS, P{k}/DG{kv}, n=4")
    import cudolfinx as cufem
    from mifem.functionals import mixed_functional
    mesh = create_unit_cube(4, jitter=0.2, seed=4)
    V, D = functionspace(mesh, ("P", k)), functionspace(mesh, ("DG", kv))
    u, kap, v = Function(V), Function(V), Function(D)
    rng = np.random.default_rng(31)
    u.x[:] = rng.uniform(-1, 1, V.size)
    kap.x[:] = rng.uniform(1, 2, V.size)
    v.x[:] = rng.uniform(-1, 1, D.size)
    M = mixed_functional(u, kap, v)
    asm = cufem.CUDAAssembler()
    Mg = cufem.form(M)
    got = asm.assemble_scalar(Mg)
    again = asm.assemble_scalar(Mg)
    ref = cpu.OracleAssembler(M).assemble_scalar()
    _say(f"  value {got:.15e} oracle {ref:.15e}")
    assert abs(got - ref) <= 1e-13 * abs(ref), (got, ref)
    assert got == again, "repeated assembly differs"
    # a host-side coefficient change reaches the device on the next call
    u.x[:] *= 0.5
    u.version += 1
    ref2 = cpu.OracleAssembler(M).assemble_scalar()
    got2 = asm.assemble_scalar(Mg)
    assert abs(got2 - ref2) <= 1e-13 * abs(ref2), (got2, ref2)


@pytest.mark.parametrize("k", [1, 2])
def test_sipg_energy_functional(k):
    _say(f"SIPG p-Laplacian energy functional DG{k}, n=6 jittered; device output, accumulate")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(6, degree=k)
    u = P.extras["u"]
    rng = np.random.default_rng(41)
    u.x[:] = rng.uniform(-1, 1, P.V.size)
    E_ = sipg_energy(P.V, u)
    ref = cpu.OracleAssembler(E_).assemble_scalar()
    asm = cufem.CUDAAssembler()
    Eg = cufem.form(E_)
    got = asm.assemble_scalar(Eg)
    _say(f"  value {got:.15e} oracle {ref:.15e}")
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    out = torch.full((1,), 1.5, dtype=torch.float64, device="cuda")
    asm.assemble_scalar_device(Eg, out, zero=False)
    asm.synchronize()
    assert abs(out.item() - (1.5 + got)) <= 1e-12 * abs(1.5 + got)
    asm.assemble_scalar_device(Eg, out)
    asm.synchronize()
    assert out.item() == got


def test_graph_newton_step_with_energy():
    _say("HIP graph: capture SIPG J + F + energy E (device-bound u), 3 replays vs eager and oracle")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(3)
    n = P.V.size
    s = torch.cuda.Stream()
    E_ = sipg_energy(P.V, P.extras["u"])
    J, F, Eg = cufem.form(P.a), cufem.form(P.L), cufem.form(E_)
    asm = cufem.CUDAAssembler(stream=s)
    ud = torch.zeros(n, dtype=torch.float64, device="cuda")
    for f in (J, F, Eg):
        f.set_coefficient_device(0, ud.data_ptr(), n)
    A, b = asm.create_matrix(J), asm.create_vector(F)
    ev = torch.zeros(1, dtype=torch.float64, device="cuda")

    def seq():
        asm.assemble_matrix(J, A)
        asm.assemble_vector(F, b)
        asm.assemble_scalar_device(Eg, ev)

    seq()
    asm.synchronize()
    g = cufem.CUDAGraph(s)
    with g:
        seq()
    rng = np.random.default_rng(22)
    for rep in range(3):
        u = rng.uniform(-1, 1, n)
        ud.copy_(torch.from_numpy(u).cuda())
        torch.cuda.synchronize()
        g.launch()
        asm.synchronize()
        eg = ev.item()
        seq()
        asm.synchronize()
        assert ev.item() == eg, f"replay {rep}: energy differs from eager"
        P.extras["u"].x[:] = u
        ref = cpu.OracleAssembler(E_).assemble_scalar()
        _say(f"  replay {rep}: E={eg:.15e} oracle {ref:.15e}")
        assert abs(eg - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("cell,fam0,k0,fam1,k1,bs", [
    ("tetrahedron", "P", 1, "P", 2, 1), ("tetrahedron", "P", 2, "P", 3, 3),
    ("tetrahedron", "DG", 1, "P", 1, 1), ("tetrahedron", "P", 3, "DG", 1, 1),
    ("hexahedron", "P", 1, "P", 3, 1)])
def test_same_map_interpolation(cell, fam0, k0, fam1, k1, bs):
    _say(f"interpolation {fam0}{k0} -> {fam1}{k1} bs={bs} on {cell} n=6 (random source)")
    import cudolfinx as cufem
    from cudolfinx.coefficient import interpolation_operator
    mesh = create_unit_cube(6, cell, jitter=0.15, seed=6)
    shape = (bs,) if bs > 1 else None
    V0 = functionspace(mesh, (fam0, k0), shape=shape)
    V1 = functionspace(mesh, (fam1, k1), shape=shape)
    f0, f1 = Function(V0), Function(V1)
    rng = np.random.default_rng(17)
    f0.x[:] = rng.uniform(-1, 1, V0.size)
    c0, c1 = cufem.Coefficient(f0), cufem.Coefficient(f1)
    im = interpolation_operator(V1.element, V0.element)
    for rep in range(2):
        c1.interpolate(c0)
        got = c1.values()
        ref = cpu.interpolate_same_map(V1, V0, im, f0.x)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        _say(f"  call {rep}: max rel diff {err:.2e}")
        assert err <= 1e-14, err
        f0.x[:] = rng.uniform(-1, 1, V0.size)
        c0.copy_host_values_to_device()


def test_many_instances_block_separate():
    _say("batch of 6 Poisson instances (P1-P3, n=2..4) merged per degree, owner-computes")
    from mifem import batch
    import cudolfinx as cufem
    B = batch.many_poisson(6, seed=3, nmin=2, nmax=4)
    for g in B.groups:
        class _P:
            pass
        P = _P()
        P.name, P.a, P.L, P.bcs, P.V = f"batch P{g.degree}", g.a, g.L, g.bcs, g.V
        ref = _oracle(P)
        got = _gpu(P)
        _compare(P, got, ref)
        for j in range(len(g.instances)):
            r1, c1, v1 = g.block_csr(got[0], got[1], got[2], j)
            r2, c2, v2 = g.block_csr(ref[0], ref[1], ref[2], j)
            assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
            assert _rel(v1, v2) <= TOL


def test_many_instances_one_launch():
    _say("batch of 12 Poisson instances (P1-P3, n=2..5), all groups in one launch")
    from mifem import batch
    import cudolfinx as cufem
    B = batch.many_poisson(12, seed=3, nmin=2, nmax=5)
    asm = cufem.CUDAAssembler()
    forms_ = [cufem.form(g.a) for g in B.groups]
    bcs = [asm.pack_bcs(g.bcs) for g in B.groups]
    mats = [asm.create_matrix(a) for a in forms_]
    sep = [asm.create_matrix(a) for a in forms_]
    fused = cufem.CUDABatch(forms_)
    assert "mf_owner_batch" in fused.source()
    for rep in range(2):
        fused.assemble(asm, mats, bcs, diagonal=1.0 + rep)
        for a, S, bc in zip(forms_, sep, bcs):
            asm.assemble_matrix(a, S, bcs=bc, diagonal=1.0 + rep)
        asm.synchronize()
        for g, A, S in zip(B.groups, mats, sep):
            class _P:
                pass
            P = _P()
            P.name, P.a, P.L, P.bcs, P.V = f"one-launch P{g.degree} call {rep}", g.a, g.L, g.bcs, g.V
            rp, cols, vals = A.csr()
            srp, scols, svals = S.csr()
            assert np.array_equal(rp, srp) and np.array_equal(cols, scols)
            orp, ocols, _, _ = cpu.OracleAssembler(P.a).sparsity()
            ovals = cpu.OracleAssembler(P.a).assemble_matrix(orp, ocols, P.bcs, diagonal=1.0 + rep)
            assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
            es, eo = _rel(vals, svals), _rel(vals, ovals)
            _say(f"{P.name}: nnz={cols.size} vs separate {es:.2e}, vs oracle {eo:.2e}")
            assert es <= 1e-14 and eo <= TOL, (es, eo)
    # accumulate (zero=False): a second pass doubles every entry off the constrained rows/columns
    A0 = [m.csr()[2].copy() for m in mats]
    fused.assemble(asm, mats, bcs, diagonal=2.0, zero=False)
    asm.synchronize()
    for g, A, v0 in zip(B.groups, mats, A0):
        rp, cols, vals = A.csr()
        bc = np.zeros(g.V.size, dtype=bool)
        for d in g.bcs:
            bc[d.dofs] = True
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
        free = ~(bc[rows] | bc[cols])
        assert np.allclose(vals[free], 2 * v0[free], rtol=1e-14, atol=0)
        assert np.array_equal(vals[~free], v0[~free])


def test_batch_autotuner_shapes_agree():
    _say("batch autotuner on 12 Poisson instances (P1-P3): every member shape it times and the "
         "winner give the matrices of separate assembly")
    from mifem import batch
    import cudolfinx as cufem
    from cudolfinx.autotune import autotune_batch
    B = batch.many_poisson(12, seed=7, nmin=2, nmax=5)
    asm = cufem.CUDAAssembler()
    forms_ = [cufem.form(g.a) for g in B.groups]
    bcs = [asm.pack_bcs(g.bcs) for g in B.groups]
    mats = [asm.create_matrix(a) for a in forms_]
    sep = [asm.create_matrix(a) for a in forms_]
    for a, S, bc in zip(forms_, sep, bcs):
        asm.assemble_matrix(a, S, bcs=bc)
    cands = [{"chunk_values": 1024, "owner_pairs": 256}, {"chunk_values": 2048, "owner_pairs": 384}]
    opts, fused, members, timings = autotune_batch([g.a for g in B.groups], asm, mats, bcs,
                                                   candidates=cands, reps=2)
    # the first choice plus one trial per member and other candidate
    assert len(timings) == 1 + len(B.groups) * (len(cands) - 1)
    assert len(opts) == len(members) == len(B.groups)
    for o in opts:
        assert {k: o[k] for k in ("chunk_values", "owner_pairs")} in cands
    fused.assemble(asm, mats, bcs)
    asm.synchronize()
    for g, A, S in zip(B.groups, mats, sep):
        rp, cols, vals = A.csr()
        srp, scols, svals = S.csr()
        assert np.array_equal(rp, srp) and np.array_equal(cols, scols)
        e = _rel(vals, svals)
        _say(f"P{g.degree}: nnz={cols.size}, tuned {opts[B.groups.index(g)]}, vs separate {e:.2e}")
        assert e <= 1e-14
    # plain lists of DirichletBC: each call packs its own collections
    fused.assemble(asm, mats, [list(g.bcs) for g in B.groups])
    asm.synchronize()
    for A, S in zip(mats, sep):
        assert _rel(A.csr()[2], S.csr()[2]) <= 1e-14


def test_many_right_hand_sides_one_launch():
    _say("right-hand sides of 12 Poisson instances (P1-P3) + elasticity (body force and traction): "
         "cell integrals in one launch, set_bc in one launch")
    from mifem import batch
    import cudolfinx as cufem
    B = batch.many_poisson(12, seed=5, nmin=2, nmax=5)
    E = problems.elasticity_p2(3)                      # a cell and an exterior-facet integral
    asm = cufem.CUDAAssembler()
    Ls = [cufem.form(g.L) for g in B.groups] + [cufem.form(E.L)]
    spaces = [g.V for g in B.groups] + [E.V]
    bclists = [g.bcs for g in B.groups] + [E.bcs]
    bcs = [asm.pack_bcs(b) for b in bclists]
    vb = cufem.CUDABatch(Ls)
    assert "mf_vown_batch" in vb.source()
    vecs = [asm.create_vector(L) for L in Ls]
    for rep in range(2):
        vb.assemble_vectors(asm, vecs)
        # packed collections, then plain lists of DirichletBC (packed per call)
        vb.set_bc(asm, vecs, bcs if rep == 0 else bclists, scale=1.0 + rep)
        asm.synchronize()
        for L, V, bl, v in zip(Ls, spaces, bclists, vecs):
            ref = asm.assemble_vector(L)
            asm.set_bc(ref, asm.pack_bcs(bl), V, scale=1.0 + rep)
            asm.synchronize()
            e = _rel(v.to_host(), ref.to_host())
            _say(f"  call {rep}, {V.element.family}{V.element.degree} bs {V.bs}: vs separate {e:.2e}")
            assert e <= 1e-14
    # accumulate: the cell and facet parts added on top
    before = [v.to_host().copy() for v in vecs]
    vb.assemble_vectors(asm, vecs, zero=False)
    asm.synchronize()
    for L, v, v0 in zip(Ls, vecs, before):
        ref = asm.assemble_vector(L)
        asm.synchronize()
        assert _rel(v.to_host(), v0 + ref.to_host()) <= 1e-14
    # against the oracle (without the Dirichlet values)
    vb.assemble_vectors(asm, vecs)
    asm.synchronize()
    mforms = [g.L for g in B.groups] + [E.L]
    for mf_, v in zip(mforms, vecs):
        assert _rel(v.to_host(), cpu.OracleAssembler(mf_).assemble_vector()) <= TOL


def test_one_launch_mixed_block_sizes():
    _say("one launch for vector P2 elasticity (bs 3) and scalar P1 Poisson (bs 1)")
    import cudolfinx as cufem
    Ps = [problems.elasticity_p2(3), problems.poisson_p1(5)]
    asm = cufem.CUDAAssembler()
    fs = [cufem.form(P.a) for P in Ps]
    bcs = [asm.pack_bcs(P.bcs) for P in Ps]
    mats = [asm.create_matrix(a) for a in fs]
    cufem.CUDABatch(fs).assemble(asm, mats, bcs)
    asm.synchronize()
    for P, A in zip(Ps, mats):
        rp, cols, vals = A.csr()
        orp, ocols, ovals, _ = _oracle(P)
        assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
        ea = _rel(vals, ovals)
        _say(f"  {P.name}: nnz={cols.size} relerr A={ea:.3e}")
        assert ea <= TOL, ea


def test_one_launch_errors():
    _say("one-launch batch: unsupported members and mismatched arguments are rejected")
    import cudolfinx as cufem
    H = problems.helmholtz_q3(2)
    with pytest.raises(NotImplementedError):
        cufem.CUDABatch([cufem.form(H.a)])
    P = problems.poisson_p1(3)
    a = cufem.form(P.a)
    asm = cufem.CUDAAssembler()
    fused = cufem.CUDABatch([a])
    with pytest.raises(ValueError):
        fused.assemble(asm, [asm.create_matrix(a), asm.create_matrix(a)])
    Q = problems.poisson_p1(4)
    with pytest.raises(ValueError):           # matrix of a different mesh
        fused.assemble(asm, [asm.create_matrix(cufem.form(Q.a))])


def test_reference_tensor_fp32():
    _say("scalar reference-tensor owner kernel in FP32 (precision=32): P1-P3 Poisson groups, "
         "alone and in one launch")
    from mifem import batch
    import cudolfinx as cufem
    B = batch.many_poisson(6, seed=3, nmin=2, nmax=4)
    asm = cufem.CUDAAssembler()
    forms32, mats, bcs, refs = [], [], [], []
    for g in B.groups:
        a = cufem.form(g.a, options={"matrix_kernel": 2, "precision": 32})
        dbc = asm.pack_bcs(g.bcs)
        A = asm.assemble_matrix(a, bcs=dbc)
        asm.synchronize()
        Oa = cpu.OracleAssembler(g.a)
        orp, ocols, _, _ = Oa.sparsity()
        ovals = Oa.assemble_matrix(orp, ocols, g.bcs)
        rp, cols, vals = A.csr()
        assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
        ea = _rel(vals, ovals)
        _say(f"  P{g.degree} alone: nnz={cols.size} relerr A={ea:.3e}")
        assert 0 < ea <= 1e-6, ea               # FP32 arithmetic, not the FP64 kernel
        forms32.append(a)
        mats.append(asm.create_matrix(a))
        bcs.append(dbc)
        refs.append(ovals)
    fused = cufem.CUDABatch(forms32)
    fused.assemble(asm, mats, bcs)
    asm.synchronize()
    for g, A, ovals in zip(B.groups, mats, refs):
        ea = _rel(A.csr()[2], ovals)
        _say(f"  P{g.degree} in one launch: relerr A={ea:.3e}")
        assert 0 < ea <= 1e-6, ea


def test_precision_errors():
    _say("precision 32 outside the scalar reference-tensor owner path is rejected")
    import cudolfinx as cufem
    P = problems.elasticity_p2(2)
    asm = cufem.CUDAAssembler()
    a = cufem.form(P.a, options={"precision": 32})
    with pytest.raises(NotImplementedError):
        asm.assemble_matrix(a, bcs=P.bcs)
    Q = problems.poisson_p1(3)
    q = cufem.form(Q.a, options={"precision": 32, "matrix_kernel": 1})
    with pytest.raises(NotImplementedError):
        asm.assemble_matrix(q, bcs=Q.bcs)


@pytest.mark.parametrize("which", ["poisson_p1", "elasticity_p2", "helmholtz_q3"])
def test_native_spmv_diagonal_and_cg(which):
    _say(f"native CSR SpMV / diagonal / Jacobi-CG on the assembled {which} matrix")
    import torch
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from cudolfinx.solve import pcg
    P = {"poisson_p1": lambda: problems.poisson_p1(6), "elasticity_p2": lambda: problems.elasticity_p2(3),
         "helmholtz_q3": lambda: problems.helmholtz_q3(3)}[which]()
    rp, cols, vals, b, (a, L, asm, A, dbc) = _gpu(P)
    S = sp.csr_matrix((vals, cols, rp), shape=(A.nrows, A.ncols))
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, A.ncols)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    A.mult(xd, yd)
    y1 = yd.cpu().numpy()
    assert _rel(y1, S @ x) <= 1e-14
    A.mult(xd, yd)
    assert np.array_equal(yd.cpu().numpy(), y1)              # repeatable, fixed order
    dd = torch.empty(A.nrows, dtype=torch.float64, device="cuda")
    A.diagonal(dd)
    assert np.array_equal(dd.cpu().numpy(), S.diagonal())
    with pytest.raises(ValueError):
        A.mult(xd, xd)
    bd = torch.from_numpy(b).cuda()
    u, it, res = pcg(A, bd, rtol=1e-12, maxiter=5000)
    ref = spla.spsolve(S.tocsc(), b)
    err = _rel(u.cpu().numpy(), ref)
    _say(f"  nnz={S.nnz}: SpMV ok, CG {it} its, residual {res:.1e}, error vs direct solve {err:.1e}")
    assert res < 1e-12 and err <= 1e-8, (res, err)


def test_block_jacobi_cg_on_dg_jacobian():
    _say("block-Jacobi CG (each cell's 4 x 4 block inverted) on the SIPG DG1 Jacobian, n=4")
    import torch
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import cudolfinx as cufem
    from cudolfinx.solve import pcg
    P = problems.dg_plaplace(4)
    a = cufem.form(P.a)
    asm = cufem.CUDAAssembler()
    A = asm.assemble_matrix(a)
    asm.synchronize()
    rp, cols, vals = A.csr()
    S = sp.csr_matrix((vals, cols, rp), shape=(A.nrows, A.ncols))
    b = np.random.default_rng(5).uniform(-1, 1, A.nrows)
    ref = spla.spsolve(S.tocsc(), b)
    its = {}
    for blk in (1, 4):
        u, it, res = pcg(A, torch.from_numpy(b).cuda(), rtol=1e-12, maxiter=20000, block=blk)
        err = _rel(u.cpu().numpy(), ref)
        its[blk] = it
        _say(f"  block {blk}: {it} its, residual {res:.1e}, error vs direct solve {err:.1e}")
        assert res < 1e-12 and err <= 1e-8, (blk, res, err)
    assert its[4] < its[1]


def test_coefficient_updates_repack():
    _say("host coefficient changed between assemblies: repacked, matrices follow the new values")
    import cudolfinx as cufem
    from mifem import batch
    mesh, V, kap = batch.poisson_instance(4, 2)
    D = functionspace(mesh, ("DG", 0))
    kf = Function(D)
    kf.set_cellwise(kap)
    bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
    form = forms.stiffness(V, kf)
    a = cufem.form(form)
    asm = cufem.CUDAAssembler()
    A = asm.create_matrix(a)
    Oa = cpu.OracleAssembler(form)
    orp, ocols, _, _ = Oa.sparsity()
    rng = np.random.default_rng(3)
    for rep in range(3):
        # 0: first pack; 1: new host values, packed again; 2: unchanged, the pack is skipped
        if rep == 1:
            kf.x[:] = rng.uniform(0.5, 2.0, kf.x.size)
            kf.version += 1
        asm.assemble_matrix(a, A, bcs=[bc])
        asm.synchronize()
        ref = cpu.OracleAssembler(form).assemble_matrix(orp, ocols, [bc])
        err = _rel(A.csr()[2], ref)
        _say(f"  assembly {rep}: relerr A={err:.2e}")
        assert err <= TOL, (rep, err)


# ------------------------------------------------------------ block / restricted assembly
def _selector(V, keep):
    """Scalar-dof selection matrix of the kept nodes (in order), bs-unrolled."""
    import scipy.sparse as sp
    keep = np.arange(V.num_nodes) if keep is None else np.asarray(keep)
    rows = np.arange(keep.size * V.bs)
    cols = (V.bs * np.repeat(keep, V.bs) + np.tile(np.arange(V.bs), keep.size))
    return sp.csr_matrix((np.ones(rows.size), (rows, cols)), shape=(rows.size, V.size))


@pytest.mark.parametrize("restricted", [False, True])
def test_block_vector(restricted):
    _say(f"block vector: P1 + P2 + vector P1 sources, restricted={restricted}")
    import cudolfinx as cufem
    mesh = create_unit_cube(3, jitter=0.2, seed=8)
    V1, V2 = functionspace(mesh, ("P", 1)), functionspace(mesh, ("P", 2))
    W = functionspace(mesh, ("P", 1), shape=(3,))
    f1 = Function(V1)
    f1.interpolate(lambda x: x[0] ** 2 + x[1] ** 3)
    Ls = [forms.source(V1, f1), forms.source(V2, Constant([2.0])),
          forms.source(W, Constant([0.0, 1.0, -1.0]))]
    keep = None
    if restricted:
        rng = np.random.default_rng(4)
        keep = [rng.permutation(V.num_nodes)[: V.num_nodes // 2] for V in (V1, V2, W)]
    asm = cufem.CUDAAssembler()
    bf = cufem.form(Ls, restriction=keep)
    assert isinstance(bf, cufem.BlockCUDAForm)
    vec = asm.create_vector_block(bf)
    asm.assemble_vector_block(bf, vec)
    got = vec.to_host()
    ref = np.concatenate([_selector(L.function_spaces[0], None if keep is None else keep[i])
                          @ cpu.OracleAssembler(L).assemble_vector() for i, L in enumerate(Ls)])
    assert got.size == ref.size == bf.local_size
    err = _rel(got, ref)
    _say(f"  size {got.size}, offsets {bf.offsets}, relerr {err:.2e}")
    assert err <= TOL


def test_block_matrix_with_restriction_and_bcs():
    _say("block matrix [[K(P2), M(P2,P1)], [M(P1,P2), M(P1,P1)]], P1 restricted to x < 0.55, "
         "Dirichlet on the P2 boundary, diagonal 2.0")
    import scipy.sparse as sp
    import cudolfinx as cufem
    mesh = create_unit_cube(3, jitter=0.2, seed=9)
    V, Q = functionspace(mesh, ("P", 2)), functionspace(mesh, ("P", 1))
    D = functionspace(mesh, ("DG", 0))
    kap = Function(D)
    kap.set_cellwise(np.linspace(1.0, 3.0, mesh.num_cells))
    a = [[forms.stiffness(V, kap), forms.mass(V, Q)], [forms.mass(Q, V, kap), forms.mass(Q)]]
    keepQ = np.flatnonzero(Q.node_coordinates[:, 0] < 0.55)
    bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
    bcs = [[[bc], [bc]], [[bc], []]]
    asm = cufem.CUDAAssembler()
    bf = cufem.form(a, restriction=([None, keepQ], [None, keepQ]))
    A = asm.create_matrix_block(bf)
    for i in range(2):
        for j in range(2):
            asm.assemble_matrix(bf.forms[i][j], A, bcs=bcs[i][j], diagonal=2.0, zero=(i == 0 and j == 0))
    asm.synchronize()
    rp, cols, vals = A.csr()
    # oracle: each block with its bcs, restricted and stacked
    blocks, pats = [], []
    S = [_selector(V, None), _selector(Q, keepQ)]
    for i in range(2):
        brow, prow = [], []
        for j in range(2):
            oa = cpu.OracleAssembler(a[i][j])
            orp, ocols, _, _ = oa.sparsity()
            ov = oa.assemble_matrix(orp, ocols, bcs[i][j], diagonal=2.0)
            n0, n1 = a[i][j].function_spaces[0].size, a[i][j].function_spaces[1].size
            M = sp.csr_matrix((ov, ocols, orp), shape=(n0, n1))
            P = sp.csr_matrix((np.ones(ocols.size), ocols, orp), shape=(n0, n1))
            brow.append(S[i] @ M @ S[j].T)
            prow.append(S[i] @ P @ S[j].T)
        blocks.append(brow)
        pats.append(prow)
    ref = sp.bmat(blocks).tocsr()
    pat = sp.bmat(pats).tocsr()
    pat.sort_indices()
    assert np.array_equal(rp, pat.indptr) and np.array_equal(cols, pat.indices), "block CSR differs"
    got = sp.csr_matrix((vals, cols, rp), shape=ref.shape)
    from scipy.sparse.linalg import norm as spnorm
    err = spnorm(got - ref) / spnorm(ref)
    _say(f"  {ref.shape[0]} rows, nnz {cols.size}, relerr {err:.2e}")
    assert err <= TOL
    # the block-assembler entry point gives the same values
    A2 = asm.assemble_matrix_block(bf, bcs=bcs)
    asm.synchronize()
    v2 = A2.csr()[2]
    exp = sp.csr_matrix((vals, cols, rp), shape=ref.shape)
    # assemble_matrix_block uses diagonal 1.0 (DOLFINx default): compare off the diagonal rows
    bcrows = np.unique(bc.dofs)
    mask = np.ones(ref.shape[0], dtype=bool)
    mask[bcrows] = False
    d = (sp.csr_matrix((v2, cols, rp), shape=ref.shape) - exp)[mask]
    assert abs(d).max() <= 1e-13 * abs(exp).max()


# ------------------------------------------------------------ error behaviour and edge cases
def test_error_paths_raise_like_the_reference():
    """Argument checks raise TypeError / ValueError as python/cudolfinx/assemble.py
    does; C-ABI status codes surface as ValueError (bad shape / argument) or
    NotImplementedError (unsupported); the library keeps working afterwards."""
    _say("error behaviour: wrong types, length and shape mismatches, unsupported calls")
    import cudolfinx as cufem
    from cudolfinx.la import CUDAVector
    P = problems.poisson_p1(2)
    a, L = cufem.form(P.a), cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    with pytest.raises(TypeError):
        asm.assemble_matrix(P.a)
    with pytest.raises(TypeError):
        asm.assemble_vector("L")
    with pytest.raises(TypeError):
        asm.create_matrix_block(a)
    b = asm.create_vector(L)
    with pytest.raises(ValueError):
        asm.apply_lifting(b, [a], [])
    with pytest.raises(ValueError):
        asm.assemble_scalar(L)
    with pytest.raises(ValueError):
        asm.assemble_vector(L, CUDAVector(a._ctx, P.V.size + 1))
    P3 = problems.poisson_p1(3)
    with pytest.raises(ValueError):
        asm.assemble_matrix(cufem.form(P3.a), asm.create_matrix(a))
    mesh2 = create_unit_cube(2)
    f_other = Function(functionspace(mesh2, ("P", 1)))
    with pytest.raises(ValueError):
        cufem.Coefficient(Function(P.V)).interpolate(cufem.Coefficient(f_other))
    # lifting with a block member form is rejected
    blk = cufem.form([[forms.stiffness(P.V)]])
    with pytest.raises(NotImplementedError):
        asm.apply_lifting(b, [blk.forms[0][0]], [P.bcs])
    # still healthy: a normal assembly matches the oracle
    _compare(P, _gpu(P), _oracle(P))


def test_edge_cases_one_cube_and_empty_integrals():
    _say("edge cases: one-cube meshes (P1, P2, vector P2), an empty facet integral, an empty form")
    import dataclasses
    import cudolfinx as cufem
    for k in (1, 2):
        mesh = create_unit_cube(1)
        V = functionspace(mesh, ("P", k))
        f = Function(V)
        f.interpolate(lambda x: 1.0 + x[0])
        bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets[:3]), V)

        class _P:
            pass
        P = _P()
        P.name, P.V, P.bcs = f"one cube P{k}", V, [bc]
        P.a, P.L = forms.stiffness(V), forms.source(V, f)
        _compare(P, _gpu(P), _oracle(P))
    Pe = problems.elasticity_p2(1)
    _compare(Pe, _gpu(Pe), _oracle(Pe))
    # a facet integral over no facets adds nothing
    P = problems.poisson_p1(3)
    empty = forms.facet_source(P.V, Constant([1.0]), np.zeros((0, 2), dtype=np.int32))
    P.L = forms.linear_sum(P.L, empty)
    assert P.L.integrals[-1].num_entities == 0
    _compare(P, _gpu(P), _oracle(P))
    # a bilinear form whose only integral has no cells: empty pattern, nothing assembled
    a0 = forms.stiffness(P.V)
    a0.integrals[0] = dataclasses.replace(a0.integrals[0], entities=np.zeros(0, dtype=np.int32))
    ag = cufem.form(a0)
    asm = cufem.CUDAAssembler()
    A = asm.assemble_matrix(ag, bcs=[])
    asm.synchronize()
    rp, cols, vals = A.csr()
    assert A.nnz == 0 and cols.size == 0 and np.all(rp == 0)


# ------------------------------------------------------------ fused Newton system (J + F + E)
@pytest.mark.parametrize("k", [1, 2])
def test_fused_system_matches_separate_forms_and_oracle(k):
    _say(f"fused J + F + E launch, SIPG DG{k} n=4: vs separate launches, oracle, repeat, accumulate")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(4, degree=k)
    u = P.extras["u"]
    rng = np.random.default_rng(51)
    u.x[:] = rng.uniform(-1, 1, P.V.size)
    u.version += 1
    E_ = sipg_energy(P.V, u)
    J, F, E = cufem.form(P.a), cufem.form(P.L), cufem.form(E_)
    asm = cufem.CUDAAssembler()
    A1 = asm.assemble_matrix(J)
    b1 = asm.assemble_vector(F)
    e1 = asm.assemble_scalar(E)
    system = cufem.CUDASystem(J, F, E)
    A2, b2 = asm.create_matrix(J), asm.create_vector(F)
    ev = torch.zeros(1, dtype=torch.float64, device="cuda")
    system.assemble(asm, A2, b2, ev)
    asm.synchronize()
    v1, v2 = A1.csr()[2], A2.csr()[2]
    eA, eb = _rel(v2, v1), _rel(b2.to_host(), b1.to_host())
    e2 = ev.item()
    _say(f"  fused vs separate: A {eA:.2e}, b {eb:.2e}, E {abs(e2 - e1) / abs(e1):.2e}")
    assert eA <= 1e-14 and eb <= 1e-14 and abs(e2 - e1) <= 1e-13 * abs(e1)
    oa = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = oa.sparsity()
    assert _rel(v2, oa.assemble_matrix(rp, cols)) <= TOL
    assert _rel(b2.to_host(), cpu.OracleAssembler(P.L).assemble_vector()) <= TOL
    ref = cpu.OracleAssembler(E_).assemble_scalar()
    assert abs(e2 - ref) <= 1e-12 * abs(ref)
    # repeat: identical; accumulate: doubled A, b and E
    system.assemble(asm, A2, b2, ev)
    asm.synchronize()
    assert ev.item() == e2 and np.array_equal(A2.csr()[2], v2)
    system.assemble(asm, A2, b2, ev, zero=False, accumulate_value=True)
    asm.synchronize()
    assert _rel(A2.csr()[2], 2 * v2) <= 1e-15 and abs(ev.item() - 2 * e2) <= 1e-14 * abs(e2)


@pytest.mark.parametrize("opts", [{"owner_balance": 0},
                                  {"owner_balance": 1, "chunk_values": 2048, "owner_pairs": 128}])
def test_owner_unit_mappings(opts):
    _say(f"elasticity_p2 n=6 owner kernel {opts}")
    P = problems.elasticity_p2(6)
    _compare(P, _gpu(P, {"matrix_kernel": 2, **opts}), _oracle(P))


@pytest.mark.parametrize("stage,prefetch", [(0, 0), (0, 1), (1, 0)])
def test_dg_kernel_variants_match_default(stage, prefetch):
    """Direct stores (dg_stage 0) and no item prefetch (dg_prefetch 0) give the
    values of the default DG kernels (staged, prefetching): the standalone
    matrix kernel bit for bit, the fused system's A and b bit for bit and its
    energy to round-off (the workgroup count, which orders the energy's
    partial sums, depends on the kernel's resources)."""
    _say(f"SIPG DG1 n=4: dg_stage={stage} dg_prefetch={prefetch} against the defaults")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(4)
    u = P.extras["u"]
    u.x[:] = np.random.default_rng(73).uniform(-1, 1, P.V.size)
    E_ = sipg_energy(P.V, u)
    out = []
    for o in (None, {"dg_stage": stage, "dg_prefetch": prefetch}):
        J, F, E = cufem.form(P.a, options=o), cufem.form(P.L, options=o), cufem.form(E_, options=o)
        asm = cufem.CUDAAssembler()
        A1 = asm.assemble_matrix(J)
        system = cufem.CUDASystem(J, F, E)
        A2, b2 = asm.create_matrix(J), asm.create_vector(F)
        ev = torch.zeros(1, dtype=torch.float64, device="cuda")
        system.assemble(asm, A2, b2, ev)
        asm.synchronize()
        out.append((A1.csr()[2].copy(), A2.csr()[2].copy(), b2.to_host().copy(), ev.item()))
    (a0, s0, b0, e0), (a1, s1, b1, e1) = out
    assert np.array_equal(a0, a1) and np.array_equal(s0, s1) and np.array_equal(b0, b1)
    assert abs(e1 - e0) <= 1e-14 * abs(e0)
    oa = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = oa.sparsity()
    assert _rel(a1, oa.assemble_matrix(rp, cols)) <= TOL


def test_fused_system_dirichlet_and_graph_replay():
    _say("fused J + F + E with Dirichlet rows (diagonal 2.5) and HIP graph replays, DG1 n=3")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(3)
    n = P.V.size
    rng = np.random.default_rng(61)
    dofs = np.arange(0, n, 5, dtype=np.int32)
    bc = dirichletbc(0.0, dofs, P.V)
    s = torch.cuda.Stream()
    E_ = sipg_energy(P.V, P.extras["u"])
    J, F, E = cufem.form(P.a), cufem.form(P.L), cufem.form(E_)
    asm = cufem.CUDAAssembler(stream=s)
    ud = torch.zeros(n, dtype=torch.float64, device="cuda")
    for f in (J, F, E):
        f.set_coefficient_device(0, ud.data_ptr(), n)
    system = cufem.CUDASystem(J, F, E)
    A, b = asm.create_matrix(J), asm.create_vector(F)
    dbc = asm.pack_bcs([bc])
    ev = torch.zeros(1, dtype=torch.float64, device="cuda")

    def step():
        system.assemble(asm, A, b, ev, bcs=dbc, diagonal=2.5)

    step()
    asm.synchronize()
    g = cufem.CUDAGraph(s)
    with g:
        step()
    oa = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = oa.sparsity()
    for rep in range(3):
        u = rng.uniform(-1, 1, n)
        ud.copy_(torch.from_numpy(u).cuda())
        torch.cuda.synchronize()
        g.launch()
        asm.synchronize()
        vg, bg, eg = A.csr()[2].copy(), b.to_host().copy(), ev.item()
        step()
        asm.synchronize()
        assert np.array_equal(vg, A.csr()[2]) and np.array_equal(bg, b.to_host()) and eg == ev.item()
        P.extras["u"].x[:] = u
        oa = cpu.OracleAssembler(P.a)
        ea = _rel(vg, oa.assemble_matrix(rp, cols, [bc], diagonal=2.5))
        eb = _rel(bg, cpu.OracleAssembler(P.L).assemble_vector())
        ref = cpu.OracleAssembler(E_).assemble_scalar()
        _say(f"  replay {rep}: A {ea:.2e} b {eb:.2e} E {abs(eg - ref) / abs(ref):.2e}")
        assert ea <= TOL and eb <= TOL and abs(eg - ref) <= 1e-12 * abs(ref)


# ------------------------------------------------------------ the reference test suite, on tetrahedra
# (python/test/test_cuda_assembly.py: test_cuda_assembly, test_reassembly, test_lifting,
# test_diagonal_assembly; its 2D unit square replaced by a jittered unit cube)
def _ref_suite(k=1, domain="tet"):
    """The reference suite's forms on a jittered tetrahedral cube, or on its
    own domain (test_cuda_assembly.py:40-46): the 19 x 27 unit square of
    triangles."""
    from mifem.mesh import create_unit_square
    from mifem.nonlinear import reference_test_forms
    mesh = (create_unit_cube(4, jitter=0.15, seed=3) if domain == "tet"
            else create_unit_square(19, 27))
    return reference_test_forms(mesh, k)


@pytest.mark.parametrize("domain", ["tet", "tri"])
@pytest.mark.parametrize("k", [1, 2])
def test_reference_suite_cuda_assembly(k, domain):
    _say(f"reference suite ({domain}): functional, 3 residuals, 3 Jacobians with bcs, P{k}/DG{k}")
    import cudolfinx as cufem
    F = _ref_suite(k, domain)
    asm = cufem.CUDAAssembler()
    for form in F["scalar"]:
        v1 = cpu.OracleAssembler(form).assemble_scalar()
        v2 = asm.assemble_scalar(cufem.form(form))
        assert abs(v2 - v1) <= 1e-12 * abs(v1), (v1, v2)
    for form in F["vector"]:
        ref = cpu.OracleAssembler(form).assemble_vector()
        got = asm.assemble_vector(cufem.form(form)).to_host()
        assert _rel(got, ref) <= TOL, form.description
    for form in F["matrix"]:
        oa = cpu.OracleAssembler(form)
        rp, cols, _, _ = oa.sparsity()
        ref = oa.assemble_matrix(rp, cols, F["bcs"])
        A = asm.assemble_matrix(cufem.form(form), bcs=F["bcs"])
        A.assemble()
        grp, gcols, vals = A.mat.getValuesCSR()
        assert np.array_equal(grp, rp) and np.array_equal(gcols, cols), form.description
        err = _rel(vals, ref)
        _say(f"  {form.description}: nnz {cols.size} relerr {err:.2e}")
        assert err <= TOL


@pytest.mark.parametrize("domain", ["tet", "tri"])
def test_reference_suite_reassembly(domain):
    _say(f"reference suite: reassembly after kappa.interpolate(x^d + y^d), d = 2, 3 ({domain})")
    import cudolfinx as cufem
    F = _ref_suite(1, domain)
    coeff = F["coeff"]
    cform = cufem.form(F["vector"][0])
    asm = cufem.CUDAAssembler()
    vec = asm.assemble_vector(cform)
    assert _rel(vec.to_host(), cpu.OracleAssembler(F["vector"][0]).assemble_vector()) <= TOL
    for d in (2, 3):
        coeff.interpolate(lambda x: x[0] ** d + x[1] ** d)
        cform.to_device()
        asm.assemble_vector(cform, vec)
        assert _rel(vec.to_host(), cpu.OracleAssembler(F["vector"][0]).assemble_vector()) <= TOL


@pytest.mark.parametrize("domain", ["tet", "tri"])
def test_reference_suite_lifting(domain):
    _say(f"reference suite: set_bc then apply_lifting on the exterior residual / Jacobian ({domain})")
    import cudolfinx as cufem
    F = _ref_suite(1, domain)
    L, a = F["vector"][1], F["matrix"][1]
    asm = cufem.CUDAAssembler()
    vec = asm.assemble_vector(cufem.form(L))
    ref = cpu.OracleAssembler(L).assemble_vector()
    assert _rel(vec.to_host(), ref) <= TOL
    cpu.set_bc(ref, F["bcs"], L.function_spaces[0])
    asm.set_bc(vec, F["bcs"], L.function_spaces[0])
    assert _rel(vec.to_host(), ref) <= TOL
    cpu.OracleAssembler(a).apply_lifting(ref, F["bcs"])
    asm.apply_lifting(vec, [cufem.form(a)], [F["bcs"]])
    assert _rel(vec.to_host(), ref) <= TOL


@pytest.mark.parametrize("domain", ["tet", "tri"])
def test_reference_suite_diagonal_assembly(domain):
    _say(f"reference suite: diagonal part of each Jacobian assembled as a vector ({domain})")
    import cudolfinx as cufem
    from mifem.nonlinear import diagonal
    F = _ref_suite(1, domain)
    asm = cufem.CUDAAssembler()
    for a in F["matrix"]:
        d = diagonal(a)
        got = asm.assemble_vector(cufem.form(d)).to_host()
        ref = cpu.OracleAssembler(d).assemble_vector()
        assert _rel(got, ref) <= TOL, a.description


def test_native_autotuner_then_parity():
    _say("C-ABI owner-kernel autotuner (mf_autotune_matrix) on elasticity P2 n=4, then parity")
    import cudolfinx as cufem
    from cudolfinx.autotune import autotune_matrix_native
    P = problems.elasticity_p2(4)
    a, L = cufem.form(P.a), cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    A = asm.create_matrix(a)
    desc, ms = autotune_matrix_native(a, asm, A, P.bcs, reps=2)
    _say(f"  chosen {desc} ({ms:.3f} ms)")
    assert "chunk_values=" in desc and ms > 0
    asm.assemble_matrix(a, A, bcs=P.bcs)
    b = asm.assemble_vector(L)
    dbc = asm.pack_bcs(P.bcs)
    asm.apply_lifting(b, [a], [dbc])
    asm.set_bc(b, dbc, P.V)
    asm.synchronize()
    rp, cols, vals = A.csr()
    _compare(P, (rp, cols, vals, b.to_host()), _oracle(P))


def test_graph_lifting_recorded_while_g_is_zero():
    """apply_lifting captured while every constrained value is 0 is still in
    the graph: after CUDADirichletBC.update sets g != 0, each replay equals
    the eager sequence and the oracle."""
    _say("HIP graph: lifting captured with g = 0, replayed after bc update to g != 0")
    import torch
    import cudolfinx as cufem
    P = problems.poisson_p1(6)
    s = torch.cuda.Stream()
    a, L = cufem.form(P.a), cufem.form(P.L)
    asm = cufem.CUDAAssembler(stream=s)
    g = Function(P.V)
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)
    dbc = asm.pack_bcs([bc])
    A = asm.assemble_matrix(a, bcs=dbc)
    b = asm.create_vector(L)

    def seq():
        asm.assemble_vector(L, b)
        asm.apply_lifting(b, [a], [dbc])
        asm.set_bc(b, dbc, P.V)

    seq()                                   # g = 0: builds the lifting lists, adds nothing
    asm.synchronize()
    graph = cufem.CUDAGraph(s)
    with graph:
        seq()
    rng = np.random.default_rng(5)
    for rep in range(3):
        g.x[:] = rng.uniform(-1, 1, P.V.size)
        dbc.update([bc])
        graph.launch()
        asm.synchronize()
        bg = b.to_host().copy()
        seq()
        asm.synchronize()
        # the vector and lifting kernels add with atomics: equal up to summation order
        assert _rel(bg, b.to_host()) <= 1e-14, f"replay {rep} differs from eager"
        _, _, _, ob = _oracle(P, [bc])
        eb = _rel(bg, ob)
        _say(f"  replay {rep}: relerr b={eb:.3e}")
        assert eb <= TOL


def test_lifting_with_plain_bc_lists_keeps_device_memory_flat():
    """apply_lifting called with a plain list (a new device condition per call,
    as the reference's own test does) reuses one lifting entity list per
    constrained set: device memory does not grow with the number of calls."""
    _say("apply_lifting with plain bc lists, 30 calls: device memory flat")
    import gc
    import torch
    import cudolfinx as cufem
    P = problems.poisson_p1(24)
    g = Function(P.V)
    g.x[:] = 1.0
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)
    a, L = cufem.form(P.a), cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    b = asm.assemble_vector(L)
    for _ in range(2):
        asm.apply_lifting(b, [a], [[bc]])
    asm.synchronize()
    gc.collect()
    free0 = torch.cuda.mem_get_info()[0]
    for _ in range(30):
        asm.apply_lifting(b, [a], [[bc]])
    asm.synchronize()
    gc.collect()
    grown = free0 - torch.cuda.mem_get_info()[0]
    _say(f"  device memory grown by {grown / 1e6:.2f} MB over 30 calls "
         f"(a list per call would be {30 * 4 * P.mesh.num_cells / 1e6:.1f} MB)")
    assert grown < 4 * P.mesh.num_cells * 10


def test_host_coefficient_writes_reach_the_next_assembly():
    """Writing a coefficient's values in place (f.x[:] = ..., f.x *= ...)
    without touching anything else is seen by the next assembly."""
    _say("coefficient written through f.x only, then reassembled")
    import cudolfinx as cufem
    P = problems.elasticity_p2(3)
    a = cufem.form(P.a)
    asm = cufem.CUDAAssembler()
    dbc = asm.pack_bcs(P.bcs)
    A = asm.assemble_matrix(a, bcs=dbc)
    asm.synchronize()
    mu = P.extras["mu"]
    for step, write in enumerate((lambda: mu.x.__setitem__(slice(None), mu.x * 1.5),
                                  lambda: mu.x.__imul__(0.25))):
        write()
        asm.assemble_matrix(a, A, bcs=dbc)
        asm.synchronize()
        _, _, vals = A.csr()
        orp, ocols, ovals, _ = _oracle(P)
        ea = _rel(vals, ovals)
        _say(f"  write {step}: relerr A={ea:.3e}")
        assert ea <= TOL


def test_autotune_on_one_matrix_keeps_other_matrices_correct():
    """Two matrices of one form: tuning (recompiling) the form on one leaves
    the schedule of the other stale; its next assembly rebuilds it and still
    matches the oracle."""
    _say("autotune with a second matrix of the same form, then reassemble that matrix")
    import cudolfinx as cufem
    from cudolfinx.autotune import autotune_matrix_native
    P = problems.elasticity_p2(4)
    a = cufem.form(P.a, options={"chunk_values": 6144, "owner_pairs": 512, "owner_lds": 98304})
    asm = cufem.CUDAAssembler()
    dbc = asm.pack_bcs(P.bcs)
    A1, A2 = asm.create_matrix(a), asm.create_matrix(a)
    asm.assemble_matrix(a, A1, bcs=dbc)
    asm.assemble_matrix(a, A2, bcs=dbc)
    asm.synchronize()
    desc, ms = autotune_matrix_native(a, asm, A1, P.bcs, reps=1)
    _say(f"  tuned on A1: {desc}")
    asm.assemble_matrix(a, A2, bcs=dbc)
    asm.synchronize()
    orp, ocols, ovals, _ = _oracle(P)
    rp, cols, vals = A2.csr()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert _rel(vals, ovals) <= TOL


def _oracle_csr(P):
    import scipy.sparse as sp
    rp, cols, vals, b = _oracle(P)
    n = P.V.size
    return sp.csr_matrix((vals, cols, rp), shape=(n, n)), b


def test_adopted_csr_scipy_matches_oracle():
    """CUDAMatrix(ctx, scipy CSR) with the oracle's pattern (bs-unrolled block
    structure: owner-computes path): after the handoff the caller's matrix
    holds the oracle's values."""
    _say("adopted scipy CSR (oracle pattern, vector P2 elasticity n=3)")
    import cudolfinx as cufem
    from cudolfinx.context import get_cuda_context
    P = problems.elasticity_p2(3)
    ref, _ = _oracle_csr(P)
    mine = ref.copy()
    mine.data[:] = -7.0
    a = cufem.form(P.a, options={"matrix_kernel": 2})
    asm = cufem.CUDAAssembler()
    M = cufem.CUDAMatrix(get_cuda_context(), mine)
    asm.assemble_matrix(a, M, bcs=P.bcs)
    M.assemble()
    assert M.mat is mine
    ea = _rel(mine.data, ref.data)
    _say(f"  relerr A={ea:.3e}")
    assert ea <= TOL


def test_adopted_csr_torch_superset_borrowed_values():
    """A torch sparse CSR on the GPU whose pattern is the oracle's plus extra
    scalar entries (no longer bs-unrolled blocks: per-entity path): values are
    assembled in place in the tensor's own memory; pattern entries equal the
    oracle, extra entries are zero."""
    _say("adopted torch CSR (superset pattern, borrowed device values)")
    import scipy.sparse as sp
    import torch
    import cudolfinx as cufem
    from cudolfinx.context import get_cuda_context
    P = problems.elasticity_p2(3)
    ref, _ = _oracle_csr(P)
    n = P.V.size
    rng = np.random.default_rng(9)
    refpat = sp.csr_matrix((np.ones(ref.nnz), ref.indices, ref.indptr), shape=(n, n))
    extra = sp.csr_matrix((np.ones(40), (rng.integers(0, n, 40), rng.integers(0, n, 40))), shape=(n, n))
    pat = (refpat + extra).tocsr()          # all entries positive: nothing is pruned
    pat.sort_indices()
    assert pat.nnz > ref.nnz
    t = torch.sparse_csr_tensor(torch.from_numpy(pat.indptr.astype(np.int64)),
                                torch.from_numpy(pat.indices.astype(np.int64)),
                                torch.full((pat.nnz,), 5.0, dtype=torch.float64), size=(n, n)).cuda()
    vptr = t.values().data_ptr()
    a = cufem.form(P.a)
    asm = cufem.CUDAAssembler()
    M = cufem.CUDAMatrix(get_cuda_context(), t)
    asm.assemble_matrix(a, M, bcs=P.bcs)
    M.assemble()
    assert t.values().data_ptr() == vptr
    vals = t.values().cpu().numpy()                  # in the pattern's CSR order
    rows = np.repeat(np.arange(n), np.diff(pat.indptr))
    cols = pat.indices
    mask = np.asarray(refpat[rows, cols]).ravel() != 0
    expect = np.asarray(ref[rows, cols]).ravel()
    ea = _rel(vals[mask], expect[mask])
    _say(f"  relerr A={ea:.3e}, {int((~mask).sum())} extra entries, max |extra| "
         f"{np.abs(vals[~mask]).max():.1e}")
    assert ea <= TOL and np.all(vals[~mask] == 0.0)


def test_adopted_csr_missing_entry_is_rejected():
    """A pattern lacking an entry the form touches is rejected with
    MF_ENOTFOUND at the first assembly; a wrong shape with ValueError."""
    _say("adopted CSR with a missing entry / wrong shape")
    import scipy.sparse as sp
    import cudolfinx as cufem
    from cudolfinx._lib import MifemError
    from cudolfinx.context import get_cuda_context
    P = problems.poisson_p1(3)
    ref, _ = _oracle_csr(P)
    ones = sp.csr_matrix((np.ones(ref.nnz), ref.indices, ref.indptr), shape=ref.shape)
    lil = ones.tolil()
    lil[5, 5] = 0.0
    holed = lil.tocsr()
    holed.eliminate_zeros()
    assert holed.nnz == ref.nnz - 1
    a = cufem.form(P.a)
    asm = cufem.CUDAAssembler()
    M = cufem.CUDAMatrix(get_cuda_context(), holed)
    with pytest.raises(MifemError, match="missing"):
        asm.assemble_matrix(a, M)
    small = sp.identity(4, format="csr")
    with pytest.raises(ValueError):
        asm.assemble_matrix(a, cufem.CUDAMatrix(get_cuda_context(), small))


@pytest.mark.parametrize("which", ["elasticity_p2", "dg_plaplace"])
def test_forms_from_ffcx_laid_out_source(which, tmp_path):
    """Forms built from an FFCx-laid-out C file (cudolfinx.jit extraction,
    integrals matched by the file's form_integrals order) assemble to the
    oracle's A and b on the per-entity kernels, which run the extracted
    tabulate_tensor bodies."""
    _say(f"{which}: tabulate_tensor bodies extracted from an FFCx-laid-out .c file")
    import cudolfinx as cufem
    from ffcx_layout import body_of, ffcx_c_file, integral_id
    P = getattr(problems, which)(3)
    text = ffcx_c_file({"a": [(integral_id(I.name), body_of(I.source)) for I in P.a.integrals],
                        "L": [(integral_id(I.name), body_of(I.source)) for I in P.L.integrals]})
    path = tmp_path / "form_module.c"
    path.write_text(text)
    from cudolfinx import jit
    wa, oa = jit.get_wrapped_tabulate_tensors(str(path), form_name="a")
    a = cufem.CUDAForm(P.a, options={"matrix_kernel": 1, "lift_kernel": 1},
                       ffcx_source=_OneForm(path, "a"))
    L = cufem.CUDAForm(P.L, options={"vector_kernel": 1}, ffcx_source=_OneForm(path, "L"))
    assert a.source().count("tabulate_tensor_integral_") >= len(P.a.integrals)
    asm = cufem.CUDAAssembler()
    dbc = asm.pack_bcs(P.bcs)
    A = asm.assemble_matrix(a, bcs=dbc)
    b = asm.assemble_vector(L)
    asm.apply_lifting(b, [a], [dbc])
    asm.set_bc(b, dbc, P.V)
    asm.synchronize()
    rp, cols, vals = A.csr()
    _compare(P, (rp, cols, vals, b.to_host()), _oracle(P))


class _OneForm:
    """C text of one form of a multi-form FFCx file: the file with only that
    form's order array (what a single-form FFCx module holds)."""

    def __init__(self, path, name):
        import re
        text = path.read_text()
        keep = f"form_integrals_form_{name}["
        self.text = "\n".join(l for l in text.splitlines()
                              if not re.match(r"\s*ufcx_integral\* form_integrals_form_", l)
                              or keep in l)

    def __str__(self):
        return self.text


@pytest.mark.parametrize("k", [1, 2])
def test_interior_facets_with_facet_permutations(k):
    """Cells list their vertices in random order, so the two cells of an
    interior facet see it in different orders (all six facet permutations
    occur): the per-entity kernels, the cell-owner kernel and the fused
    J + F + E launch gather each side's permutation (CUDAFormIntegral.cpp:
    1913-1920) and match the oracle, which pairs the sides through physical
    space; the mixed functional's interior term likewise."""
    _say(f"SIPG DG{k} n=3 with shuffled cell vertices (non-zero facet permutations)")
    import torch
    import cudolfinx as cufem
    from mifem.dgforms import sipg_energy
    from mifem.functionals import mixed_functional
    P = problems.dg_plaplace(3, degree=k, shuffle_vertices=7)
    assert len(np.unique(P.mesh.facet_permutations())) == 6
    u = P.extras["u"]
    u.x[:] = np.random.default_rng(61).uniform(-1, 1, P.V.size)
    ref = _oracle(P)
    for kernel in (1, 3):
        _compare(P, _gpu(P, {"matrix_kernel": kernel}, loptions={"vector_kernel": kernel}), ref)
    E_ = sipg_energy(P.V, u)
    J, F, E = cufem.form(P.a), cufem.form(P.L), cufem.form(E_)
    asm = cufem.CUDAAssembler()
    system = cufem.CUDASystem(J, F, E)
    A, b = asm.create_matrix(J), asm.create_vector(F)
    ev = torch.zeros(1, dtype=torch.float64, device="cuda")
    system.assemble(asm, A, b, ev)
    asm.synchronize()
    _compare(P, (*A.csr(), b.to_host()), ref)
    eref = cpu.OracleAssembler(E_).assemble_scalar()
    _say(f"  fused energy rel err {abs(ev.item() - eref) / abs(eref):.2e}")
    assert abs(ev.item() - eref) <= 1e-12 * abs(eref)
    V, D = functionspace(P.mesh, ("P", k)), functionspace(P.mesh, ("DG", k))
    uu, kap, v = Function(V), Function(V), Function(D)
    rng = np.random.default_rng(62)
    uu.x[:] = rng.uniform(-1, 1, V.size)
    kap.x[:] = rng.uniform(1, 2, V.size)
    v.x[:] = rng.uniform(-1, 1, D.size)
    M = mixed_functional(uu, kap, v)
    got = asm.assemble_scalar(cufem.form(M))
    mref = cpu.OracleAssembler(M).assemble_scalar()
    _say(f"  mixed functional rel err {abs(got - mref) / abs(mref):.2e}")
    assert abs(got - mref) <= 1e-12 * abs(mref)
