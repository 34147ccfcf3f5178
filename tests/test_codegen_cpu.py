"""CPU checks of the engine's generated element code and of the C ABI library.

* Every generated UFCx tabulate_tensor body is compiled with gcc and its
  element tensors are compared with the oracle's (independent bases,
  quadrature and integrands).
* The staged description (reference gradients + combine) reproduces the
  same element tensor when evaluated in numpy.
* libmifem.so loads and exports every symbol include/mifem.h declares.
"""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import cpu
from mifem import forms, problems
from mifem.fem import Constant, Function, functionspace
from mifem.mesh import create_unit_cube

ROOT = Path(__file__).resolve().parent.parent


def _compile(src: str, tmp_path: Path, name: str):
    c = tmp_path / f"{name}.c"
    so = tmp_path / f"{name}.so"
    c.write_text("#include <math.h>\n#include <stdint.h>\n" + src)
    subprocess.run(["gcc", "-O1", "-std=c99", "-shared", "-fPIC", str(c), "-o", str(so), "-lm"],
                   check=True)
    return C.CDLL(str(so))


def _call(lib, name, n, w, c, xc, eli=None, nA=None, qperm=None):
    fn = getattr(lib, name)
    A = np.zeros(nA)
    w = np.ascontiguousarray(w if len(w) else [0.0], dtype=np.float64)
    c = np.ascontiguousarray(c if len(c) else [0.0], dtype=np.float64)
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    e = (C.c_int * 2)(*(eli or [0, 0]))
    q = (C.c_uint8 * 2)(*qperm) if qperm is not None else None
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    fn(dp(A), dp(w), dp(c), dp(xc), e, q)
    return A


def _check_form(form, tmp_path, ncells=5, seed=0):
    rng = np.random.default_rng(seed)
    mesh = form.mesh
    perms = mesh.facet_permutations()
    packed = form.pack_coefficients()
    consts = form.constant_values()
    for k, I in enumerate(form.integrals):
        lib = _compile(I.source, tmp_path, f"f{k}_{I.name}")
        V0 = form.function_spaces[0] if form.rank else None
        nd = V0.element.ndofs * V0.bs if V0 else 1
        V1 = form.function_spaces[-1] if form.rank == 2 else V0
        nd1 = V1.element.ndofs * V1.bs if V1 else 1
        ne = I.num_entities
        stride = {"cell": 1, "exterior_facet": 2, "interior_facet": 4}[I.type]
        for e in rng.choice(ne, size=min(ncells, ne), replace=False):
            ent = I.entities[e * stride:(e + 1) * stride]
            c = ent[0]
            xc = mesh.x[mesh.cells[c]]
            w = packed[c]
            eli = [int(ent[1])] if I.type == "exterior_facet" else None
            qperm = None
            n, n1 = nd, nd1
            if I.type == "interior_facet":
                c1 = ent[2]
                xc = np.concatenate([xc, mesh.x[mesh.cells[c1]]])
                off = form.coefficient_offsets
                w = np.concatenate([np.concatenate([packed[c][off[j]:off[j + 1]],
                                                    packed[c1][off[j]:off[j + 1]]])
                                    for j in range(len(off) - 1)]) if len(off) > 1 else w
                eli = [int(ent[1]), int(ent[3])]
                qperm = [int(perms[c, ent[1]]), int(perms[c1, ent[3]])]
                n = 2 * nd if form.rank else 1
                n1 = 2 * nd1
            nA = n * n1 if form.rank == 2 else n
            got = _call(lib, I.name, nd, w, consts, xc, eli, nA, qperm)
            ref = cpu.tabulate(I.spec, w, consts, xc, eli).reshape(-1)
            # tensors that vanish up to round-off (e.g. a jump of a continuous
            # field) are compared on an absolute floor
            scale = max(np.abs(ref).max(), 1e-3)
            assert np.abs(got - ref).max() <= 1e-13 * scale, (I.name, e)


def test_generated_p1_poisson(tmp_path):
    P = problems.poisson_p1(3)
    _check_form(P.a, tmp_path)
    _check_form(P.L, tmp_path)


def test_generated_p2_elasticity(tmp_path):
    P = problems.elasticity_p2(3)
    _check_form(P.a, tmp_path)
    _check_form(P.L, tmp_path)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_generated_pk_poisson_with_dg0(tmp_path, k):
    mesh = create_unit_cube(2, jitter=0.2, seed=5)
    V = functionspace(mesh, ("P", k))
    D = functionspace(mesh, ("DG", 0))
    kap = Function(D)
    kap.set_cellwise(np.linspace(1, 3, mesh.num_cells))
    _check_form(forms.stiffness(V, kap), tmp_path)
    f = Function(V)
    f.interpolate(lambda x: x[0] + x[1] ** 2)
    _check_form(forms.source(V, f), tmp_path)


@pytest.mark.parametrize("k", [1, 3])
def test_generated_hex_helmholtz(tmp_path, k):
    P = problems.helmholtz_q3(2, degree=k)
    _check_form(P.a, tmp_path, ncells=3)
    _check_form(P.L, tmp_path, ncells=3)


@pytest.mark.parametrize("k", [1, 2])
def test_generated_sipg_plaplace(tmp_path, k):
    P = problems.dg_plaplace(2, degree=k)
    rng = np.random.default_rng(1)
    P.extras["u"].x[:] = rng.uniform(-1, 1, P.V.size)
    _check_form(P.a, tmp_path, ncells=4)
    _check_form(P.L, tmp_path, ncells=4)


@pytest.mark.parametrize("k", [1, 2])
def test_generated_interior_facet_bodies_with_facet_permutations(tmp_path, k):
    """Cells listing their vertices in random order (every facet permutation
    occurs): the generated interior-facet bodies, given each side's
    permutation, equal the oracle, which pairs the two sides through physical
    space instead (SIPG J, F and energy; the mixed functional; the reference
    suite's interior residual and Jacobian)."""
    from mifem.dgforms import sipg_energy
    from mifem.functionals import mixed_functional
    from mifem.nonlinear import reference_test_forms
    P = problems.dg_plaplace(2, degree=k, shuffle_vertices=11)
    assert len(np.unique(P.mesh.facet_permutations())) == 6
    rng = np.random.default_rng(3)
    P.extras["u"].x[:] = rng.uniform(-1, 1, P.V.size)
    for f in (P.a, P.L, sipg_energy(P.V, P.extras["u"])):
        _check_form(f, tmp_path, ncells=24)
    mesh = P.mesh
    V, D = functionspace(mesh, ("P", k)), functionspace(mesh, ("DG", k))
    u, kap, v = Function(V), Function(V), Function(D)
    u.x[:] = rng.uniform(-1, 1, V.size)
    kap.x[:] = rng.uniform(1, 2, V.size)
    v.x[:] = rng.uniform(-1, 1, D.size)
    _check_form(mixed_functional(u, kap, v), tmp_path, ncells=24)
    F = reference_test_forms(mesh, k)
    for form in F["vector"] + F["matrix"]:
        _check_form(form, tmp_path, ncells=12)


@pytest.mark.parametrize("kv,kw", [(1, 1), (2, 1), (1, 3)])
def test_generated_mixed_mass(tmp_path, kv, kw):
    mesh = create_unit_cube(2, jitter=0.2, seed=3)
    V, W = functionspace(mesh, ("P", kv)), functionspace(mesh, ("P", kw))
    D = functionspace(mesh, ("DG", 0))
    kap = Function(D)
    kap.set_cellwise(np.linspace(1, 2, mesh.num_cells))
    _check_form(forms.mass(V, W, kap), tmp_path)


@pytest.mark.parametrize("k", [1, 2])
def test_generated_sipg_energy(tmp_path, k):
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(2, degree=k)
    rng = np.random.default_rng(2)
    u = P.extras["u"]
    u.x[:] = rng.uniform(-1, 1, P.V.size)
    _check_form(sipg_energy(P.V, u), tmp_path, ncells=4)


@pytest.mark.parametrize("k,kv", [(1, 1), (2, 1), (1, 2)])
def test_generated_mixed_functional(tmp_path, k, kv):
    from mifem.functionals import mixed_functional
    mesh = create_unit_cube(2, jitter=0.2, seed=4)
    V, D = functionspace(mesh, ("P", k)), functionspace(mesh, ("DG", kv))
    u, kap, v = Function(V), Function(V), Function(D)
    rng = np.random.default_rng(5)
    u.x[:] = rng.uniform(-1, 1, V.size)
    kap.x[:] = rng.uniform(1, 2, V.size)
    v.x[:] = rng.uniform(-1, 1, D.size)
    _check_form(mixed_functional(u, kap, v), tmp_path, ncells=4)


@pytest.mark.parametrize("k", [1, 2])
def test_generated_reference_test_forms(tmp_path, k):
    """The reference test suite's residuals, Jacobians, functional and
    diagonal parts (test_cuda_assembly.py:48-91, 200-209) on tetrahedra."""
    from mifem.nonlinear import diagonal, reference_test_forms
    mesh = create_unit_cube(2, jitter=0.2, seed=7)
    F = reference_test_forms(mesh, k)
    for form in F["vector"] + F["matrix"] + F["scalar"] + [diagonal(a) for a in F["matrix"]]:
        _check_form(form, tmp_path, ncells=3)


@pytest.mark.parametrize("k", [1, 2])
@pytest.mark.parametrize("shuffle", [None, 11])
def test_generated_reference_test_forms_on_triangles(tmp_path, k, shuffle):
    """The same forms on the reference suite's own domain, a unit square of
    triangles; with shuffle, cells list their vertices in random order, so
    interior edges are seen in opposite directions (edge permutation 1)."""
    from mifem.mesh import create_unit_square
    from mifem.nonlinear import diagonal, reference_test_forms
    mesh = create_unit_square(4, 3, shuffle_vertices=shuffle)
    if shuffle is not None:
        assert mesh.facet_permutations().max() == 1
    F = reference_test_forms(mesh, k)
    for form in F["vector"] + F["matrix"] + F["scalar"] + [diagonal(a) for a in F["matrix"]]:
        _check_form(form, tmp_path, ncells=6)


def test_staged_description_matches_oracle():
    P = problems.elasticity_p2(2)
    I = P.a.integrals[0]
    S = I.staged
    packed = P.a.pack_coefficients()
    c = 5
    xc = P.mesh.x[P.mesh.cells[c]]
    J = (xc[1:] - xc[0]).T
    K = np.linalg.inv(J)
    g = np.einsum("qir,rd->qid", S.dphi, K)        # physical gradients
    mu, lam = packed[c]
    nd = S.nd
    A = np.zeros((3 * nd, 3 * nd))
    for i in range(nd):
        for j in range(nd):
            G = np.einsum("q,qa,qb->ab", S.weights * abs(np.linalg.det(J)), g[:, i], g[:, j])
            blk = mu * (np.trace(G) * np.eye(3) + G.T) + lam * G
            A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
    ref = cpu.tabulate(I.spec, packed[c], [], xc)
    np.testing.assert_allclose(A, ref, atol=1e-13 * abs(ref).max())


@pytest.mark.parametrize("k", [1, 2, 3])
def test_reference_tensor_trace_matches_oracle(k):
    """Scalar stiffness on the reference-tensor path: kappa |det J| sum_rs M_rs R_ij[rs]
    with M = K K^T reproduces the oracle's element matrix."""
    mesh = create_unit_cube(2, jitter=0.25, seed=9)
    V = functionspace(mesh, ("P", k))
    D = functionspace(mesh, ("DG", 0))
    kap = Function(D)
    kap.set_cellwise(np.linspace(0.5, 2.0, mesh.num_cells))
    form = forms.stiffness(V, kap)
    I = form.integrals[0]
    S = I.staged
    assert S.trace_only
    RS = S.reference_tensor()
    packed = form.pack_coefficients()
    for c in (0, 7, 31):
        xc = mesh.x[mesh.cells[c]]
        J = (xc[1:] - xc[0]).T
        K = np.linalg.inv(J)
        M = abs(np.linalg.det(J)) * K @ K.T
        m = np.array([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]])
        A = packed[c][0] * RS @ m
        ref = cpu.tabulate(I.spec, packed[c], [], xc)
        np.testing.assert_allclose(A, ref, atol=1e-13 * abs(ref).max())


@pytest.mark.parametrize("k", [1, 2])
def test_full_reference_tensor_elasticity_matches_oracle(k):
    """Elasticity on the full reference-tensor path: G_ij = |det J| K^T R_ij K
    (K = J^-1), combined as the kernels do (mu (delta_ab tr G + G^T) + lam G),
    reproduces the oracle's element matrix."""
    mesh = create_unit_cube(2, jitter=0.25, seed=5)
    V = functionspace(mesh, ("P", k), shape=(3,))
    D = functionspace(mesh, ("DG", 0))
    mu, lam = Function(D), Function(D)
    mu.set_cellwise(np.linspace(1.0, 3.0, mesh.num_cells))
    lam.set_cellwise(np.linspace(5.0, 2.0, mesh.num_cells))
    form = forms.elasticity(V, mu, lam, reference_tensor=True)
    I = form.integrals[0]
    S = I.staged
    assert S.full_tensor and "SK_FULLRT" in S.source()
    RT = S.reference_tensor_full().reshape(S.nd, S.nd, 3, 3)
    packed = form.pack_coefficients()
    nd = S.nd
    for c in (0, 9, 40):
        xc = mesh.x[mesh.cells[c]]
        J = (xc[1:] - xc[0]).T
        K = np.linalg.inv(J)                      # K[r, p] = dX_r / dx_p
        det = abs(np.linalg.det(J))
        m, l = packed[c][0], packed[c][1]
        A = np.zeros((3 * nd, 3 * nd))
        for i in range(nd):
            for j in range(nd):
                G = det * K.T @ RT[i, j] @ K
                blk = m * (np.trace(G) * np.eye(3) + G.T) + l * G
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blk
        ref = cpu.tabulate(I.spec, packed[c], [], xc)
        np.testing.assert_allclose(A, ref, atol=1e-13 * abs(ref).max())


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "mifem.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|const char\*)\s+(mf_\w+)\s*\(", header, re.M))
    lib_path = ROOT / "cuda-dolfinx_amd" / "lib" / "libmifem.so"
    if not lib_path.exists():
        pytest.skip("libmifem.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared and declared <= exported, declared - exported
    from cudolfinx import _lib
    assert declared == set(_lib.SIGNATURES)
    L = _lib.lib()
    assert L.mf_version() >= 100
    # without a device, context creation fails with a status code, not a crash
    h = C.c_void_p()
    rc = L.mf_context_create(0, None, C.byref(h))
    if rc != 0:
        assert L.mf_last_error()


def test_null_arguments_rejected_without_a_device():
    """Entry points validate their arguments before any device work: null or
    empty arguments return MF_EINVAL with a message (no GPU needed)."""
    lib_path = ROOT / "cuda-dolfinx_amd" / "lib" / "libmifem.so"
    if not lib_path.exists():
        pytest.skip("libmifem.so not built")
    from cudolfinx import _lib
    L = _lib.lib()
    n = C.c_int64(0)
    calls = {
        "mf_batch_create": lambda: L.mf_batch_create(None, 0, None),
        "mf_batch_assemble": lambda: L.mf_batch_assemble(None, None, None, 1.0, 1, None),
        "mf_batch_source": lambda: L.mf_batch_source(None, None, C.byref(n)),
        "mf_batch_assemble_vector": lambda: L.mf_batch_assemble_vector(None, None, None, 1, None),
        "mf_set_bc_batch": lambda: L.mf_set_bc_batch(None, 1, None, None, None, 1.0, None),
        "mf_system_create": lambda: L.mf_system_create(None, None, None, 0, None),
        "mf_form_set_option": lambda: L.mf_form_set_option(None, b"chunk_values", 1),
        "mf_matrix_info": lambda: L.mf_matrix_info(None, None, None, None),
        "mf_interpolator_create": lambda: L.mf_interpolator_create(None, None, None, None, None),
        "mf_assemble_scalar": lambda: L.mf_assemble_scalar(None, None, None),
        "mf_form_set_restriction": lambda: L.mf_form_set_restriction(None, 0, None, 0, 0),
        "mf_matrix_create_block": lambda: L.mf_matrix_create_block(None, 0, 0, None),
    }
    for name, call in calls.items():
        assert call() == _lib.MF_EINVAL, name
        assert L.mf_last_error(), name
    assert L.mf_batch_destroy(None) == _lib.MF_OK


# ----------------------------------------------------------------- gfx950 compile check
_ITYPES = {"cell": 0, "exterior_facet": 1, "interior_facet": 2}
_READELF = Path("/opt/rocm/llvm/bin/llvm-readelf")


def _describe(form, path):
    """The engine-side view of a form for tests/codegen_check.cpp: rank, shapes,
    coefficient layout, and each integral's UFCx body + staged description."""
    if form.rank:
        V0, V1 = form.function_spaces[0], form.function_spaces[-1]
        sh = (V0.element.ndofs, V0.bs, V1.element.ndofs, V1.bs)
    else:
        sh = (1, 1, 1, 1)
    off = form.coefficient_offsets
    with open(path, "w") as o:
        o.write(f"{form.rank} {form.mesh.nv_per_cell} {form.mesh.facet_table.shape[0]} {off[-1]} "
                f"{sh[0]} {sh[1]} {sh[2]} {sh[3]} {len(form.integrals)} {len(form.coefficients)}\n")
        for j in range(len(form.coefficients)):
            o.write(f"{off[j]} {off[j + 1] - off[j]}\n")
        for i, I in enumerate(form.integrals):
            st = I.staged.source() if I.staged is not None else ""
            o.write(f"{_ITYPES[I.type]} {i} {I.name} {len(I.source.encode())} {len(st.encode())}\n")
            o.write(I.source)
            o.write(st)


def _kernels(co: Path):
    """{kernel name: (vgprs, scratch bytes, LDS bytes)} from a code object's metadata."""
    notes = subprocess.run([str(_READELF), "--notes", str(co)], capture_output=True, text=True,
                           check=True).stdout
    out = {}
    for block in notes.split("  - .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block).group(1)
        get = lambda key: int(re.search(rf"\.{key}:\s+(\d+)", block).group(1))
        out[name] = (get("vgpr_count"), get("private_segment_fixed_size"),
                     get("group_segment_fixed_size"))
    return out


def test_generated_kernels_compile_for_gfx950(tmp_path):
    """Every kernel family's generated source compiles with hipRTC for gfx950
    (no GPU needed): owner-computes FP64 + FP32 variants (scalar P1 / P3),
    vector P2 elasticity, the Q3 matrix-core kernels, the SIPG cell-owner
    kernel, the fused Newton-system kernel and the one-launch batch of three
    element groups.  The owner
    kernels use no scratch memory; the batch reserves its largest group's LDS."""
    import shutil
    from mifem import dgforms  # noqa: F401
    from mifem.fem import Function as Fn
    build = ROOT / "cuda-dolfinx_amd" / "_build"
    objs = [build / f"{n}.o" for n in ("codegen", "runtime", "owner_kernel", "hex_kernel",
                                        "dg_kernel", "sparsity", "util_kernels", "matrix")]
    if not shutil.which("hipcc") or not _READELF.exists() or not all(o.exists() for o in objs):
        pytest.skip("hipcc / llvm-readelf / library objects not available")
    drv = tmp_path / "codegen_check"
    subprocess.run(["hipcc", "-c", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-O1",
                    f"-I{ROOT / 'cuda-dolfinx_amd' / 'csrc'}", f"-I{ROOT / 'include'}",
                    str(ROOT / "tests" / "codegen_check.cpp"), "-o", str(tmp_path / "check.o")],
                   check=True)
    subprocess.run(["hipcc", str(tmp_path / "check.o"), *map(str, objs), "-lhiprtc", "-o", str(drv)],
                   check=True)
    m = create_unit_cube(2)
    D0 = functionspace(m, ("DG", 0))
    kap = Fn(D0)
    forms_ = {
        "p1": forms.stiffness(functionspace(m, ("P", 1)), kap),
        "p2": forms.stiffness(functionspace(m, ("P", 2)), kap),
        "p3": forms.stiffness(functionspace(m, ("P", 3)), kap),
        "elasticity": problems.elasticity_p2(2).a,
        "q3": problems.helmholtz_q3(2).a,
        }
    Pdg = problems.dg_plaplace(2)
    forms_["sipg"] = Pdg.a
    forms_["sipg_F"] = Pdg.L
    forms_["sipg_E"] = dgforms.sipg_energy(Pdg.V, Pdg.extras["u"])
    for k, f in forms_.items():
        _describe(f, tmp_path / f"{k}.txt")
    runs = {k: [str(drv), "form", str(tmp_path / f"{k}.txt"), str(tmp_path / f"{k}.co")]
            for k in ("p1", "p3", "elasticity", "q3", "sipg")}
    runs["batch"] = [str(drv), "batch", str(tmp_path / "batch.co")] + \
        [str(tmp_path / f"{k}.txt") for k in ("p1", "p2", "p3")]
    runs["system"] = [str(drv), "system", str(tmp_path / "system.co")] + \
        [str(tmp_path / f"{k}.txt") for k in ("sipg", "sipg_F", "sipg_E")]
    procs = {k: subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
             for k, c in runs.items()}
    outs = {k: p.communicate(timeout=600)[0] for k, p in procs.items()}
    for k, o in outs.items():
        assert o.startswith("OK"), (k, o[-2000:])
    K = {k: _kernels(tmp_path / f"{k}.co") for k in runs}
    for k in ("p1", "p3"):
        assert {"mf_prep_cell_0", "mf_owner_cell_0", "mf_owner32_cell_0", "mf_mat_cell_0"} <= set(K[k])
        assert K[k]["mf_owner_cell_0"][1] == 0 and K[k]["mf_owner32_cell_0"][1] == 0
        assert K[k]["mf_owner32_cell_0"][2] < K[k]["mf_owner_cell_0"][2]    # FP32 tables
    assert "mf_owner32_cell_0" not in K["elasticity"] and K["elasticity"]["mf_owner_cell_0"][1] == 0
    assert {"mf_hprep_cell_0", "mf_howner64_cell_0", "mf_howner32_cell_0",
            "mf_howner16_cell_0"} <= set(K["q3"])
    # matrix-core FP32 / BF16 kernels: no scratch; the FP64 one keeps 48 B (the per-entity
    # kernels hold the whole 64 x 64 element tensor per thread in scratch, as the reference does)
    assert K["q3"]["mf_howner32_cell_0"][1] == 0 and K["q3"]["mf_howner16_cell_0"][1] == 0
    assert K["q3"]["mf_howner64_cell_0"][1] <= 64
    assert "mf_dgown" in K["sipg"]
    assert K["system"]["mf_dgsys"][1] == 0                      # fused J + F + E, no scratch
    # the J, F and E bodies' constant tables become shared file-scope arrays
    # (identical tables once each); none stays function-local
    declared, shared, local = map(int, outs["system"].split()[-3:])
    assert 0 < shared < declared and local == 0, (declared, shared, local)
    assert K["sipg"]["mf_dgown"][1] == 0
    lds = int(outs["batch"].split()[-1])
    assert {"mf_owner_batch", "mf_prep_g0", "mf_prep_g1", "mf_prep_g2"} <= set(K["batch"])
    assert K["batch"]["mf_owner_batch"][1] == 0 and K["batch"]["mf_owner_batch"][2] == lds
    assert lds == K["p3"]["mf_owner_cell_0"][2]                # the largest group's LDS
