"""Pins of the CPU oracle against closed forms, library routines and brute force.

Each check is chosen so that a plausible mistake in the oracle (wrong
quadrature weight, basis index, Jacobian transpose, dropped elasticity
term, bc masking order) fails it.
"""
import math

import numpy as np
import pytest
from scipy.special import roots_jacobi

from oracle import cpu
from mifem import element as E
from mifem import forms
from mifem import problems
from mifem.fem import Constant, Function, dirichletbc, functionspace, locate_dofs_topological
from mifem.mesh import create_unit_cube


# ------------------------------------------------------------ quadrature
@pytest.mark.parametrize("alpha", [0.0, 1.0, 2.0])
@pytest.mark.parametrize("m", [1, 2, 3, 5, 8])
def test_gauss_jacobi_matches_scipy(alpha, m):
    t, w = cpu.gauss_jacobi(m, alpha)
    x, wx = roots_jacobi(m, alpha, 0.0)
    order = np.argsort(t)
    np.testing.assert_allclose(t[order], (x + 1) / 2, atol=1e-15)
    np.testing.assert_allclose(w[order], wx / 2 ** (alpha + 1), rtol=1e-13)


@pytest.mark.parametrize("m", [1, 2, 3, 4])
def test_tet_rule_exactness(m):
    pts, wts = cpu.tet_rule(m)
    for a in range(2 * m):
        for b in range(2 * m - a):
            for c in range(2 * m - a - b):
                exact = math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(wts * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) < 1e-15, (a, b, c)
    # one degree more is not integrated exactly (rule really has degree 2m-1)
    d = 2 * m
    exact = math.factorial(d) / math.factorial(d + 3)
    assert abs(np.sum(wts * pts[:, 0] ** d) - exact) > 1e-12


@pytest.mark.parametrize("m", [1, 2, 4])
def test_hex_and_triangle_rules(m):
    pts, wts = cpu.hex_rule(m)
    for a in range(2 * m):
        assert abs(np.sum(wts * pts[:, 1] ** a * pts[:, 2]) - 1 / ((a + 1) * 2)) < 1e-15
    tp, tw = cpu.tri_rule(m)
    for a in range(2 * m):
        for b in range(2 * m - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            assert abs(np.sum(tw * tp[:, 0] ** a * tp[:, 1] ** b) - exact) < 1e-15


# ------------------------------------------------------------ bases
@pytest.mark.parametrize("cell,k", [(0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3)])
def test_basis_nodal_partition_and_gradients(cell, k):
    el = E.Element("P", "tetrahedron" if cell == 0 else "hexahedron", k)
    X = el.nodes
    V = np.array([cpu.basis(cell, k, x)[0] for x in X])
    np.testing.assert_allclose(V, np.eye(len(X)), atol=1e-14)
    rng = np.random.default_rng(0)
    P = rng.random((7, 3)) * (0.3 if cell == 0 else 1.0)
    for p in P:
        phi, dphi = cpu.basis(cell, k, p)
        assert abs(phi.sum() - 1) < 1e-14
        np.testing.assert_allclose(dphi.sum(axis=0), 0, atol=1e-12)
        # gradient of the interpolant of a degree-k polynomial is exact
        f = lambda x: x[..., 0] ** k + 2 * x[..., 1] * x[..., 2] ** (k - 1) - x[..., 2]
        gf = np.array([k * p[0] ** (k - 1), 2 * p[2] ** (k - 1),
                       2 * (k - 1) * p[1] * p[2] ** max(k - 2, 0) - 1])
        vals = f(X)
        np.testing.assert_allclose(vals @ dphi, gf, atol=1e-11)
        assert abs(vals @ phi - f(p)) < 1e-12


# ------------------------------------------------------------ element kernels
def _tet(rng):
    x = rng.random((4, 3))
    while abs(np.linalg.det((x[1:] - x[0]).T)) < 0.05:
        x = rng.random((4, 3))
    return x


def test_p1_stiffness_and_mass_closed_form():
    rng = np.random.default_rng(1)
    x = _tet(rng)
    vol = abs(np.linalg.det((x[1:] - x[0]).T)) / 6
    # barycentric gradients: rows of inv([[1, x]])
    Gm = np.linalg.inv(np.hstack([np.ones((4, 1)), x]))[1:].T    # (4, 3)
    spec = dict(kind="stiffness", cell=0, degree=1, bs=1, rank=2, qdeg=0, kappa=("val", 2.5), k2=0.0)
    A = cpu.tabulate(spec, [], [], x)
    np.testing.assert_allclose(A, 2.5 * vol * Gm @ Gm.T, rtol=1e-12, atol=1e-14)
    spec = dict(kind="stiffness", cell=0, degree=1, bs=1, rank=2, qdeg=2, kappa=("val", 0.0), k2=-1.0)
    M = cpu.tabulate(spec, [], [], x)
    np.testing.assert_allclose(M, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-12)


def test_p2_elasticity_element_rigid_modes_and_psd():
    rng = np.random.default_rng(2)
    x = _tet(rng)
    spec = dict(kind="elasticity", cell=0, degree=2, bs=3, rank=2, qdeg=2,
                mu=("w", 0), lam=("w", 1))
    A = cpu.tabulate(spec, [2.0, 3.0], [], x)
    np.testing.assert_allclose(A, A.T, atol=1e-12 * abs(A).max())
    ev = np.linalg.eigvalsh(A)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 6 and ev.min() > -1e-10 * ev.max()
    # nodes of P2 in physical space
    lam = np.concatenate([1 - E.Element("P", "tetrahedron", 2).nodes.sum(1, keepdims=True),
                          E.Element("P", "tetrahedron", 2).nodes], axis=1)
    X = lam @ x
    for mode in (lambda p: np.array([1.0, 0, 0]), lambda p: np.array([-p[1], p[0], 0.0]),
                 lambda p: np.array([0.0, -p[2], p[1]])):
        u = np.concatenate([mode(p) for p in X])
        assert np.abs(A @ u).max() < 1e-10 * abs(A).max()
    # scaling: the stiffness of a tetrahedron scales like h
    A2 = cpu.tabulate(spec, [2.0, 3.0], [], 2 * x)
    np.testing.assert_allclose(A2, 2 * A, rtol=1e-12, atol=1e-13 * abs(A).max())


def _p2_nodes(x):
    """Physical coordinates of the P2 nodes of the tetrahedron x (4 x 3)."""
    ref = E.Element("P", "tetrahedron", 2).nodes
    lam = np.concatenate([1 - ref.sum(1, keepdims=True), ref], axis=1)
    return lam @ x


# Linear displacement fields and their elastic energy densities
# 2 mu eps(u):eps(u) + lambda (div u)^2 (constant for a linear field):
#   u = (x, 0, 0): eps = diag(1, 0, 0)          -> 2 mu + lambda
#   u = (y, 0, 0): eps_xy = eps_yx = 1/2, div 0 -> mu
#   u = (x, y, z): eps = I, div u = 3            -> 6 mu + 9 lambda
_ENERGY_FIELDS = [
    (lambda p: np.array([p[0], 0.0, 0.0]), lambda mu, lam: 2 * mu + lam),
    (lambda p: np.array([p[1], 0.0, 0.0]), lambda mu, lam: mu),
    (lambda p: np.array([p[0], p[1], p[2]]), lambda mu, lam: 6 * mu + 9 * lam),
]


@pytest.mark.parametrize("mu,lam", [(2.0, 3.0), (3.0, 2.0), (1.0, 7.5)])
def test_p2_elasticity_element_closed_form_energies(mu, lam):
    """u^T A u of linear fields on one random tetrahedron equals the closed-form
    energy times |T|: the three fields tell mu from lambda, so a swapped pair,
    a dropped lambda div-div term or a dropped mu trace term fails."""
    rng = np.random.default_rng(7)
    x = _tet(rng)
    vol = abs(np.linalg.det((x[1:] - x[0]).T)) / 6
    spec = dict(kind="elasticity", cell=0, degree=2, bs=3, rank=2, qdeg=2,
                mu=("w", 0), lam=("w", 1))
    A = cpu.tabulate(spec, [mu, lam], [], x)
    X = _p2_nodes(x)
    for field, energy in _ENERGY_FIELDS:
        u = np.concatenate([field(p) for p in X])
        assert abs(u @ A @ u - energy(mu, lam) * vol) < 1e-12 * energy(mu, lam) * vol


def test_elasticity_global_energy_cellwise_lame():
    """Global oracle assembly with the seeded per-cell Lame fields: u^T A u of
    each linear field is sum_c energy(mu_c, lambda_c) |T_c| (pins that each
    cell reads its own coefficients, in the right slots)."""
    P = problems.elasticity_p2(2)
    O = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = O.sparsity()
    A = _dense(rp, cols, O.assemble_matrix(rp, cols), P.V.size)
    mesh = P.mesh
    dg = P.extras["mu"].function_space.dofmap[:, 0]
    mu_c, lam_c = P.extras["mu"].x[dg], P.extras["lambda"].x[dg]
    assert np.ptp(mu_c) > 1 and np.ptp(lam_c) > 1
    xc = mesh.x[mesh.cells]
    vol = np.abs(np.linalg.det(np.transpose(xc[:, 1:] - xc[:, :1], (0, 2, 1)))) / 6
    X = P.V.node_coordinates
    for field, energy in _ENERGY_FIELDS:
        u = np.concatenate([field(p) for p in X])
        exact = float(np.sum(energy(mu_c, lam_c) * vol))
        assert abs(u @ A @ u - exact) < 1e-12 * exact


def _hex_dense(P, spec):
    """Dense global matrix of a hexahedral form from per-cell oracle tabulation."""
    mesh, V = P.mesh, P.V
    A = np.zeros((V.size, V.size))
    for c in range(mesh.num_cells):
        d = V.dofmap[c]
        A[np.ix_(d, d)] += cpu.tabulate(spec, [], [], mesh.x[mesh.cells[c]])
    return A


@pytest.mark.parametrize("degree", [2, 3])
def test_jittered_hex_linear_patch(degree):
    """Trilinear (non-affine) hexahedra: for u = a . x, which Q_k holds exactly,
    the stiffness rows of interior nodes vanish and u^T K u = |a|^2 vol(=1);
    the mass gives u^T M u = int_(unit cube) (a . x)^2.  Every integrand is a
    polynomial the Gauss rule integrates exactly, so these hold to round-off;
    a transposed Jacobian or a wrong inverse on non-affine cells breaks them."""
    P = problems.helmholtz_q3(3, degree=degree)
    spec = dict(P.a.integrals[0].spec)
    a = np.array([0.3, -0.7, 0.5])
    X = P.V.node_coordinates
    u = X @ a
    spec.update(kappa=("val", 1.0), k2=0.0)
    K = _hex_dense(P, spec)
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    assert interior.sum() > 0
    assert np.abs((K @ u)[interior]).max() < 1e-13
    assert abs(u @ K @ u - a @ a) < 1e-13
    spec.update(kappa=("val", 0.0), k2=-1.0)
    M = _hex_dense(P, spec)
    # int (a . x)^2 over [0,1]^3 = sum_i a_i^2 / 3 + sum_{i != j} a_i a_j / 4
    exact = (a @ a) / 3 + (a.sum() ** 2 - a @ a) / 4
    assert abs(u @ M @ u - exact) < 1e-13


def _gll1d_matrices(k, m):
    t = E.gll_points(k)
    V = np.vander(t, k + 1, increasing=True)
    C = np.linalg.inv(V)
    xg, wg = np.polynomial.legendre.leggauss(m)
    xg, wg = (xg + 1) / 2, wg / 2
    P = np.vander(xg, k + 1, increasing=True) @ C
    dP = np.zeros_like(P)
    for j in range(1, k + 1):
        dP += j * np.outer(xg ** (j - 1), C[j])
    return dP.T @ (wg[:, None] * dP), P.T @ (wg[:, None] * P)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_hex_kronecker_structure(k):
    x = np.array([(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=float)
    x = x * np.array([0.5, 0.25, 2.0])          # affine box
    spec = dict(kind="stiffness", cell=1, degree=k, bs=1, rank=2, qm=k + 1, kappa=("val", 1.0), k2=4.0)
    A = cpu.tabulate(spec, [], [], x)
    Kx, Mx = _gll1d_matrices(k, k + 1)
    hx, hy, hz = 0.5, 0.25, 2.0
    K = (np.kron(np.kron(Mx * hz, Mx * hy), Kx / hx) + np.kron(np.kron(Mx * hz, Kx / hy), Mx * hx)
         + np.kron(np.kron(Kx / hz, Mx * hy), Mx * hx))
    M = np.kron(np.kron(Mx * hz, Mx * hy), Mx * hx)
    np.testing.assert_allclose(A, K - 4.0 * M, rtol=1e-11, atol=1e-12 * abs(K).max())


# ------------------------------------------------------------ global assembly
def _dense(rp, cols, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, cols[rp[r]:rp[r + 1]]] += vals[rp[r]:rp[r + 1]]
    return A


def test_c1_closed_forms_full_size():
    """32^3 x 6 P1 Poisson: nnz closed form and the 7-point stencil times h."""
    n = 32
    P = problems.poisson_p1(n)
    O = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = O.sparsity()
    E_ = 3 * n * (n + 1) ** 2 + 3 * n ** 2 * (n + 1) + n ** 3
    assert cols.size == (n + 1) ** 3 + 2 * E_ == 513313
    vals = O.assemble_matrix(rp, cols)
    X = P.V.node_coordinates
    h = 1.0 / n
    interior = np.nonzero(np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1))[0]
    for r in interior[::997]:
        c, v = cols[rp[r]:rp[r + 1]], vals[rp[r]:rp[r + 1]] / h
        assert c.size == 15
        d = np.abs(X[c] - X[r]).sum(axis=1) * n
        assert abs(v[c == r][0] - 6) < 1e-12
        axis_nb = np.isclose(d, 1.0)
        assert axis_nb.sum() == 6 and np.allclose(v[axis_nb], -1, atol=1e-12)
        assert np.allclose(v[~axis_nb & (c != r)], 0, atol=1e-12)


def test_sparsity_brute_force():
    P = problems.elasticity_p2(2)
    O = cpu.OracleAssembler(P.a)
    rp, cols, rpb, colsb = O.sparsity()
    dm = P.V.unrolled_dofmap()
    pairs = set()
    for c in range(dm.shape[0]):
        for r in dm[c]:
            for q in dm[c]:
                pairs.add((int(r), int(q)))
    got = {(r, int(q)) for r in range(P.V.size) for q in cols[rp[r]:rp[r + 1]]}
    assert got == pairs
    for r in range(P.V.size):
        assert np.all(np.diff(cols[rp[r]:rp[r + 1]]) > 0)


def test_bc_semantics_brute_force():
    """A_bc = P A P + D, lifting b - s A[:, bc](g - x0) and set_bc, by dense algebra."""
    P = problems.poisson_p1(3)
    rng = np.random.default_rng(3)
    g = Function(P.V)
    g.x[:] = rng.random(P.V.size)
    bc = dirichletbc(g, locate_dofs_topological(P.V, 2, P.mesh.exterior_facets), P.V)
    O = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = O.sparsity()
    n = P.V.size
    A = _dense(rp, cols, O.assemble_matrix(rp, cols), n)
    Abc = _dense(rp, cols, O.assemble_matrix(rp, cols, [bc], diagonal=3.0), n)
    keep = np.ones(n)
    keep[bc.dofs] = 0
    ref = A * keep[:, None] * keep[None, :]
    ref[bc.dofs, bc.dofs] = 3.0
    np.testing.assert_allclose(Abc, ref, atol=1e-14)
    OL = cpu.OracleAssembler(P.L)
    b = OL.assemble_vector()
    x0 = rng.random(n)
    b2 = O.apply_lifting(b.copy(), [bc], x0=x0, scale=0.7)
    gd = np.zeros(n)
    gd[bc.dofs] = g.x[bc.dofs] - x0[bc.dofs]
    np.testing.assert_allclose(b2, b - 0.7 * A @ gd, atol=1e-13)
    b3 = cpu.set_bc(b2.copy(), [bc], P.V, x0=x0, scale=0.7)
    exp = b2.copy()
    exp[bc.dofs] = 0.7 * (g.x[bc.dofs] - x0[bc.dofs])
    np.testing.assert_allclose(b3, exp)


def test_elasticity_patch_and_load_totals():
    mesh = create_unit_cube(3)
    V = functionspace(mesh, ("P", 2), shape=(3,))
    a = forms.elasticity(V, 2.0, 5.0)
    O = cpu.OracleAssembler(a)
    rp, cols, _, _ = O.sparsity()
    A = _dense(rp, cols, O.assemble_matrix(rp, cols), V.size)
    X = V.node_coordinates
    u = (X @ np.array([[0.3, 0.1, -0.2], [0.0, 0.5, 0.4], [-0.1, 0.2, 0.7]])).reshape(-1)
    r = (A @ u).reshape(-1, 3)
    interior = np.all((X > 1e-9) & (X < 1 - 1e-9), axis=1)
    assert np.abs(r[interior]).max() < 1e-11 * np.abs(A).max()
    P = problems.elasticity_p2(3)
    b = cpu.OracleAssembler(P.L).assemble_vector().reshape(-1, 3)
    # body force (0,0,-1) over volume 1 plus traction (0,0,-1) over the unit x=1 face
    np.testing.assert_allclose(b.sum(axis=0), [0.0, 0.0, -2.0], atol=1e-13)


def test_jittered_hex_mass_sums_to_volume():
    P = problems.helmholtz_q3(3)
    spec = dict(P.a.integrals[0].spec)
    spec.update(kappa=("val", 0.0), k2=-1.0)
    mesh = P.mesh
    total = 0.0
    for c in range(mesh.num_cells):
        total += cpu.tabulate(spec, [], [], mesh.x[mesh.cells[c]]).sum()
    assert abs(total - 1.0) < 1e-13


# ------------------------------------------------------------ SIPG p-Laplacian (DG1)
_TF = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]


def _dg1_energy(mesh, u, p, eps, f, alpha, penalty, mu0):
    """E(u) whose derivative is the SIPG residual, from vertex values only.

    Cell gradients come from solving the vertex-difference system, facet
    integrals of products of linear functions from the exact triangle rule
    int_T a b = |T| / 12 (sum a_k b_k + sum a_k sum b_k); nothing here shares
    code or tables with the oracle.
    """
    cells, x = mesh.cells, mesh.x
    U = u.reshape(-1, 4)
    E_ = 0.0
    grads, diam = [], []
    for c in range(mesh.num_cells):
        xc = x[cells[c]]
        D = xc[1:] - xc[0]
        g = np.linalg.solve(D, U[c, 1:] - U[c, 0])
        vol = abs(np.linalg.det(D)) / 6.0
        s = eps * eps + g @ g
        E_ += vol * (s ** (p / 2) / p + 0.5 * mu0 * (g @ g) - f * U[c].mean())
        grads.append(g)
        diam.append(max(np.linalg.norm(xc[a] - xc[b]) for a in range(4) for b in range(a + 1, 4)))

    def tri(c, lf):
        v = [cells[c][k] for k in _TF[lf]]
        P = x[v]
        n = np.cross(P[1] - P[0], P[2] - P[0])
        area = 0.5 * np.linalg.norm(n)
        n = n / np.linalg.norm(n)
        if n @ (x[cells[c][lf]] - P[0]) > 0:
            n = -n
        return v, area, n

    def int2(a, b, area):
        return area / 12.0 * (a @ b + a.sum() * b.sum())

    for c0, l0, c1, l1 in mesh.interior_facets:
        v, area, n = tri(c0, l0)
        u0 = np.array([U[c0][list(cells[c0]).index(g)] for g in v])
        u1 = np.array([U[c1][list(cells[c1]).index(g)] for g in v])
        j = u0 - u1
        hF = 0.5 * (diam[c0] + diam[c1])
        avg_dn = 0.5 * (grads[c0] + grads[c1]) @ n
        E_ += 0.5 * penalty / hF * int2(j, j, area) - avg_dn * area * j.mean()
    for c, lf in mesh.exterior_facets:
        v, area, n = tri(c, lf)
        uv = np.array([U[c][list(cells[c]).index(g)] for g in v])
        E_ += 0.5 * alpha * int2(uv, uv, area)
    return E_


def _dg_problem(n=2, p=3.0):
    P = problems.dg_plaplace(n, p=p)
    rng = np.random.default_rng(7)
    P.extras["u"].x[:] = rng.uniform(-1, 1, P.V.size)
    return P


def _dg_residual(P, u):
    P.extras["u"].x[:] = u
    return cpu.OracleAssembler(P.L).assemble_vector()


@pytest.mark.parametrize("p", [2.0, 3.0])
def test_sipg_residual_is_energy_gradient(p):
    P = _dg_problem(2, p)
    u = P.extras["u"].x.copy()
    F = _dg_residual(P, u)
    spec = P.L.integrals[0].spec
    args = (P.mesh, None, spec["p"], spec["eps"], spec["fval"], spec["alpha"], spec["penalty"],
            spec["mu0"])
    rng = np.random.default_rng(3)
    for _ in range(4):
        d = rng.standard_normal(u.size)
        h = 1e-5
        dE = (_dg1_energy(P.mesh, u + h * d, *args[2:]) - _dg1_energy(P.mesh, u - h * d, *args[2:])) / (2 * h)
        assert abs(dE - F @ d) <= 1e-7 * max(1.0, abs(dE)), (dE, F @ d)


def test_sipg_jacobian_is_residual_derivative():
    P = _dg_problem(2)
    u = P.extras["u"].x.copy()
    oa = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = oa.sparsity()
    vals = oa.assemble_matrix(rp, cols)
    import scipy.sparse as sp
    J = sp.csr_matrix((vals, cols, rp), shape=(P.V.size, P.V.size))
    assert abs(J - J.T).max() <= 1e-13 * abs(vals).max()
    rng = np.random.default_rng(4)
    d = rng.standard_normal(u.size)
    h = 1e-6
    fd = (_dg_residual(P, u + h * d) - _dg_residual(P, u - h * d)) / (2 * h)
    Jd = J @ d
    assert np.linalg.norm(fd - Jd) <= 1e-7 * np.linalg.norm(Jd)


def test_sipg_consistency_for_linear_fields():
    """p = 2, f = 0, alpha = 0: for a continuous linear u the cell and interior
    facet terms sum to the boundary flux int_dOmega (grad u . n) v."""
    from mifem.dgforms import sipg_plaplace
    mesh = create_unit_cube(2, jitter=0.2, seed=2)
    V = functionspace(mesh, ("DG", 1))
    u = Function(V)
    gvec = np.array([0.3, -1.1, 0.7])
    u.interpolate(lambda x: 0.5 + gvec @ x)
    _, F = sipg_plaplace(V, u, f=0.0, alpha=0.0, p=2.0, eps=0.0, mu0=0.0)
    r = cpu.OracleAssembler(F).assemble_vector().reshape(-1, 4)
    ref = np.zeros_like(r)
    for c, lf in mesh.exterior_facets:
        v = mesh.cells[c][list(_TF[lf])]
        P = mesh.x[v]
        nrm = np.cross(P[1] - P[0], P[2] - P[0])
        area = 0.5 * np.linalg.norm(nrm)
        nrm /= np.linalg.norm(nrm)
        if nrm @ (mesh.x[mesh.cells[c][lf]] - P[0]) > 0:
            nrm = -nrm
        for k in _TF[lf]:
            ref[c, k] += (gvec @ nrm) * area / 3.0
    np.testing.assert_allclose(r, ref, atol=1e-13)


# ------------------------------------------------------------ many-instance batch
def test_batch_blocks_equal_independent_instances():
    """Each block of the merged block-diagonal system equals the instance
    assembled on its own (CSR bit-exact, values bit-exact: same cell order)."""
    from mifem import batch
    B = batch.many_poisson(6, seed=3, nmin=2, nmax=4)
    assert sorted(i for g in B.groups for i in g.instances) == list(range(6))
    for g in B.groups:
        oa = cpu.OracleAssembler(g.a)
        rp, cols, _, _ = oa.sparsity()
        vals = oa.assemble_matrix(rp, cols, g.bcs)
        b = cpu.OracleAssembler(g.L).assemble_vector()
        oa.apply_lifting(b, g.bcs)
        cpu.set_bc(b, g.bcs, g.V)
        for j, i in enumerate(g.instances):
            mesh, V, kap = batch.poisson_instance(B.sizes[i], g.degree)
            D = functionspace(mesh, ("DG", 0))
            kf = Function(D)
            kf.set_cellwise(kap)
            bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
            o = cpu.OracleAssembler(forms.stiffness(V, kf))
            r2, c2, _, _ = o.sparsity()
            v2 = o.assemble_matrix(r2, c2, [bc])
            r1, c1, v1 = g.block_csr(rp, cols, vals, j)
            assert np.array_equal(r1, r2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
            b2 = cpu.OracleAssembler(forms.source(V, Constant([1.0]))).assemble_vector()
            cpu.set_bc(b2, [bc], V)
            assert np.array_equal(b[g.dof_offsets[j]:g.dof_offsets[j + 1]], b2)


# ------------------------------------------------------------ functionals (rank 0)
def _sub(form, k):
    """The functional restricted to its k-th integral."""
    return forms.Form(0, [], [form.integrals[k]], form.coefficients, [], domain=form.mesh)


def _fn_setup(n, jitter, k=1, kv=1, seed=4):
    from mifem.functionals import mixed_functional
    mesh = create_unit_cube(n, jitter=jitter, seed=seed)
    V, D = functionspace(mesh, ("P", k)), functionspace(mesh, ("DG", kv))
    u, kap, v = Function(V), Function(V), Function(D)
    return mesh, u, kap, v, mixed_functional(u, kap, v)


def _tri_area(P):
    return 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))


def _int_tri_product(P, fs):
    """Exact integral over triangle P of a product of linear functions given by
    their vertex values fs[m][0..2]: int_T lambda^a = 2|T| a! / (|a| + 2)!."""
    area = _tri_area(P)
    tot = 0.0
    for idx in np.ndindex(*(3,) * len(fs)):
        a = np.bincount(idx, minlength=3)
        mono = 2 * area * np.prod([math.factorial(t) for t in a]) / math.factorial(len(fs) + 2)
        tot += np.prod([fs[m][idx[m]] for m in range(len(fs))]) * mono
    return tot


def test_functional_cell_term_closed_form():
    """int exp(u) kappa dx with u linear, kappa = 1: per tetrahedron
    int_T e^l = 6 |T| sum_i e^{l_i} / prod_{j != i} (l_i - l_j)."""
    mesh, u, kap, v, M = _fn_setup(3, 0.0)
    a = np.array([0.3, -0.5, 0.7])
    u.interpolate(lambda x: 0.1 + a @ x)
    kap.x[:] = 1.0
    got = cpu.OracleAssembler(_sub(M, 0)).assemble_scalar()
    ref = 0.0
    for c in mesh.cells:
        X = mesh.x[c]
        vol = abs(np.linalg.det(X[1:] - X[0])) / 6.0
        l = 0.1 + X @ a
        ref += 6 * vol * sum(np.exp(l[i]) / np.prod([l[i] - l[j] for j in range(4) if j != i])
                             for i in range(4))
    assert abs(got - ref) <= 1e-7 * abs(ref), (got, ref)
    # constant u: exp(c) times the volume of the (jittered) unit cube
    mesh, u, kap, v, M = _fn_setup(3, 0.2)
    u.x[:] = 0.4
    kap.x[:] = 1.0
    assert abs(cpu.OracleAssembler(_sub(M, 0)).assemble_scalar() - math.exp(0.4)) <= 1e-13


def test_functional_exterior_term_closed_form():
    """int_dOmega u kappa ds with u = y, kappa = z: 1/4 + 1/4 (x faces) + 1/2 (y = 1)
    + 1/2 (z = 1) = 3/2."""
    mesh, u, kap, v, M = _fn_setup(3, 0.2)
    u.interpolate(lambda x: x[1])
    kap.interpolate(lambda x: x[2])
    assert abs(cpu.OracleAssembler(_sub(M, 1)).assemble_scalar() - 1.5) <= 1e-13


@pytest.mark.parametrize("kv", [1, 2])
def test_functional_interior_term_exact(kv):
    """int_F avg(v^2) avg(kappa) over interior facets: kappa = x and a
    continuous linear v give a cubic, integrated exactly from barycentric
    monomials; v = 1, kappa = 1 gives the total interior facet area."""
    mesh, u, kap, v, M = _fn_setup(3, 0.2, kv=kv)
    b = np.array([0.6, -0.2, 0.9])
    kap.interpolate(lambda x: x[0])
    v.interpolate(lambda x: 0.3 + b @ x)
    got = cpu.OracleAssembler(_sub(M, 2)).assemble_scalar()
    ref = 0.0
    for c0, l0, _, _ in mesh.interior_facets:
        P = mesh.x[mesh.cells[c0][list(_TF[l0])]]
        vv = 0.3 + P @ b
        ref += _int_tri_product(P, [P[:, 0], vv, vv])
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    kap.x[:] = 1.0
    v.x[:] = 1.0
    area = sum(_tri_area(mesh.x[mesh.cells[c0][list(_TF[l0])]])
               for c0, l0, _, _ in mesh.interior_facets)
    assert abs(cpu.OracleAssembler(_sub(M, 2)).assemble_scalar() - area) <= 1e-12 * area


def test_sipg_energy_matches_independent_energy():
    """DG1: the oracle's energy functional equals the vertex-value energy
    computed without the oracle (exact for DG1: constant gradients, linear
    traces)."""
    from mifem.dgforms import sipg_energy
    P = _dg_problem(2)
    spec = P.L.integrals[0].spec
    E_ = sipg_energy(P.V, P.extras["u"], f=spec["fval"], alpha=spec["alpha"],
                     penalty=spec["penalty"], p=spec["p"], eps=spec["eps"], mu0=spec["mu0"])
    got = cpu.OracleAssembler(E_).assemble_scalar()
    ref = _dg1_energy(P.mesh, P.extras["u"].x, spec["p"], spec["eps"], spec["fval"], spec["alpha"],
                      spec["penalty"], spec["mu0"])
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)


@pytest.mark.parametrize("k", [1, 2])
def test_sipg_energy_gradient_is_residual(k):
    """d/dh E(u + h d) at h = 0 equals F(u) . d (central differences)."""
    from mifem.dgforms import sipg_energy
    P = problems.dg_plaplace(2, degree=k)
    u = P.extras["u"]
    rng = np.random.default_rng(11)
    u.x[:] = rng.uniform(-1, 1, P.V.size)
    E_ = sipg_energy(P.V, u)
    u0 = u.x.copy()
    F = cpu.OracleAssembler(P.L).assemble_vector()
    d = rng.standard_normal(u0.size)
    h = 1e-5
    u.x[:] = u0 + h * d
    ep = cpu.OracleAssembler(E_).assemble_scalar()
    u.x[:] = u0 - h * d
    em = cpu.OracleAssembler(E_).assemble_scalar()
    u.x[:] = u0
    dE = (ep - em) / (2 * h)
    assert abs(dE - F @ d) <= 1e-7 * max(1.0, abs(dE)), (dE, F @ d)


# ------------------------------------------------------------ same-map interpolation
def _poly(deg):
    c = np.array([0.7, -0.3, 0.5, 1.1])

    def f(x):
        v = c[0] + c[1] * x[0] + c[2] * x[1] + c[3] * x[2]
        if deg >= 2:
            v = v + 0.4 * x[0] * x[1] - 0.9 * x[2] ** 2
        if deg >= 3:
            v = v + 0.6 * x[0] * x[1] * x[2] - 0.2 * x[1] ** 3
        return v
    return f


@pytest.mark.parametrize("cell,fam0,k0,fam1,k1,bs", [
    ("tetrahedron", "P", 1, "P", 2, 1), ("tetrahedron", "P", 2, "P", 3, 1),
    ("tetrahedron", "P", 1, "P", 3, 1), ("tetrahedron", "P", 2, "DG", 2, 1),
    ("tetrahedron", "DG", 1, "DG", 3, 1), ("tetrahedron", "P", 2, "P", 3, 3),
    ("tetrahedron", "P", 2, "P", 1, 1), ("hexahedron", "P", 1, "P", 3, 1),
    ("hexahedron", "P", 2, "P", 3, 1)])
def test_interpolation_exact_for_polynomials(cell, fam0, k0, fam1, k1, bs):
    """Interpolating the degree-k0 interpolant of a polynomial of degree <= k0
    into a space that contains it (or onto a subset of its nodes) reproduces the
    target's own nodal interpolant of that polynomial."""
    from cudolfinx.coefficient import interpolation_operator
    mesh = create_unit_cube(2, cell, jitter=0.15, seed=6)
    shape = (bs,) if bs > 1 else None
    V0 = functionspace(mesh, (fam0, k0), shape=shape)
    V1 = functionspace(mesh, (fam1, k1), shape=shape)
    deg = min(k0, k1) if cell == "tetrahedron" else 1
    p = _poly(deg)
    f0, f1 = Function(V0), Function(V1)
    g = (lambda x: np.stack([p(x), 2 * p(x), -p(x)])) if bs > 1 else p
    f0.interpolate(g)
    f1.interpolate(g)
    u1 = cpu.interpolate_same_map(V1, V0, interpolation_operator(V1.element, V0.element), f0.x)
    np.testing.assert_allclose(u1, f1.x, atol=1e-13 * np.abs(f1.x).max())


def test_interpolation_to_dg0_is_vertex_mean_of_p1():
    from cudolfinx.coefficient import interpolation_operator
    mesh = create_unit_cube(3, jitter=0.2, seed=2)
    V0, V1 = functionspace(mesh, ("P", 1)), functionspace(mesh, ("DG", 0))
    u0 = np.random.default_rng(3).uniform(-1, 1, V0.size)
    u1 = cpu.interpolate_same_map(V1, V0, interpolation_operator(V1.element, V0.element), u0)
    ref = u0[V0.dofmap].mean(axis=1)
    np.testing.assert_allclose(u1[V1.dofmap[:, 0]], ref, atol=1e-15)


def test_interpolation_shared_nodes_take_last_cell():
    """From a discontinuous field into P1: each vertex takes the value of the
    highest-numbered cell containing it (brute force over the cell list)."""
    from cudolfinx.coefficient import interpolation_operator
    mesh = create_unit_cube(2, jitter=0.1, seed=1)
    V0, V1 = functionspace(mesh, ("DG", 1)), functionspace(mesh, ("P", 1))
    u0 = np.random.default_rng(5).uniform(-1, 1, V0.size)
    u1 = cpu.interpolate_same_map(V1, V0, interpolation_operator(V1.element, V0.element), u0)
    ref = np.full(V1.size, np.nan)
    for c in range(mesh.num_cells):
        for i in range(4):
            ref[V1.dofmap[c, i]] = u0[V0.dofmap[c, i]]
    np.testing.assert_allclose(u1, ref, atol=1e-15)


# ------------------------------------------------------------ mass between two spaces
@pytest.mark.parametrize("kv,kw", [(1, 1), (1, 2), (3, 2)])
def test_mixed_mass_element_sums(kv, kw):
    """Partition of unity in both spaces: the entries sum to |T|, row i sums to
    int phi_i (= |T|/4 for P1 rows); P1 x P1 is |T|/20 (1 + delta_ij)."""
    x = np.array([[0.1, 0.0, 0.2], [1.1, 0.3, 0.0], [0.2, 0.9, 0.1], [0.3, 0.2, 1.3]])
    vol = abs(np.linalg.det(x[1:] - x[0])) / 6
    spec = dict(kind="mass", cell=0, degree=kv, trial_degree=kw, bs=1, rank=2,
                qm=max(1, (kv + kw + 2) // 2), kappa=("val", 1.0))
    A = cpu.tabulate(spec, [], [], x)
    assert A.shape == (cpu.lib().or_ndofs(0, kv), cpu.lib().or_ndofs(0, kw))
    assert abs(A.sum() - vol) <= 1e-14
    if kv == 1:
        np.testing.assert_allclose(A.sum(axis=1), vol / 4, rtol=1e-13)
    if kv == 1 and kw == 1:
        np.testing.assert_allclose(A, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-13)


# ------------------------------------------------------------ the reference test suite's forms
def _ref_forms(n=2, k=1, seed=7):
    from mifem.nonlinear import reference_test_forms
    mesh = create_unit_cube(n, jitter=0.2, seed=seed)
    return mesh, reference_test_forms(mesh, k)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_nonlinear_jacobians_are_residual_derivatives(which):
    """J = dF/du (cell, exterior) and dF/du_dg (interior) by central differences."""
    import scipy.sparse as sp
    mesh, F = _ref_forms(2, 1)
    L, a = F["vector"][which], F["matrix"][which]
    state = F["u_dg"] if which == 2 else F["u"]
    rng = np.random.default_rng(9 + which)
    state.x[:] = rng.uniform(-0.5, 0.5, state.x.size)
    x0 = state.x.copy()
    oa = cpu.OracleAssembler(a)
    rp, cols, _, _ = oa.sparsity()
    V = a.function_spaces[0]
    J = sp.csr_matrix((oa.assemble_matrix(rp, cols), cols, rp), shape=(V.size, V.size))
    d = rng.standard_normal(x0.size)
    h = 1e-6
    state.x[:] = x0 + h * d
    fp = cpu.OracleAssembler(L).assemble_vector()
    state.x[:] = x0 - h * d
    fm = cpu.OracleAssembler(L).assemble_vector()
    state.x[:] = x0
    fd = (fp - fm) / (2 * h)
    assert np.linalg.norm(fd - J @ d) <= 1e-7 * max(1.0, np.linalg.norm(J @ d))


def test_nonlinear_residual_sums_closed_forms():
    """Summed over all test functions (partition of unity): the cell residual is
    int exp(u) kappa dx (u = 0.4, kappa = 1: e^0.4 x volume 1); the exterior
    residual with u = y, kappa = z is int y z n_y ds = 1/2 (the face y = 1);
    the interior residual equals the functional's interior term."""
    mesh, F = _ref_forms(3, 1)
    u, kap, udg = F["u"], F["coeff"], F["u_dg"]
    u.x[:] = 0.4
    kap.x[:] = 1.0
    assert abs(cpu.OracleAssembler(F["vector"][0]).assemble_vector().sum() - math.exp(0.4)) <= 1e-13
    u.interpolate(lambda x: x[1])
    kap.interpolate(lambda x: x[2])
    assert abs(cpu.OracleAssembler(F["vector"][1]).assemble_vector().sum() - 0.5) <= 1e-13
    kap.interpolate(lambda x: x[0])
    udg.interpolate(lambda x: 0.3 + 0.6 * x[0] - 0.2 * x[1] + 0.9 * x[2])
    got = cpu.OracleAssembler(F["vector"][2]).assemble_vector().sum()
    ref = cpu.OracleAssembler(_sub(F["scalar"][0], 2)).assemble_scalar()
    assert abs(got - ref) <= 1e-13 * abs(ref)


def test_diagonal_part_is_matrix_diagonal():
    """Assembled diagonal part = diagonal of the assembled matrix (the element
    diagonals are what lands on the global diagonal)."""
    from mifem.nonlinear import diagonal
    mesh, F = _ref_forms(2, 1)
    for a in F["matrix"]:
        oa = cpu.OracleAssembler(a)
        rp, cols, _, _ = oa.sparsity()
        vals = oa.assemble_matrix(rp, cols)
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
        diag = np.zeros(rp.size - 1)
        diag[rows[rows == cols]] = vals[rows == cols]
        got = cpu.OracleAssembler(diagonal(a)).assemble_vector()
        np.testing.assert_allclose(got, diag, atol=1e-14 * max(1.0, abs(diag).max()))


@pytest.mark.parametrize("which", ["elasticity", "dg"])
def test_multithreaded_oracle_matches_serial(which):
    """The coloured multi-core assembly (CPU baseline) equals the serial oracle
    to rounding: same entries, contributions summed colour by colour."""
    P = problems.elasticity_p2(3) if which == "elasticity" else problems.dg_plaplace(3)
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    rp, cols, _, _ = Oa.sparsity()
    ref = Oa.assemble_matrix(rp, cols, P.bcs)
    got = Oa.assemble_matrix_mt(rp, cols, P.bcs, threads=4)
    assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)
    bref, bgot = OL.assemble_vector(), OL.assemble_vector_mt(threads=4)
    assert np.linalg.norm(bgot - bref) <= 1e-14 * np.linalg.norm(bref)
    g = np.random.default_rng(5).uniform(-1, 1, P.V.size)
    for bc in P.bcs:
        bc.values[bc.dofs] = g[bc.dofs]           # nonzero Dirichlet data
    lref = Oa.apply_lifting(bref.copy(), P.bcs, scale=0.7)
    lgot = Oa.apply_lifting(bref.copy(), P.bcs, scale=0.7, threads=4)
    assert np.linalg.norm(lgot - lref) <= 1e-14 * np.linalg.norm(lref)
    assert (np.linalg.norm(lref - bref) > 0) == bool(P.bcs)   # DG: Robin terms only
    for (order, starts, nc), I in zip(Oa.colouring(), P.a.integrals):
        assert starts[-1] == I.num_entities and nc <= 64
        assert np.array_equal(np.sort(order), np.arange(I.num_entities))
    if which == "elasticity":
        # no two cells of one colour share a row dof
        dm = P.V.unrolled_dofmap()
        order, starts, nc = Oa.colouring()[0]
        for c in range(nc):
            d = dm[order[starts[c]:starts[c + 1]]].ravel()
            assert np.unique(d).size == d.size


def test_jacobi_cg_matches_direct_solve():
    """cudolfinx.solve.pcg (its torch sparse branch, on CPU tensors) solves the
    oracle's P1 Poisson system with Dirichlet rows to a direct solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    from cudolfinx.solve import pcg
    P = problems.poisson_p1(6)
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    rp, cols, _, _ = Oa.sparsity()
    vals = Oa.assemble_matrix(rp, cols, P.bcs)
    b = OL.assemble_vector()
    Oa.apply_lifting(b, P.bcs)
    cpu.set_bc(b, P.bcs, P.V)
    S = sp.csr_matrix((vals, cols, rp), shape=(P.V.size, P.V.size))
    M = torch.sparse_csr_tensor(torch.from_numpy(rp.astype(np.int64)), torch.from_numpy(cols.astype(np.int64)),
                                torch.from_numpy(vals), size=S.shape)
    x, it, res = pcg(M, torch.from_numpy(b), rtol=1e-12, maxiter=2000)
    ref = spla.spsolve(S.tocsc(), b)
    assert res < 1e-12 and 0 < it < 2000
    assert np.linalg.norm(x.numpy() - ref) <= 1e-9 * np.linalg.norm(ref)


@pytest.mark.parametrize("k", [1, 2])
def test_sipg_invariant_under_cell_vertex_order(k):
    """The oracle's SIPG energy and residual do not depend on the order in
    which cells list their vertices: the same physical state on the ascending
    mesh and on a mesh with every cell's vertices shuffled (all six facet
    permutations occur) gives the same energy and, node by node (matched by
    cell and physical position), the same residual.  A side of an interior
    facet evaluated at the wrong points would change both."""
    from mifem.dgforms import sipg_energy
    # p = 2: every integrand is a polynomial the rules integrate exactly, so
    # moving the (non-symmetric) cell quadrature points by reordering a cell's
    # vertices changes nothing but round-off
    P0 = problems.dg_plaplace(2, degree=k, p=2.0)
    P1 = problems.dg_plaplace(2, degree=k, p=2.0, shuffle_vertices=13)
    assert len(np.unique(P1.mesh.facet_permutations())) == 6
    # node k of cell c on the shuffled mesh <-> node match[c][k] on the ascending
    # one (same cell, same physical position)
    X0, X1 = P0.V.node_coordinates, P1.V.node_coordinates
    d0, d1 = P0.V.dofmap, P1.V.dofmap
    perm = np.empty(P0.V.size, dtype=np.int64)          # dof of mesh 1 -> dof of mesh 0
    for c in range(d0.shape[0]):
        a, b = X0[d0[c]], X1[d1[c]]
        match = np.argmin(np.linalg.norm(b[:, None] - a[None], axis=2), axis=1)
        assert np.allclose(b, a[match], atol=1e-13)
        perm[d1[c]] = d0[c][match]
    # a discontinuous state (random per cell and node): interior jumps are non-zero
    u0 = np.random.default_rng(17).uniform(-1, 1, P0.V.size)
    P0.extras["u"].x[:] = u0
    P1.extras["u"].x[:] = u0[perm]
    E = [cpu.OracleAssembler(sipg_energy(P.V, P.extras["u"], p=2.0)).assemble_scalar()
         for P in (P0, P1)]
    F0 = cpu.OracleAssembler(P0.L).assemble_vector()
    F1 = cpu.OracleAssembler(P1.L).assemble_vector()
    assert abs(E[0] - E[1]) <= 1e-12 * abs(E[0])
    assert np.abs(F1 - F0[perm]).max() <= 1e-12 * np.abs(F0).max()
    # the interior-facet part is not trivially zero for this state
    inter = [I for I in sipg_energy(P0.V, P0.extras["u"], p=2.0).integrals
             if I.type == "interior_facet"][0]
    from mifem.forms import Form
    Ei = cpu.OracleAssembler(Form(0, [], [inter], [P0.extras["u"]], [], "int",
                                  domain=P0.mesh)).assemble_scalar()
    assert abs(Ei) > 1e-3 * abs(E[0])


# ------------------------------------------------------------ triangles (the reference suite's domain)
@pytest.mark.parametrize("k", [1, 2, 3])
def test_triangle_basis_nodal_partition_and_gradients(k):
    """P_k on the reference triangle: nodal at the Basix-ordered nodes of the
    engine's element, a partition of unity, gradients summing to zero, and
    the interpolant of a degree-k polynomial reproducing it and its gradient."""
    el = E.Element("P", "triangle", k)
    X = np.hstack([el.nodes, np.zeros((el.ndofs, 1))])
    V = np.array([cpu.basis(2, k, x)[0] for x in X])
    np.testing.assert_allclose(V, np.eye(el.ndofs), atol=1e-14)
    f = lambda x: x[..., 0] ** k - 2 * x[..., 0] * x[..., 1] ** (k - 1) + 0.5 * x[..., 1]
    vals = f(X)
    for p in np.random.default_rng(4).random((7, 2)) * 0.45:
        p3 = np.array([p[0], p[1], 0.0])
        phi, dphi = cpu.basis(2, k, p3)
        assert abs(phi.sum() - 1) < 1e-14
        np.testing.assert_allclose(dphi.sum(axis=0), 0, atol=1e-12)
        gf = np.array([k * p[0] ** (k - 1) - 2 * p[1] ** (k - 1),
                       -2 * (k - 1) * p[0] * p[1] ** max(k - 2, 0) + 0.5, 0.0])
        np.testing.assert_allclose(vals @ dphi, gf, atol=1e-11)
        assert abs(vals @ phi - f(p3)) < 1e-12


def test_line_rule_exactness():
    """The m-point Gauss rule on [0, 1] (triangle edges) integrates t^a exactly
    for a <= 2m - 1."""
    for m in range(1, 6):
        t, w = cpu.gauss_jacobi(m, 0.0)
        for a in range(2 * m):
            assert abs(np.sum(w * t ** a) - 1.0 / (a + 1)) < 1e-14


def _tri_forms(nx=5, ny=7, k=1):
    from mifem.mesh import create_unit_square
    from mifem.nonlinear import reference_test_forms
    mesh = create_unit_square(nx, ny)
    return mesh, reference_test_forms(mesh, k)


def test_triangle_p1_stiffness_closed_form():
    """The cell Jacobian with kappa = 0 is the P1 stiffness matrix |T| G G^T
    (G: barycentric gradients of the physical triangle)."""
    rng = np.random.default_rng(3)
    x = np.hstack([rng.random((3, 2)), np.zeros((3, 1))])
    area = 0.5 * abs(np.linalg.det(x[1:, :2] - x[0, :2]))
    G = np.linalg.inv(np.hstack([np.ones((3, 1)), x[:, :2]]))[1:].T      # (3, 2)
    spec = dict(kind="nl_cell", cell=2, degree=1, bs=1, rank=2, qm=3, u_off=0, kappa_off=3)
    A = cpu.tabulate(spec, np.zeros(6), [], x)
    np.testing.assert_allclose(A, area * G @ G.T, rtol=1e-12, atol=1e-14)


def test_triangle_reference_forms_closed_forms():
    """On the unit square of triangles, summed over all test functions (a
    partition of unity): cell residual = e^0.4 (u = 0.4, kappa = 1); exterior
    residual = int u kappa (grad u . n) ds = 1/2 (u = y, kappa = x: the edge
    y = 1 only); interior residual = the functional's interior term.  The
    functional: e^0.4 for constant u; int_dOmega y ds = 2; int avg(v^2)
    avg(kappa) over interior edges with v = kappa = 1 = their total length."""
    mesh, F = _tri_forms()
    u, kap, udg = F["u"], F["coeff"], F["u_dg"]
    u.x[:] = 0.4
    kap.x[:] = 1.0
    assert abs(cpu.OracleAssembler(F["vector"][0]).assemble_vector().sum() - math.exp(0.4)) <= 1e-13
    assert abs(cpu.OracleAssembler(_sub(F["scalar"][0], 0)).assemble_scalar() - math.exp(0.4)) <= 1e-13
    u.interpolate(lambda x: x[1])
    assert abs(cpu.OracleAssembler(_sub(F["scalar"][0], 1)).assemble_scalar() - 2.0) <= 1e-13
    kap.interpolate(lambda x: x[0])
    assert abs(cpu.OracleAssembler(F["vector"][1]).assemble_vector().sum() - 0.5) <= 1e-13
    udg.interpolate(lambda x: 0.3 + 0.6 * x[0] - 0.2 * x[1])
    got = cpu.OracleAssembler(F["vector"][2]).assemble_vector().sum()
    ref = cpu.OracleAssembler(_sub(F["scalar"][0], 2)).assemble_scalar()
    assert abs(got - ref) <= 1e-13 * abs(ref)
    kap.x[:] = 1.0
    udg.x[:] = 1.0
    ft = mesh.facet_table
    length = sum(np.linalg.norm(np.subtract(*mesh.x[mesh.cells[c0][list(ft[l0])]]))
                 for c0, l0, _, _ in mesh.interior_facets)
    assert abs(cpu.OracleAssembler(_sub(F["scalar"][0], 2)).assemble_scalar() - length) <= 1e-12 * length


def test_triangle_functional_cell_term_closed_form():
    """int_T e^l for linear l: 2 |T| sum_i e^{l_i} / prod_{j != i} (l_i - l_j)."""
    mesh, F = _tri_forms(4, 3)
    a = np.array([0.3, -0.5, 0.0])
    F["u"].interpolate(lambda x: 0.1 + a @ x)
    F["coeff"].x[:] = 1.0
    got = cpu.OracleAssembler(_sub(F["scalar"][0], 0)).assemble_scalar()
    ref = 0.0
    for c in mesh.cells:
        X = mesh.x[c]
        area = 0.5 * abs(np.linalg.det(X[1:, :2] - X[0, :2]))
        l = 0.1 + X @ a
        ref += 2 * area * sum(np.exp(l[i]) / np.prod([l[i] - l[j] for j in range(3) if j != i])
                              for i in range(3))
    assert abs(got - ref) <= 1e-7 * abs(ref), (got, ref)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_triangle_jacobians_are_residual_derivatives(which):
    """J = dF/du (cell, exterior) and dF/du_dg (interior) by central differences, on triangles."""
    import scipy.sparse as sp
    mesh, F = _tri_forms(4, 5)
    L, a = F["vector"][which], F["matrix"][which]
    state = F["u_dg"] if which == 2 else F["u"]
    rng = np.random.default_rng(19 + which)
    state.x[:] = rng.uniform(-0.5, 0.5, state.x.size)
    x0 = state.x.copy()
    oa = cpu.OracleAssembler(a)
    rp, cols, _, _ = oa.sparsity()
    V = a.function_spaces[0]
    J = sp.csr_matrix((oa.assemble_matrix(rp, cols), cols, rp), shape=(V.size, V.size))
    d = rng.standard_normal(x0.size)
    h = 1e-6
    state.x[:] = x0 + h * d
    fp = cpu.OracleAssembler(L).assemble_vector()
    state.x[:] = x0 - h * d
    fm = cpu.OracleAssembler(L).assemble_vector()
    state.x[:] = x0
    assert np.linalg.norm((fp - fm) / (2 * h) - J @ d) <= 1e-7 * max(1.0, np.linalg.norm(J @ d))


def test_sipg_default_penalty_is_coercive():
    """The Jacobian of the SIPG p-Laplacian at u = 0 (the linear SIPG operator
    with the Robin term) is positive definite with the default penalty 30 k^2
    on the jittered 12^3 x 6 mesh, and indefinite with 10 k^2 -- the penalty
    with which Newton stalled once the iterate's gradients became small."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from mifem.dgforms import sipg_plaplace
    P = problems.dg_plaplace(12)
    u = P.extras["u"]
    u.x[:] = 0.0
    negative = {}
    for pen in (None, 10.0):
        a, _ = sipg_plaplace(P.V, u, f=1.0, alpha=2.0, p=3.0, eps=1e-3, penalty=pen)
        O = cpu.OracleAssembler(a)
        rp, cols, _, _ = O.sparsity()
        n = P.V.size
        A = sp.csr_matrix((O.assemble_matrix(rp, cols), cols, rp), shape=(n, n))
        assert abs(A - A.T).max() <= 1e-13 * abs(A).max()          # symmetric
        # inertia (Sylvester): a symmetric permutation P A P^T = L D L^T, diagonal
        # pivots only; the number of negative pivots is the number of negative
        # eigenvalues
        lu = spla.splu(A.tocsc(), permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                       options=dict(SymmetricMode=True))
        assert np.array_equal(lu.perm_r, lu.perm_c)
        negative[pen] = int((lu.U.diagonal() < 0).sum())
    assert negative[None] == 0 and negative[10.0] > 0, negative
