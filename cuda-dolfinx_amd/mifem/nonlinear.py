"""The nonlinear residual / Jacobian forms the reference's assembly tests use
(python/test/test_cuda_assembly.py:48-91), on affine tetrahedra or triangles
(the reference's own domain, a unit square of triangles), and the
"diagonal part" of a bilinear form (FFCx form_compiler_options
{"part": "diagonal"}, test_cuda_assembly.py:200-209).

With u, kappa in a continuous P_k space V, u_dg in DG_k and n the outward
unit normal:

  cell residual      F_c(p)  = int (exp(u) kappa p + grad u . grad p) dx
  exterior residual  F_e(p)  = int_ds u kappa p (grad u . n)
  interior residual  F_i(q)  = int_dS avg(q) avg(kappa) avg(u_dg^2)      (q in DG_k)
  Jacobians          J_c = dF_c/du, J_e = dF_e/du, J_i = dF_i/du_dg

Quadrature: collapsed Gauss-Jacobi rules with m points per direction
(Gauss-Legendre with m points on triangle edges),
cells m = k + 2 (the exp term is not polynomial), exterior facets
m = (4k + 1) // 2 (exact for the degree 4k - 1 integrand), interior facets
m = (3 k_dg + k + 2) // 2 (degree 3 k_dg + k) -- the rules the oracle
evaluates.
Element bodies use the UFCx layout: coefficients [u | kappa] on cell and
exterior integrals, [kappa | u_dg] interlaced per coefficient as (cell 0,
cell 1) on interior facets.
"""
from __future__ import annotations

import numpy as np

from .dgforms import _QPERM
from .fem import Function, FunctionSpace, dirichletbc, functionspace
from .forms import CELL, EXTERIOR_FACET, INTERIOR_FACET, Form, Integral, _SIG, _carr, _uid
from .simplex import Kit


def _cell_forms(V: FunctionSpace, u: Function, kappa: Function, rank: int):
    el, nd, k = V.element, V.element.ndofs, V.element.degree
    kit = Kit(V.mesh.cell_type)
    D = kit.tdim
    m = k + 2
    q = kit.cell_rule(m)
    phi, dphi = kit.cell_tables(el, q)
    if rank == 2:
        loop = f"""    for (int i = 0; i < {nd}; ++i)
      for (int j = 0; j < {nd}; ++j)
        A[i * {nd} + j] += fw * (ek * FE_phi[iq][j] * FE_phi[iq][i]
                                 + {kit.dot("g[j][{}]", "g[i][{}]")});
"""
    else:
        loop = f"""    for (int i = 0; i < {nd}; ++i)
      A[i] += fw * (ek * FE_phi[iq][i] + {kit.dot("gu[{}]", "g[i][{}]")});
"""
    name = f"tabulate_tensor_integral_{_uid('nl_cell', rank, k, D)}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phi) + _carr("FE_dphi", dphi)
            + kit.geom("", "coordinate_dofs") + f"""
  (void)h;
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
{kit.grad("", "FE_dphi[iq]", nd)}    double uq = 0.0, kq = 0.0, gu[{D}] = {{0.0}};
    for (int i = 0; i < {nd}; ++i) {{
      uq += w[i] * FE_phi[iq][i];
      kq += w[{nd} + i] * FE_phi[iq][i];
      for (int d = 0; d < {D}; ++d) gu[d] += w[i] * g[i][d];
    }}
    const double ek = exp(uq) * kq;
    const double fw = FE_w[iq] * adet;
{loop}  }}
""")
    spec = dict(kind="nl_cell", cell=kit.oracle_cell, degree=k, bs=1, rank=rank, qm=m, u_off=0,
                kappa_off=nd)
    return Integral(CELL, -1, np.arange(V.mesh.num_cells, dtype=np.int32), name,
                    _SIG.format(name=name) + "\n{\n" + body + "}\n", None, spec)


def _ext_forms(V: FunctionSpace, rank: int):
    el, nd, k = V.element, V.element.ndofs, V.element.degree
    kit = Kit(V.mesh.cell_type)
    D = kit.tdim
    m = (4 * k + 1) // 2
    q = kit.facet_rule(m)
    phi, dphi = kit.facet_tables(el, q)
    if rank == 2:
        loop = f"""    for (int i = 0; i < {nd}; ++i)
      for (int j = 0; j < {nd}; ++j) {{
        const double gjn = {kit.normal_dot("g[j][{}]")};
        A[i * {nd} + j] += fw * kq * ph[i] * (ph[j] * gun + uq * gjn);
      }}
"""
    else:
        loop = f"""    for (int i = 0; i < {nd}; ++i) A[i] += fw * uq * kq * ph[i] * gun;
"""
    name = f"tabulate_tensor_integral_{_uid('nl_ext', rank, k, D)}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phi) + _carr("FE_dphi", dphi) + kit.FACET
            + kit.geom("", "coordinate_dofs") + f"""
  (void)h; (void)adet;
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
    const double* ph = FE_phi[lf0][iq];
{kit.grad("", "FE_dphi[lf0][iq]", nd)}    double uq = 0.0, kq = 0.0, gu[{D}] = {{0.0}};
    for (int i = 0; i < {nd}; ++i) {{
      uq += w[i] * ph[i];
      kq += w[{nd} + i] * ph[i];
      for (int d = 0; d < {D}; ++d) gu[d] += w[i] * g[i][d];
    }}
    const double gun = {kit.normal_dot("gu[{}]")};
    const double fw = FE_w[iq] * ds;
{loop}  }}
""")
    spec = dict(kind="nl_ext", cell=kit.oracle_cell, degree=k, bs=1, rank=rank, qm=m, u_off=0,
                kappa_off=nd)
    ent = np.ascontiguousarray(V.mesh.exterior_facets.reshape(-1), dtype=np.int32)
    return Integral(EXTERIOR_FACET, -1, ent, name, _SIG.format(name=name) + "\n{\n" + body + "}\n",
                    None, spec)


def _int_forms(Vc: FunctionSpace, Vd: FunctionSpace, rank: int):
    """Interior residual / Jacobian on the DG space Vd; kappa on the continuous Vc."""
    kc, kd = Vc.element.degree, Vd.element.degree
    nc, ndg = Vc.element.ndofs, Vd.element.ndofs
    kit = Kit(Vd.mesh.cell_type)
    m = (3 * kd + kc + 2) // 2
    q = kit.facet_rule(m)
    phc = kit.facet_tables_permuted(Vc.element, q)[0]
    phd = kit.facet_tables_permuted(Vd.element, q)[0]
    ko, uo = 0, 2 * nc                                 # interlaced: k0 k1 | u0 u1
    n2 = 2 * ndg
    if rank == 2:
        loop = f"""    for (int I = 0; I < {n2}; ++I) {{
      const double pI = 0.5 * (I < {ndg} ? FE_psi[lf0][qp0][iq][I] : FE_psi[lf1][qp1][iq][I - {ndg}]);
      for (int J = 0; J < {n2}; ++J) {{
        const double uJ = J < {ndg} ? ua * FE_psi[lf0][qp0][iq][J] : ub * FE_psi[lf1][qp1][iq][J - {ndg}];
        A[I * {n2} + J] += fw * pI * kav * uJ;
      }}
    }}
"""
    else:
        loop = f"""    const double uav = 0.5 * (ua * ua + ub * ub);
    for (int I = 0; I < {n2}; ++I) {{
      const double pI = 0.5 * (I < {ndg} ? FE_psi[lf0][qp0][iq][I] : FE_psi[lf1][qp1][iq][I - {ndg}]);
      A[I] += fw * pI * kav * uav;
    }}
"""
    name = f"tabulate_tensor_integral_{_uid('nl_int', rank, kc, kd, kit.tdim)}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phc) + _carr("FE_psi", phd) + kit.FACET
            + "  const int lf1 = entity_local_index[1];\n" + _QPERM + f"""
  (void)nx; (void)ny;
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
    double ka = 0.0, kb = 0.0, ua = 0.0, ub = 0.0;
    for (int i = 0; i < {nc}; ++i) {{
      ka += w[{ko} + i] * FE_phi[lf0][qp0][iq][i];
      kb += w[{ko + nc} + i] * FE_phi[lf1][qp1][iq][i];
    }}
    for (int i = 0; i < {ndg}; ++i) {{
      ua += w[{uo} + i] * FE_psi[lf0][qp0][iq][i];
      ub += w[{uo + ndg} + i] * FE_psi[lf1][qp1][iq][i];
    }}
    const double kav = 0.5 * (ka + kb);
    const double fw = FE_w[iq] * ds;
{loop}  }}
""")
    spec = dict(kind="nl_int", cell=kit.oracle_cell, degree=kd, bs=1, rank=rank, qm=m, kappa_off=0,
                v_off=nc, v_degree=kd, kappa_degree=kc)
    ent = np.ascontiguousarray(Vd.mesh.interior_facets.reshape(-1), dtype=np.int32)
    return Integral(INTERIOR_FACET, -1, ent, name, _SIG.format(name=name) + "\n{\n" + body + "}\n",
                    None, spec)


def reference_test_forms(mesh, k: int = 1):
    """The reference test's forms (test_cuda_assembly.py:48-91) on a tetrahedral
    or triangular mesh: a dict with "coeff", "u", "u_dg", "bcs", "scalar",
    "vector", "matrix"."""
    from .functionals import mixed_functional
    V = functionspace(mesh, ("P", k))
    Vd = functionspace(mesh, ("DG", k))
    u, u_dg, kappa = Function(V, "u"), Function(Vd, "u_dg"), Function(V, "kappa")
    u.interpolate(lambda x: x[0] ** 2 + x[1])
    u_dg.interpolate(lambda x: x[0] ** 2 + x[1])
    kappa.interpolate(lambda x: np.sin(x[0]) * np.cos(x[1]))
    f = Function(V, "f")
    f.interpolate(lambda x: x[0] + x[1])
    dofs = np.flatnonzero(np.isclose(V.node_coordinates[:, 0], 0.0)).astype(np.int32)
    bc = dirichletbc(f, dofs, V)
    vec = [Form(1, [V], [_cell_forms(V, u, kappa, 1)], [u, kappa], [], "nl_cell_residual"),
           Form(1, [V], [_ext_forms(V, 1)], [u, kappa], [], "nl_exterior_residual"),
           Form(1, [Vd], [_int_forms(V, Vd, 1)], [kappa, u_dg], [], "nl_interior_residual")]
    mat = [Form(2, [V, V], [_cell_forms(V, u, kappa, 2)], [u, kappa], [], "nl_cell_jacobian"),
           Form(2, [V, V], [_ext_forms(V, 2)], [u, kappa], [], "nl_exterior_jacobian"),
           Form(2, [Vd, Vd], [_int_forms(V, Vd, 2)], [kappa, u_dg], [], "nl_interior_jacobian")]
    return {"coeff": kappa, "u": u, "u_dg": u_dg, "bcs": [bc],
            "scalar": [mixed_functional(u, kappa, u_dg)], "vector": vec, "matrix": mat}


def diagonal(a: Form) -> Form:
    """The diagonal part of a square bilinear form as a linear form: entry i is
    the element matrix's (i, i) entry (FFCx "part": "diagonal")."""
    if a.rank != 2 or a.function_spaces[0] is not a.function_spaces[1]:
        raise ValueError("the diagonal part needs a square bilinear form")
    V = a.function_spaces[0]
    nd = V.element.ndofs * V.bs
    ints = []
    for I in a.integrals:
        n = 2 * nd if I.type == INTERIOR_FACET else nd
        name = f"tabulate_tensor_integral_{_uid('diag', I.name)}"
        src = (I.source + "\n" + _SIG.format(name=name) + f"""
{{
  double T[{n * n}];
  for (int k = 0; k < {n * n}; ++k) T[k] = 0.0;
  {I.name}(T, w, c, coordinate_dofs, entity_local_index, quadrature_permutation);
  for (int i = 0; i < {n}; ++i) A[i] += T[i * {n} + i];
}}
""")
        spec = dict(I.spec, rank=1, diag_part=1)
        ints.append(Integral(I.type, I.id, I.entities, name, src, None, spec))
    return Form(1, [V], ints, a.coefficients, a.constants, a.description + "_diagonal")
