"""Functionals (rank-0 forms) over cells, exterior and interior facets.

`mixed_functional(u, kappa, v)` is the functional the reference's assembly
test integrates (python/test/test_cuda_assembly.py:66-71):

    M = int_Omega exp(u) kappa dx + int_dOmega u kappa ds
      + sum_{F interior} int_F avg(v^2) avg(kappa) dS

with u, kappa on one continuous P_k space and v on a DG space, all on affine
tetrahedra or triangles.  Quadrature degrees follow the usual estimate (exp adds two to
its argument's degree): 2k + 2 on cells, 2k on exterior facets and
2 k_v + k on interior facets.  The element bodies use the UFCx argument
layout: w = [u | kappa | v] per cell, interlaced per coefficient as
(cell 0, cell 1) on interior facets.
"""
from __future__ import annotations

import numpy as np

from .fem import Function
from .forms import CELL, EXTERIOR_FACET, INTERIOR_FACET, Form, Integral, _SIG, _carr, _uid
from .dgforms import _QPERM
from .simplex import Kit


def mixed_functional(u: Function, kappa: Function, v: Function) -> Form:
    Vp, Vd = u.function_space, v.function_space
    mesh = Vp.mesh
    if kappa.function_space is not Vp:
        raise ValueError("u and kappa must share one function space")
    if Vp.bs != 1 or Vd.bs != 1 or Vd.mesh is not mesh:
        raise NotImplementedError("scalar spaces on one mesh")
    kit = Kit(mesh.cell_type)
    k, kv = Vp.element.degree, Vd.element.degree
    np_, nv = Vp.element.ndofs, Vd.element.ndofs
    mc = k + 2                              # collapsed rule, exact to degree 2k + 3
    me = k + 1                              # exact to degree 2k + 1
    mi = (2 * kv + k + 2) // 2              # exact to degree 2 kv + k
    qc, qe, qi = kit.cell_rule(mc), kit.facet_rule(me), kit.facet_rule(mi)
    spec0 = dict(cell=kit.oracle_cell, degree=k, bs=1, u_off=0, kappa_off=np_, v_off=2 * np_,
                 v_degree=kv)

    # cells: exp(u) kappa
    name = f"tabulate_tensor_integral_{_uid('fn_cell', k, kv, kit.tdim)}"
    body = (_carr("FE_w", qc.weights) + _carr("FE_phi", Vp.element.tabulate(qc.points, 0)[0])
            + kit.geom("", "coordinate_dofs") + f"""
  (void)h; (void)K;
  for (int iq = 0; iq < {qc.weights.size}; ++iq) {{
    double uq = 0.0, kq = 0.0;
    for (int i = 0; i < {np_}; ++i) {{ uq += w[i] * FE_phi[iq][i]; kq += w[{np_} + i] * FE_phi[iq][i]; }}
    A[0] += FE_w[iq] * adet * exp(uq) * kq;
  }}
""")
    icell = Integral(CELL, -1, np.arange(mesh.num_cells, dtype=np.int32), name,
                     _SIG.format(name=name) + "\n{\n" + body + "}\n", None,
                     dict(kind="fn_cell", rank=0, qm=mc, **spec0))

    # exterior facets: u kappa
    name = f"tabulate_tensor_integral_{_uid('fn_ext', k, kv, kit.tdim)}"
    body = (_carr("FE_w", qe.weights) + _carr("FE_phi", kit.facet_tables(Vp.element, qe)[0]) + kit.FACET
            + f"""
  (void)nx; (void)ny;"""
            + f"""
  for (int iq = 0; iq < {qe.weights.size}; ++iq) {{
    const double* ph = FE_phi[lf0][iq];
    double uq = 0.0, kq = 0.0;
    for (int i = 0; i < {np_}; ++i) {{ uq += w[i] * ph[i]; kq += w[{np_} + i] * ph[i]; }}
    A[0] += FE_w[iq] * ds * uq * kq;
  }}
""")
    iext = Integral(EXTERIOR_FACET, -1,
                    np.ascontiguousarray(mesh.exterior_facets.reshape(-1), dtype=np.int32), name,
                    _SIG.format(name=name) + "\n{\n" + body + "}\n", None,
                    dict(kind="fn_ext", rank=0, qm=me, **spec0))

    # interior facets: avg(v^2) avg(kappa); interlaced w: u0 u1 | k0 k1 | v0 v1
    name = f"tabulate_tensor_integral_{_uid('fn_int', k, kv, kit.tdim)}"
    ko, vo = 2 * np_, 4 * np_
    body = (_carr("FE_w", qi.weights)
            + _carr("FE_phi", kit.facet_tables_permuted(Vp.element, qi)[0])
            + _carr("FE_psi", kit.facet_tables_permuted(Vd.element, qi)[0]) + kit.FACET
            + "  const int lf1 = entity_local_index[1];\n" + _QPERM + f"""
  (void)nx; (void)ny;"""
            + f"""
  for (int iq = 0; iq < {qi.weights.size}; ++iq) {{
    const double* pa = FE_phi[lf0][qp0][iq];
    const double* pb = FE_phi[lf1][qp1][iq];
    const double* qa = FE_psi[lf0][qp0][iq];
    const double* qb = FE_psi[lf1][qp1][iq];
    double ka = 0.0, kb = 0.0, va = 0.0, vb = 0.0;
    for (int i = 0; i < {np_}; ++i) {{
      ka += w[{ko} + i] * pa[i];
      kb += w[{ko + np_} + i] * pb[i];
    }}
    for (int i = 0; i < {nv}; ++i) {{
      va += w[{vo} + i] * qa[i];
      vb += w[{vo + nv} + i] * qb[i];
    }}
    A[0] += FE_w[iq] * ds * 0.5 * (va * va + vb * vb) * 0.5 * (ka + kb);
  }}
""")
    iint = Integral(INTERIOR_FACET, -1,
                    np.ascontiguousarray(mesh.interior_facets.reshape(-1), dtype=np.int32), name,
                    _SIG.format(name=name) + "\n{\n" + body + "}\n", None,
                    dict(kind="fn_int", rank=0, qm=mi, **spec0))
    return Form(0, [], [icell, iext, iint], [u, kappa, v], [], "mixed_functional", domain=mesh)
