"""Discontinuous Galerkin forms: symmetric interior penalty (SIPG) discretisation
of a p-Laplacian with a Robin boundary, on affine tetrahedra.

Residual F(u; v) and its Jacobian J(u; du, v) for a DG_k state u:

  F = sum_K int_K mu(grad u) grad u . grad v - f v dx
    + sum_{F interior} int_F -{grad u}.n [v] - {grad v}.n [u] + (sigma / h_F) [u][v] ds
    + sum_{F boundary} int_F alpha u v ds
  J = sum_K int_K mu grad du . grad v + mu' (grad u . grad du)(grad u . grad v) dx
    + (the facet terms with du in place of u)

with mu = mu0 + (eps^2 + |grad u|^2)^((p-2)/2), mu' = (p-2)(eps^2 + |grad u|^2)^((p-4)/2),
n the unit normal pointing out of the first cell of the facet, [w] = w+ - w-,
{q} = (q+ + q-)/2, sigma = penalty (30 k^2 by default), h_F the mean of the
two cells' diameters (largest vertex distance).  The facet terms are those
of the linear Laplacian, so J is exactly the derivative of F; the constant
part mu0 (default 1) of the coefficient keeps the cell term at least as
strong as the Laplacian consistency terms, so the SIPG form stays coercive
where |grad u| vanishes (with mu0 = 0 and a degenerate coefficient the
linear facet terms make the Jacobian indefinite there).

Interior-facet tensors follow the UFCx layout the reference assembles
(cpp/cudolfinx/fem/CUDAFormIntegral.cpp:1988-2103): rows and columns
[cell0 dofs, cell1 dofs], coefficients interlaced (w_cell0, w_cell1)
(:1855-1874), coordinate_dofs [2][4][3], entity_local_index {lf0, lf1}.
Interior-facet bodies read their facet tables through the facet quadrature
permutations (quadrature_permutation = {perm of (cell 0, lf0), perm of
(cell 1, lf1)}, mesh.Mesh.facet_permutations), so the two cells evaluate at
the same physical points whatever order they list the facet's vertices in.
"""
from __future__ import annotations

import numpy as np

from . import element as E
from .fem import Constant, Function, FunctionSpace
from .forms import (CELL, EXTERIOR_FACET, INTERIOR_FACET, Form, Integral, _SIG, _carr, _uid)


def _geom(s: str, xd: str) -> str:
    """Affine map of one tetrahedron: K = J^-1 entries K{r}{c}{s}, |det J| adet{s}, diameter h{s}."""
    return f"""
  const double J00{s} = {xd}[3] - {xd}[0], J01{s} = {xd}[6] - {xd}[0], J02{s} = {xd}[9] - {xd}[0];
  const double J10{s} = {xd}[4] - {xd}[1], J11{s} = {xd}[7] - {xd}[1], J12{s} = {xd}[10] - {xd}[1];
  const double J20{s} = {xd}[5] - {xd}[2], J21{s} = {xd}[8] - {xd}[2], J22{s} = {xd}[11] - {xd}[2];
  const double m00{s} = J11{s} * J22{s} - J12{s} * J21{s};
  const double m01{s} = J02{s} * J21{s} - J01{s} * J22{s};
  const double m02{s} = J01{s} * J12{s} - J02{s} * J11{s};
  const double det{s} = J00{s} * m00{s} + J10{s} * m01{s} + J20{s} * m02{s};
  const double id{s} = 1.0 / det{s};
  const double K{s}[3][3] = {{{{m00{s} * id{s}, m01{s} * id{s}, m02{s} * id{s}}},
    {{(J12{s} * J20{s} - J10{s} * J22{s}) * id{s}, (J00{s} * J22{s} - J02{s} * J20{s}) * id{s}, (J02{s} * J10{s} - J00{s} * J12{s}) * id{s}}},
    {{(J10{s} * J21{s} - J11{s} * J20{s}) * id{s}, (J01{s} * J20{s} - J00{s} * J21{s}) * id{s}, (J00{s} * J11{s} - J01{s} * J10{s}) * id{s}}}}};
  const double adet{s} = fabs(det{s});
  /* diameter = longest edge: one sqrt of the largest squared length (sqrt is
     monotonic, so this equals the largest of the six edge lengths) */
  double hh{s} = 0.0;
""" + "".join(f"""  {{
    const double dx = {xd}[{3 * a}] - {xd}[{3 * b}], dy = {xd}[{3 * a + 1}] - {xd}[{3 * b + 1}], dz = {xd}[{3 * a + 2}] - {xd}[{3 * b + 2}];
    const double dd = dx * dx + dy * dy + dz * dz;
    hh{s} = dd > hh{s} ? dd : hh{s};
  }}
""" for a in range(4) for b in range(a + 1, 4)) + f"""  const double h{s} = sqrt(hh{s});
"""


_FACET = """
  /* facet lf0 of the first cell has vertices (1,2,3), (0,2,3), (0,1,3) or
     (0,1,2): picked by selects, so coordinate_dofs is only read at constant
     offsets and stays in registers */
  const int lf0 = entity_local_index[0];
  const double* xa = coordinate_dofs;
  double P0[3], P1[3], P2[3], PO[3];
  for (int d = 0; d < 3; ++d) {
    const double v0 = xa[d], v1 = xa[3 + d], v2 = xa[6 + d], v3 = xa[9 + d];
    P0[d] = lf0 == 0 ? v1 : v0;
    P1[d] = lf0 <= 1 ? v2 : v1;
    P2[d] = lf0 <= 2 ? v3 : v2;
    PO[d] = lf0 == 0 ? v0 : (lf0 == 1 ? v1 : (lf0 == 2 ? v2 : v3));
  }
  const double e1x = P1[0] - P0[0], e1y = P1[1] - P0[1], e1z = P1[2] - P0[2];
  const double e2x = P2[0] - P0[0], e2y = P2[1] - P0[1], e2z = P2[2] - P0[2];
  double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double ds = sqrt(nx * nx + ny * ny + nz * nz);
  {
    const double ids = 1.0 / ds;          /* one division for the three components */
    nx *= ids; ny *= ids; nz *= ids;
  }
  {
    /* outward from the first cell: away from the vertex opposite the facet */
    const double ox = PO[0] - P0[0], oy = PO[1] - P0[1], oz = PO[2] - P0[2];
    if (nx * ox + ny * oy + nz * oz > 0.0) { nx = -nx; ny = -ny; nz = -nz; }
  }
"""


# interior facets: the facet permutation of each side (0 without an array)
_QPERM = """  const int qp0 = quadrature_permutation ? quadrature_permutation[0] : 0;
  const int qp1 = quadrature_permutation ? quadrature_permutation[1] : 0;
"""


def _grad(s: str, tab: str, nd: int) -> str:
    """g{s}[i][d] = sum_r tab[i][r] K{s}[r][d] (physical gradients of the basis)."""
    return f"""    double g{s}[{nd}][3];
    for (int i = 0; i < {nd}; ++i)
      for (int d = 0; d < 3; ++d)
        g{s}[i][d] = {tab}[i][0] * K{s}[0][d] + {tab}[i][1] * K{s}[1][d] + {tab}[i][2] * K{s}[2][d];
"""


def sipg_plaplace(V: FunctionSpace, u: Function, f: float = 1.0, alpha: float = 2.0,
                  penalty: float | None = None, p: float = 3.0, eps: float = 1e-3,
                  mu0: float = 1.0):
    """(J, F): Jacobian and residual forms of the SIPG p-Laplacian (module doc)."""
    cell, ext, inter = _sipg_parts(V, u, f, alpha, penalty, p, eps, mu0)
    J = Form(2, [V, V], [cell(2), ext(2), inter(2)], [u], [], "sipg_plaplace_J")
    F = Form(1, [V], [cell(1), ext(1), inter(1)], [u], [], "sipg_plaplace_F")
    return J, F


def sipg_energy(V: FunctionSpace, u: Function, f: float = 1.0, alpha: float = 2.0,
                penalty: float | None = None, p: float = 3.0, eps: float = 1e-3,
                mu0: float = 1.0) -> Form:
    """E(u), the functional whose first variation is the residual F of
    `sipg_plaplace` (same parameters):

      E = sum_K int_K (eps^2 + |grad u|^2)^(p/2) / p + mu0 |grad u|^2 / 2 - f u dx
        + sum_{F interior} int_F (sigma / h_F) [u]^2 / 2 - {grad u}.n [u] ds
        + sum_{F boundary} int_F alpha u^2 / 2 ds

    Newton with a line search on E uses it as the merit function.
    """
    cell, ext, inter = _sipg_parts(V, u, f, alpha, penalty, p, eps, mu0)
    return Form(0, [], [cell(0), ext(0), inter(0)], [u], [], "sipg_energy", domain=V.mesh)


def _sipg_parts(V, u, f, alpha, penalty, p, eps, mu0):
    """Builders cell(rank), ext(rank), inter(rank) of the SIPG integrals."""
    mesh = V.mesh
    el = V.element
    if mesh.cell_type != "tetrahedron" or not el.discontinuous or V.bs != 1:
        raise NotImplementedError("SIPG forms need a scalar DG space on tetrahedra")
    if u.function_space is not V:
        raise ValueError("the state u must live on V")
    k, nd = el.degree, el.ndofs
    if penalty is None:
        # 30 k^2: with 10 k^2 the SIPG part is not coercive on the jittered
        # tetrahedra of the benchmark mesh, the energy's Hessian turns
        # indefinite where the p-term vanishes and Newton stalls
        penalty = 30.0 * k * k
    mc = max(1, k)                       # cell rule: exact for degree 2k - 1 (grad.grad and f v)
    mf = k + 1                           # facet rule: exact for [u][v] (degree 2k)
    qc, qf = E.tet_collapsed(mc), E.tri_collapsed(mf)
    tc = el.tabulate(qc.points)
    phic = tc[0]
    dphic = np.stack([tc[1], tc[2], tc[3]], axis=-1)                 # (nqc, nd, 3)
    phif, dphif = [], []
    for lf in range(4):
        t = el.tabulate(E.tet_facet_points(lf, qf.points))
        phif.append(t[0])
        dphif.append(np.stack([t[1], t[2], t[3]], axis=-1))
    phif, dphif = np.stack(phif), np.stack(dphif)                   # (4, nqf, nd[, 3])
    phip, dphip = E.facet_tables_permuted(el, qf.points)            # (4, 6, nqf, nd[, 3])
    nqc, nqf = qc.weights.size, qf.weights.size
    e2 = repr(float(eps) ** 2)
    e_mu, e_dmu = repr((p - 2.0) / 2.0), repr((p - 4.0) / 2.0)
    common = dict(cell=0, degree=k, bs=1, p=float(p), eps=float(eps), alpha=float(alpha),
                  penalty=float(penalty), fval=float(f), u_off=0, mu0=float(mu0))

    def cell(rank):
        def power(base: str, e: float) -> str:
            # half-integer exponents (p an integer) without pow(): sqrt is one
            # instruction sequence, FP64 pow a long call
            t = 2.0 * e
            if t == int(t) and abs(t) <= 6:
                t = int(t)
                whole = "*".join([base] * (abs(t) // 2))
                half = f"sqrt({base})" if abs(t) % 2 else ""
                prod = " * ".join(x for x in (whole.replace("*", " * "), half) if x) or "1.0"
                return prod if t >= 0 else f"(1.0 / ({prod}))"
            return f"pow({base}, {e!r})"

        name = f"tabulate_tensor_integral_{_uid('dg_cell', rank, k, p, eps, f, mu0)}"
        out = (f"        A[i * {nd} + j] += fw * (mu * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2])\n"
               f"                               + dmu * gui[i] * gui[j]);\n" if rank == 2 else "")
        if rank == 2:
            loop = f"    for (int i = 0; i < {nd}; ++i)\n      for (int j = 0; j < {nd}; ++j)\n{out}"
        elif rank == 1:
            loop = (f"    for (int i = 0; i < {nd}; ++i)\n"
                    f"      A[i] += fw * (mu * gui[i] - {float(f)!r} * FE_phi[iq][i]);\n")
        else:
            loop = (f"    double uq = 0.0;\n    for (int i = 0; i < {nd}; ++i) uq += w[i] * FE_phi[iq][i];\n"
                    f"    A[0] += fw * ({power('s2', float(p) / 2.0)} / {float(p)!r}"
                    f" + {0.5 * float(mu0)!r} * (gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2])"
                    f" - {float(f)!r} * uq);\n")
        body = (_carr("FE_w", qc.weights) + _carr("FE_phi", phic) + _carr("FE_dphi", dphic)
                + _geom("", "coordinate_dofs") + f"""
  (void)h;
  for (int iq = 0; iq < {nqc}; ++iq) {{
{_grad("", "FE_dphi[iq]", nd)}    double gu[3] = {{0.0, 0.0, 0.0}};
    for (int i = 0; i < {nd}; ++i)
      for (int d = 0; d < 3; ++d) gu[d] += w[i] * g[i][d];
    double gui[{nd}];
    for (int i = 0; i < {nd}; ++i) gui[i] = gu[0] * g[i][0] + gu[1] * g[i][1] + gu[2] * g[i][2];
    const double s2 = {e2} + gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2];
    const double mu = {float(mu0)!r} + {power('s2', float(e_mu))};
    const double dmu = {float(p - 2.0)!r} * {power('s2', float(e_dmu))};
    (void)dmu;
    const double fw = FE_w[iq] * adet;
{loop}  }}
""")
        src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
        spec = dict(kind="dg_cell", rank=rank, qm=mc, **common)
        return Integral(CELL, -1, np.arange(mesh.num_cells, dtype=np.int32), name, src, None, spec)

    def ext(rank):
        name = f"tabulate_tensor_integral_{_uid('dg_ext', rank, k, alpha)}"
        if rank == 2:
            loop = (f"    for (int i = 0; i < {nd}; ++i)\n      for (int j = 0; j < {nd}; ++j)\n"
                    f"        A[i * {nd} + j] += fw * ph[i] * ph[j];\n")
        elif rank == 1:
            loop = (f"    double uq = 0.0;\n    for (int j = 0; j < {nd}; ++j) uq += w[j] * ph[j];\n"
                    f"    for (int i = 0; i < {nd}; ++i) A[i] += fw * uq * ph[i];\n")
        else:
            loop = (f"    double uq = 0.0;\n    for (int j = 0; j < {nd}; ++j) uq += w[j] * ph[j];\n"
                    f"    A[0] += 0.5 * fw * uq * uq;\n")
        body = (_carr("FE_w", qf.weights) + _carr("FE_phi", phif) + _FACET + f"""
  for (int iq = 0; iq < {nqf}; ++iq) {{
    const double* ph = FE_phi[lf0][iq];
    const double fw = FE_w[iq] * ds * {float(alpha)!r};
{loop}  }}
""")
        src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
        spec = dict(kind="dg_ext", rank=rank, qm=mf, **common)
        ent = np.ascontiguousarray(mesh.exterior_facets.reshape(-1), dtype=np.int32)
        return Integral(EXTERIOR_FACET, -1, ent, name, src, None, spec)

    def inter(rank):
        name = f"tabulate_tensor_integral_{_uid('dg_int', rank, k, penalty)}"
        n2 = 2 * nd
        if rank == 2:
            loop = (f"    for (int I = 0; I < {n2}; ++I)\n      for (int J = 0; J < {n2}; ++J)\n"
                    f"        A[I * {n2} + J] += fw * (-dn[J] * jmp[I] - dn[I] * jmp[J] + sig * jmp[I] * jmp[J]);\n")
        else:
            loop = (f"    double ju = 0.0, du = 0.0;\n"
                    f"    for (int J = 0; J < {n2}; ++J) {{ ju += jmp[J] * w[J]; du += dn[J] * w[J]; }}\n")
            if rank == 1:
                loop += (f"    for (int I = 0; I < {n2}; ++I)\n"
                         f"      A[I] += fw * (-du * jmp[I] - dn[I] * ju + sig * ju * jmp[I]);\n")
            else:
                loop += "    A[0] += fw * (0.5 * sig * ju * ju - du * ju);\n"
        body = (_carr("FE_w", qf.weights) + _carr("FE_phi", phif) + _carr("FE_dphi", dphif)
                + _carr("FE_phiP", phip) + _carr("FE_dphiP", dphip)
                + _FACET + "  const int lf1 = entity_local_index[1];\n" + _QPERM
                + _geom("0", "coordinate_dofs") + _geom("1", "(coordinate_dofs + 12)") + f"""
  (void)adet0; (void)adet1;
  const double sig = {float(penalty)!r} / (0.5 * (h0 + h1));
  for (int iq = 0; iq < {nqf}; ++iq) {{
    /* each side's tables at the facet points under its permutation (the
       identity tables when there is none) */
    const double* ph0 = qp0 ? FE_phiP[lf0][qp0][iq] : FE_phi[lf0][iq];
    const double* ph1 = qp1 ? FE_phiP[lf1][qp1][iq] : FE_phi[lf1][iq];
    const double (*dp0)[3] = qp0 ? FE_dphiP[lf0][qp0][iq] : FE_dphi[lf0][iq];
    const double (*dp1)[3] = qp1 ? FE_dphiP[lf1][qp1][iq] : FE_dphi[lf1][iq];
{_grad("0", "dp0", nd)}{_grad("1", "dp1", nd)}    double jmp[{n2}], dn[{n2}];
    for (int i = 0; i < {nd}; ++i) {{
      jmp[i] = ph0[i];
      jmp[{nd} + i] = -ph1[i];
      dn[i] = 0.5 * (g0[i][0] * nx + g0[i][1] * ny + g0[i][2] * nz);
      dn[{nd} + i] = 0.5 * (g1[i][0] * nx + g1[i][1] * ny + g1[i][2] * nz);
    }}
    const double fw = FE_w[iq] * ds;
{loop}  }}
""")
        src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
        spec = dict(kind="dg_int", rank=rank, qm=mf, **common)
        ent = np.ascontiguousarray(mesh.interior_facets.reshape(-1), dtype=np.int32)
        return Integral(INTERIOR_FACET, -1, ent, name, src, None, spec)

    return cell, ext, inter
