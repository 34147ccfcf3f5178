"""Variational forms and their generated element kernels (the FFCx stand-in).

A `Form` carries what `dolfinx.fem.Form` + its FFCx module hand to
cuDOLFINx: rank, function spaces, integrals (type, subdomain id, entity
list, C source of ``tabulate_tensor_integral_<id>`` in the UFCx ABI),
coefficients (packed per cell, DOLFINx layout) and constants.

The UFCx tabulate_tensor ABI (reference: python/cudolfinx/jit.py:69-76):

    void tabulate_tensor_integral_<id>(double* A, const double* w,
        const double* c, const double* coordinate_dofs,
        const int* entity_local_index, const uint8_t* quadrature_permutation)

`A` is accumulated into (the caller zeroes it), `coordinate_dofs` is
[nv x 3], `entity_local_index` is NULL for cells, &lf for exterior facets and
{lf0, lf1} for interior facets.  Interior-facet tensors are [2 nd0 x 2 nd1]
(row stride 2 nd1), coefficients interlaced per coefficient as
(w_cell0, w_cell1) (reference: cpp/cudolfinx/fem/CUDAFormIntegral.cpp:1855-1874).

Besides the UFCx body, second-order forms on affine simplices also get a
*staged* description (reference gradients, weights and a `combine` snippet
that maps the geometric tensor G_ij = sum_q w_q |det J| grad phi_i (x) grad phi_j
to the bs x bs block of entry (i, j)).  The HIP assembly templates use it to
compute element rows cooperatively instead of materialising nd x nd tensors
per thread.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Union

import numpy as np

from . import element as E
from .fem import Constant, Function, FunctionSpace
from .mesh import Mesh

CELL, EXTERIOR_FACET, INTERIOR_FACET = "cell", "exterior_facet", "interior_facet"


# ----------------------------------------------------------------- C helpers
def _carr(name: str, a: np.ndarray, qual: str = "static const double") -> str:
    a = np.asarray(a, dtype=np.float64)
    dims = "".join(f"[{d}]" for d in a.shape)
    vals = ", ".join(repr(float(v)) for v in a.reshape(-1))
    return f"{qual} {name}{dims} = {{{vals}}};\n"


_SIG = ("void {name}(double* restrict A, const double* restrict w, const double* restrict c, "
        "const double* restrict coordinate_dofs, const int* restrict entity_local_index, "
        "const uint8_t* restrict quadrature_permutation)")

_AFFINE_GEOM = """
  /* affine simplex: J[i][j] = x_{j+1,i} - x_{0,i} */
  const double* xd = coordinate_dofs;
  const double J00 = xd[3] - xd[0], J01 = xd[6] - xd[0], J02 = xd[9] - xd[0];
  const double J10 = xd[4] - xd[1], J11 = xd[7] - xd[1], J12 = xd[10] - xd[1];
  const double J20 = xd[5] - xd[2], J21 = xd[8] - xd[2], J22 = xd[11] - xd[2];
  const double m00 = J11 * J22 - J12 * J21, m01 = J02 * J21 - J01 * J22, m02 = J01 * J12 - J02 * J11;
  const double detJ = J00 * m00 + J10 * m01 + J20 * m02;
  const double idet = 1.0 / detJ;
  const double K00 = m00 * idet, K01 = m01 * idet, K02 = m02 * idet;
  const double K10 = (J12 * J20 - J10 * J22) * idet, K11 = (J00 * J22 - J02 * J20) * idet, K12 = (J02 * J10 - J00 * J12) * idet;
  const double K20 = (J10 * J21 - J11 * J20) * idet, K21 = (J01 * J20 - J00 * J21) * idet, K22 = (J00 * J11 - J01 * J10) * idet;
  const double adet = fabs(detJ);
"""


def _uid(*parts) -> str:
    h = hashlib.sha1("|".join(str(p) for p in parts).encode()).hexdigest()[:12]
    return h


# ----------------------------------------------------------------- data model
@dataclass
class StagedKernel:
    """Second-order form on affine simplices, split into stages (see module doc)."""
    nd: int
    bs: int
    nq: int
    weights: np.ndarray          # (nq,)
    dphi: np.ndarray             # (nq, nd, 3) reference gradients
    combine: str                 # body of combine(G, w, c, out)
    coeff_width: int             # packed coefficient values per cell
    symmetric: bool = True
    trace_only: bool = False     # combine reads only G[0] + G[4] + G[8]
    full_tensor: bool = False    # emit the (nd, nd, 3, 3) reference tensor (G = |det J| K^T R K)

    def reference_tensor(self) -> np.ndarray:
        """(nd, nd, 6) R_ij = sum_q w_q dphi_i (x) dphi_j in reference coordinates,
        symmetrised for a symmetric metric M: components (00, 11, 22, 01+10, 02+20, 12+21).

        For affine cells tr G_ij = |det J| sum_rs M_rs R_ij[r][s] with M = K K^T,
        K = J^-1, so the trace needs 6 multiply-adds per entry at any degree.
        """
        R = np.einsum("q,qia,qjb->ijab", self.weights, self.dphi, self.dphi)
        return np.stack([R[..., 0, 0], R[..., 1, 1], R[..., 2, 2],
                         R[..., 0, 1] + R[..., 1, 0], R[..., 0, 2] + R[..., 2, 0],
                         R[..., 1, 2] + R[..., 2, 1]], axis=-1)

    def reference_tensor_full(self) -> np.ndarray:
        """(nd, nd, 9) R_ij[3 r + s] = sum_q w_q dphi_i/dX_r dphi_j/dX_s on the reference cell.

        For affine cells G_ij[p][c] = |det J| sum_rs K_rp R_ij[r][s] K_sc with
        K = J^-1 (K_rp = dX_r / dx_p): 54 multiply-adds per entry at any degree.
        """
        R = np.einsum("q,qia,qjb->ijab", self.weights, self.dphi, self.dphi)
        return R.reshape(self.nd, self.nd, 9)

    def source(self) -> str:
        trace = ("#define SK_TRACE 1\n" + _carr("SK_RS", self.reference_tensor(), "__constant__ double")
                 if self.trace_only else "")
        if self.full_tensor and not self.trace_only:
            trace += ("#define SK_FULLRT 1\n"
                      + _carr("SK_RT", self.reference_tensor_full(), "__constant__ double"))
        return (f"#define SK_ND {self.nd}\n#define SK_BS {self.bs}\n#define SK_NQ {self.nq}\n"
                + trace
                + _carr("SK_W", self.weights, "__constant__ double")
                + _carr("SK_DPHI", self.dphi, "__constant__ double")
                + "static __device__ __forceinline__ void sk_combine(const double* G, const double* "
                  "__restrict__ w, const double* __restrict__ c, double* out)\n{\n"
                + self.combine + "}\n")


@dataclass
class Integral:
    type: str
    id: int
    entities: np.ndarray         # int32, flattened (1, 2 or 4 ints per entity)
    name: str
    source: str
    staged: Optional[StagedKernel] = None
    spec: dict = field(default_factory=dict)   # integrand description (for checkers)

    @property
    def num_entities(self) -> int:
        per = {CELL: 1, EXTERIOR_FACET: 2, INTERIOR_FACET: 4}[self.type]
        return int(self.entities.size // per)


@dataclass(eq=False)
class Form:
    rank: int
    function_spaces: List[FunctionSpace]
    integrals: List[Integral]
    coefficients: List[Function] = field(default_factory=list)
    constants: List[Constant] = field(default_factory=list)
    description: str = ""
    domain: Optional[Mesh] = None      # integration domain (needed by rank-0 forms)

    @property
    def mesh(self) -> Mesh:
        if self.domain is not None:
            return self.domain
        if self.function_spaces:
            return self.function_spaces[0].mesh
        return self.coefficients[0].function_space.mesh

    @property
    def coefficient_offsets(self) -> List[int]:
        off = [0]
        for f in self.coefficients:
            V = f.function_space
            off.append(off[-1] + V.element.ndofs * V.bs)
        return off

    def constant_values(self) -> np.ndarray:
        if not self.constants:
            return np.zeros(0)
        return np.concatenate([c.value for c in self.constants])

    def pack_coefficients(self) -> np.ndarray:
        """(Nc, W) packed coefficient values, DOLFINx cell layout."""
        nc = self.mesh.num_cells
        off = self.coefficient_offsets
        out = np.zeros((nc, off[-1]))
        for j, f in enumerate(self.coefficients):
            out[:, off[j]:off[j + 1]] = f.x[f.function_space.unrolled_dofmap()]
        return out

    def integral_ids(self, itype: str) -> List[int]:
        return [I.id for I in self.integrals if I.type == itype]


# ----------------------------------------------------------------- builders
def _coef_expr(obj, slot: int, name: str, decl: List[str], coeffs: List, consts: List,
               const_off: List[int]):
    """Return a C expression for a scalar DG0 coefficient or scalar constant."""
    if obj is None:
        return "1.0"
    if isinstance(obj, (int, float)):
        return repr(float(obj))
    if isinstance(obj, Constant):
        consts.append(obj)
        o = const_off[0]
        const_off[0] += obj.value.size
        return f"c[{o}]"
    if isinstance(obj, Function):
        V = obj.function_space
        if not (V.element.discontinuous and V.element.degree == 0 and V.bs == 1):
            raise NotImplementedError("cell coefficients in second-order forms must be scalar DG0")
        coeffs.append(obj)
        return None   # resolved by caller with the packed offset
    raise TypeError(f"unsupported coefficient {type(obj)}")


def _src_of(expr: str):
    """('w', off) / ('c', off) / ('val', v) from a generated C expression."""
    if expr.startswith("w["):
        return ("w", int(expr[2:-1]))
    if expr.startswith("c["):
        return ("c", int(expr[2:-1]))
    return ("val", float(expr))


def _check_affine(V: FunctionSpace):
    if V.mesh.cell_type != "tetrahedron":
        raise NotImplementedError("affine forms need a tetrahedral mesh")


def stiffness(V: FunctionSpace, kappa=None, qdegree: Optional[int] = None) -> Form:
    """a(u, v) = integral kappa grad u . grad v dx; kappa: None, number, Constant or DG0 Function.

    Tetrahedra use the affine map; hexahedra go through `helmholtz(V, k=0)`.
    """
    if V.mesh.cell_type == "hexahedron":
        return helmholtz(V, 0.0, kappa=kappa)
    _check_affine(V)
    if V.bs != 1:
        raise NotImplementedError("stiffness() is scalar; use elasticity() for vectors")
    el = V.element
    nd = el.ndofs
    q = E.tet_rule(2 * (el.degree - 1) if qdegree is None else qdegree)
    tab = el.tabulate(q.points)
    dphi = np.stack([tab[1], tab[2], tab[3]], axis=-1)            # (nq, nd, 3)
    coeffs, consts = [], []
    kexpr = _coef_expr(kappa, 0, "kappa", [], coeffs, consts, [0])
    if kexpr is None:
        kexpr = "w[0]"
    uid = _uid("stiffness", el, q.points.tobytes(), kexpr)
    name = f"tabulate_tensor_integral_{uid}"
    body = (_carr("FE_w", q.weights) + _carr("FE_dphi", dphi) + _AFFINE_GEOM + f"""
  const double kap = {kexpr};
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
    double g[{nd}][3];
    for (int i = 0; i < {nd}; ++i) {{
      const double d0 = FE_dphi[iq][i][0], d1 = FE_dphi[iq][i][1], d2 = FE_dphi[iq][i][2];
      g[i][0] = K00 * d0 + K10 * d1 + K20 * d2;
      g[i][1] = K01 * d0 + K11 * d1 + K21 * d2;
      g[i][2] = K02 * d0 + K12 * d1 + K22 * d2;
    }}
    const double fw = FE_w[iq] * adet * kap;
    for (int i = 0; i < {nd}; ++i)
      for (int j = 0; j < {nd}; ++j)
        A[i * {nd} + j] += fw * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    staged = StagedKernel(nd=nd, bs=1, nq=q.weights.size, weights=q.weights, dphi=dphi,
                          combine=f"  out[0] = ({kexpr}) * (G[0] + G[4] + G[8]);\n",
                          coeff_width=len(coeffs), trace_only=True)
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="stiffness", cell=0, degree=el.degree, bs=1, rank=2,
                qdeg=2 * (el.degree - 1) if qdegree is None else qdegree,
                kappa=_src_of(kexpr), k2=0.0)
    return Form(2, [V, V], [Integral(CELL, -1, cells, name, src, staged, spec)], coeffs, consts,
                "stiffness")


def elasticity(V: FunctionSpace, mu, lam, qdegree: Optional[int] = None,
               reference_tensor: bool = False) -> Form:
    """a(u, v) = integral 2 mu eps(u):eps(v) + lam div u div v dx on a vector Lagrange space.

    Entry ((i, a), (j, b)) = mu (delta_ab G_ij:I + G_ij[b][a]) + lam G_ij[a][b] with
    G_ij[p][q] = integral d_p phi_i d_q phi_j.  reference_tensor=True also emits
    the (nd, nd, 9) reference tensors so the owner-computes kernel forms
    G_ij = |det J| K^T R_ij K (54 FMAs) instead of summing over quadrature
    points; the table costs nd^2 * 72 bytes of LDS per workgroup (at P2 it
    lowers the resident workgroups per CU from 4 to 3, which costs more than
    the arithmetic saves).
    """
    _check_affine(V)
    if V.bs != 3:
        raise ValueError("elasticity needs a 3-vector space")
    el = V.element
    nd = el.ndofs
    q = E.tet_rule(2 * (el.degree - 1) if qdegree is None else qdegree)
    tab = el.tabulate(q.points)
    dphi = np.stack([tab[1], tab[2], tab[3]], axis=-1)
    coeffs, consts = [], []
    coff = [0]
    exprs = []
    for obj in (mu, lam):
        e = _coef_expr(obj, 0, "", [], coeffs, consts, coff)
        exprs.append(e if e is not None else f"w[{len(coeffs) - 1}]")
    mexpr, lexpr = exprs
    uid = _uid("elasticity", el, q.points.tobytes(), mexpr, lexpr)
    name = f"tabulate_tensor_integral_{uid}"
    n = 3 * nd
    body = (_carr("FE_w", q.weights) + _carr("FE_dphi", dphi) + _AFFINE_GEOM + f"""
  const double mu = {mexpr}, lam = {lexpr};
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
    double g[{nd}][3];
    for (int i = 0; i < {nd}; ++i) {{
      const double d0 = FE_dphi[iq][i][0], d1 = FE_dphi[iq][i][1], d2 = FE_dphi[iq][i][2];
      g[i][0] = K00 * d0 + K10 * d1 + K20 * d2;
      g[i][1] = K01 * d0 + K11 * d1 + K21 * d2;
      g[i][2] = K02 * d0 + K12 * d1 + K22 * d2;
    }}
    const double fw = FE_w[iq] * adet;
    for (int i = 0; i < {nd}; ++i)
      for (int j = 0; j < {nd}; ++j) {{
        const double gg = g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            A[(3 * i + a) * {n} + 3 * j + b] += fw * (mu * ((a == b ? gg : 0.0) + g[i][b] * g[j][a])
                                                    + lam * g[i][a] * g[j][b]);
      }}
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    combine = f"""  const double mu = {mexpr}, lam = {lexpr};
  const double tr = G[0] + G[4] + G[8];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      out[3 * a + b] = mu * ((a == b ? tr : 0.0) + G[3 * b + a]) + lam * G[3 * a + b];
"""
    staged = StagedKernel(nd=nd, bs=3, nq=q.weights.size, weights=q.weights, dphi=dphi,
                          combine=combine, coeff_width=len(coeffs), full_tensor=reference_tensor)
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="elasticity", cell=0, degree=el.degree, bs=3, rank=2,
                qdeg=2 * (el.degree - 1) if qdegree is None else qdegree,
                mu=_src_of(mexpr), lam=_src_of(lexpr))
    return Form(2, [V, V], [Integral(CELL, -1, cells, name, src, staged, spec)], coeffs, consts,
                "elasticity")


def mass(V: FunctionSpace, W: Optional[FunctionSpace] = None, kappa=None,
         qdegree: Optional[int] = None) -> Form:
    """a(u, v) = integral kappa u v dx, v in V (rows), u in W (columns, default V);
    scalar spaces on affine tetrahedra; kappa None, a number, Constant or DG0 Function."""
    W = V if W is None else W
    _check_affine(V)
    if V.mesh is not W.mesh or V.bs != 1 or W.bs != 1:
        raise NotImplementedError("mass() couples scalar spaces on one mesh")
    ev, ew = V.element, W.element
    nv, nw = ev.ndofs, ew.ndofs
    qd = ev.degree + ew.degree if qdegree is None else qdegree
    q = E.tet_rule(qd)
    coeffs, consts = [], []
    kexpr = _coef_expr(kappa, 0, "kappa", [], coeffs, consts, [0])
    if kexpr is None:
        kexpr = "w[0]"
    uid = _uid("mass", ev, ew, q.points.tobytes(), kexpr)
    name = f"tabulate_tensor_integral_{uid}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", ev.tabulate(q.points, 0)[0])
            + _carr("FE_psi", ew.tabulate(q.points, 0)[0]) + _AFFINE_GEOM + f"""
  (void)K00; (void)K01; (void)K02; (void)K10; (void)K11; (void)K12; (void)K20; (void)K21; (void)K22;
  const double kap = {kexpr};
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
    const double fw = FE_w[iq] * adet * kap;
    for (int i = 0; i < {nv}; ++i)
      for (int j = 0; j < {nw}; ++j)
        A[i * {nw} + j] += fw * FE_phi[iq][i] * FE_psi[iq][j];
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="mass", cell=0, degree=ev.degree, trial_degree=ew.degree, bs=1, rank=2,
                qdeg=qd, kappa=_src_of(kexpr))
    return Form(2, [V, W], [Integral(CELL, -1, cells, name, src, None, spec)], coeffs, consts,
                "mass")


def source(V: FunctionSpace, f, qdegree: Optional[int] = None) -> Form:
    """L(v) = integral f . v dx; f a Constant (bs values) or a Function on a scalar/bs space."""
    _check_affine(V)
    el = V.element
    nd, bs = el.ndofs, V.bs
    coeffs, consts = [], []
    if isinstance(f, Constant):
        if f.value.size != bs:
            raise ValueError("constant source must have bs components")
        consts.append(f)
        qd = el.degree if qdegree is None else qdegree
        mode = "const"
    elif isinstance(f, Function):
        Vf = f.function_space
        if Vf.bs != bs:
            raise ValueError("source Function must match the block size")
        coeffs.append(f)
        qd = el.degree + Vf.element.degree if qdegree is None else qdegree
        mode = "func"
    else:
        raise TypeError("f must be a Constant or Function")
    q = E.tet_rule(qd)
    phi = el.tabulate(q.points, 0)[0]                              # (nq, nd)
    parts = [_carr("FE_w", q.weights), _carr("FE_phi", phi)]
    if mode == "func":
        fel = f.function_space.element
        nf = fel.ndofs
        parts.append(_carr("FE_psi", fel.tabulate(q.points, 0)[0]))
        fval = f"""    double fv[{bs}] = {{0}};
    for (int k = 0; k < {nf}; ++k)
      for (int a = 0; a < {bs}; ++a) fv[a] += FE_psi[iq][k] * w[k * {bs} + a];
"""
    else:
        fval = f"    double fv[{bs}];\n    for (int a = 0; a < {bs}; ++a) fv[a] = c[a];\n"
    uid = _uid("source", el, bs, q.points.tobytes(), mode,
               f.function_space.element if mode == "func" else "")
    name = f"tabulate_tensor_integral_{uid}"
    body = "".join(parts) + _AFFINE_GEOM + f"""
  for (int iq = 0; iq < {q.weights.size}; ++iq) {{
{fval}    const double fw = FE_w[iq] * adet;
    for (int i = 0; i < {nd}; ++i)
      for (int a = 0; a < {bs}; ++a)
        A[i * {bs} + a] += fw * FE_phi[iq][i] * fv[a];
  }}
"""
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="source", cell=0, degree=el.degree, bs=bs, rank=1, qdeg=qd,
                f=("c", 0) if mode == "const" else ("w", 0),
                f_degree=f.function_space.element.degree if mode == "func" else 0)
    return Form(1, [V], [Integral(CELL, -1, cells, name, src, None, spec)], coeffs, consts,
                "source")


def facet_source(V: FunctionSpace, g: Constant, facets: np.ndarray, subdomain_id: int = 1,
                 qdegree: Optional[int] = None) -> Form:
    """L(v) = integral_Gamma g . v ds over the given exterior [cell, local facet] pairs."""
    _check_affine(V)
    el = V.element
    nd, bs = el.ndofs, V.bs
    if g.value.size != bs:
        raise ValueError("traction must have bs components")
    q = E.tri_rule(el.degree if qdegree is None else qdegree)
    nq = q.weights.size
    phi = np.stack([el.tabulate(E.tet_facet_points(lf, q.points), 0)[0] for lf in range(4)])
    # reference facet edge vectors: facet lf spans v1 - v0, v2 - v0 of its vertices
    uid = _uid("facet_source", el, bs, q.points.tobytes())
    name = f"tabulate_tensor_integral_{uid}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phi) + """
  static const int FV[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  const int lf = entity_local_index[0];
  const double* xd = coordinate_dofs;
  const int a0 = FV[lf][0], a1 = FV[lf][1], a2 = FV[lf][2];
  const double e1x = xd[3 * a1] - xd[3 * a0], e1y = xd[3 * a1 + 1] - xd[3 * a0 + 1], e1z = xd[3 * a1 + 2] - xd[3 * a0 + 2];
  const double e2x = xd[3 * a2] - xd[3 * a0], e2y = xd[3 * a2 + 1] - xd[3 * a0 + 1], e2z = xd[3 * a2 + 2] - xd[3 * a0 + 2];
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double ds = sqrt(nx * nx + ny * ny + nz * nz);
""" + f"""
  for (int iq = 0; iq < {nq}; ++iq) {{
    const double fw = FE_w[iq] * ds;
    for (int i = 0; i < {nd}; ++i)
      for (int a = 0; a < {bs}; ++a)
        A[i * {bs} + a] += fw * FE_phi[lf][iq][i] * c[a];
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    ent = np.ascontiguousarray(np.asarray(facets, dtype=np.int32).reshape(-1))
    spec = dict(kind="facet_source", cell=0, degree=el.degree, bs=bs, rank=1,
                qdeg=el.degree if qdegree is None else qdegree, f=("c", 0))
    return Form(1, [V], [Integral(EXTERIOR_FACET, subdomain_id, ent, name, src, None, spec)],
                [], [g], "facet_source")


def linear_sum(a: Form, b: Form) -> Form:
    """L = a + b for rank-1 forms: integrals concatenated in cell, ext, int order.

    The coefficient/constant arrays of the two forms are merged and the
    integral bodies rewritten so w/c offsets stay valid.
    """
    if a.rank != b.rank or a.function_spaces[0] is not b.function_spaces[0]:
        raise ValueError("forms must have equal rank and spaces")
    coffs_a = a.coefficient_offsets[-1]
    consts_a = int(sum(c.value.size for c in a.constants))
    new_b = []
    for I in b.integrals:
        src = I.source
        if coffs_a or consts_a:
            src = src.replace("{\n", "{\n  w += %d; c += %d;\n" % (coffs_a, consts_a), 1)
        spec = dict(I.spec)
        for key, val in list(spec.items()):
            if isinstance(val, tuple) and val[0] == "w":
                spec[key] = ("w", val[1] + coffs_a)
            elif isinstance(val, tuple) and val[0] == "c":
                spec[key] = ("c", val[1] + consts_a)
        new_b.append(Integral(I.type, I.id, I.entities, I.name, src, I.staged, spec))
    order = {CELL: 0, EXTERIOR_FACET: 1, INTERIOR_FACET: 2}
    ints = sorted(a.integrals + new_b, key=lambda I: order[I.type])
    return Form(a.rank, a.function_spaces, ints, a.coefficients + b.coefficients,
                a.constants + b.constants, a.description + "+" + b.description)


from .hexforms import helmholtz, hex_source  # noqa: E402
