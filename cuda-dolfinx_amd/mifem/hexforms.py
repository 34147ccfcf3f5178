"""Forms on (possibly jittered, hence non-affine) hexahedral meshes.

Geometry is the trilinear (Q1) map; the Jacobian is evaluated at every
quadrature point.  The Helmholtz operator

    a(u, v) = integral kappa grad u . grad v - k^2 u v dx

is written, per quadrature point q, as

    A_ij = sum_q  dhat_phi_i(q)^T  Gq  dhat_phi_j(q)  +  Mq phi_i(q) phi_j(q)
    Gq = kappa w_q |det J_q| K_q K_q^T   (3 x 3, reference coordinates)
    Mq = -k^2 w_q |det J_q|

which is the dense contraction [B; N]^T diag(Gq, Mq) [B; N] used by the
matrix-core kernels (`HexContraction` carries the reference tables).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import element as E
from .fem import Constant, Function, FunctionSpace
from .forms import CELL, Form, Integral, _SIG, _carr, _coef_expr, _src_of, _uid

_Q1_CORNERS = np.array([(a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=np.float64)


def q1_derivatives(points: np.ndarray) -> np.ndarray:
    """(nq, 8, 3) reference derivatives of the 8 trilinear vertex functions."""
    X = np.atleast_2d(points)
    out = np.empty((X.shape[0], 8, 3))
    for v, (a, b, c) in enumerate(_Q1_CORNERS):
        fx = X[:, 0] if a else 1 - X[:, 0]
        fy = X[:, 1] if b else 1 - X[:, 1]
        fz = X[:, 2] if c else 1 - X[:, 2]
        sx, sy, sz = (1 if a else -1), (1 if b else -1), (1 if c else -1)
        out[:, v, 0] = sx * fy * fz
        out[:, v, 1] = fx * sy * fz
        out[:, v, 2] = fx * fy * sz
    return out


_HEX_GEOM = """
    /* trilinear map: J[i][j] = sum_v x_{v,i} dpsi_v/dX_j at this point */
    double J00 = 0, J01 = 0, J02 = 0, J10 = 0, J11 = 0, J12 = 0, J20 = 0, J21 = 0, J22 = 0;
    for (int v = 0; v < 8; ++v) {
      const double x0 = coordinate_dofs[3 * v], x1 = coordinate_dofs[3 * v + 1], x2 = coordinate_dofs[3 * v + 2];
      const double d0 = FE_dpsi[iq][v][0], d1 = FE_dpsi[iq][v][1], d2 = FE_dpsi[iq][v][2];
      J00 += x0 * d0; J01 += x0 * d1; J02 += x0 * d2;
      J10 += x1 * d0; J11 += x1 * d1; J12 += x1 * d2;
      J20 += x2 * d0; J21 += x2 * d1; J22 += x2 * d2;
    }
    const double m00 = J11 * J22 - J12 * J21, m01 = J02 * J21 - J01 * J22, m02 = J01 * J12 - J02 * J11;
    const double detJ = J00 * m00 + J10 * m01 + J20 * m02;
    const double idet = 1.0 / detJ;
    const double K00 = m00 * idet, K01 = m01 * idet, K02 = m02 * idet;
    const double K10 = (J12 * J20 - J10 * J22) * idet, K11 = (J00 * J22 - J02 * J20) * idet, K12 = (J02 * J10 - J00 * J12) * idet;
    const double K20 = (J10 * J21 - J11 * J20) * idet, K21 = (J01 * J20 - J00 * J21) * idet, K22 = (J00 * J11 - J01 * J10) * idet;
    const double adet = fabs(detJ);
"""


@dataclass
class HexContraction:
    """Reference tables of the dense per-cell contraction (see module doc)."""
    nd: int
    nq: int
    weights: np.ndarray      # (nq,)
    phi: np.ndarray          # (nq, nd)
    dphi: np.ndarray         # (nq, nd, 3)
    dpsi: np.ndarray         # (nq, 8, 3) geometry derivatives
    k2: float
    kappa: float


@dataclass
class HexStaged:
    """Staged description of kappa grad.grad - k^2 mass on Q_k hexahedra.

    Consumed by the matrix-core owner kernel: 1D GLL Lagrange values and
    derivatives at the 1D Gauss points (the Q_k tables are their tensor
    products, node index a + (k+1) b + (k+1)^2 c, point index q1 + m q2 + m^2 q3),
    1D Gauss points and weights, k^2 and kappa.
    """
    degree: int
    m: int
    k2: float
    kappa: float

    def source(self) -> str:
        k, m = self.degree, self.m
        x, w = E._gauss_legendre01(m)
        el1 = E.gll_points(k)
        C1 = np.linalg.inv(np.vander(el1, k + 1, increasing=True))
        V = np.vander(x, k + 1, increasing=True) @ C1               # (m, k+1)
        D = np.zeros((m, k + 1))
        for j in range(1, k + 1):
            D += j * np.outer(x ** (j - 1), C1[j])
        return (f"#define HX_K {k}\n#define HX_NQ1 {m}\n#define HX_ND {(k + 1) ** 3}\n"
                f"#define HX_K2 {self.k2!r}\n#define HX_KAPPA {self.kappa!r}\n"
                + _carr("HX_L", V, "__constant__ double")
                + _carr("HX_DL", D, "__constant__ double")
                + _carr("HX_LF", V, "__constant__ float")
                + _carr("HX_DLF", D, "__constant__ float")
                + _carr("HX_X1", x, "__constant__ double")
                + _carr("HX_W1", w, "__constant__ double"))


def helmholtz(V: FunctionSpace, k: float, kappa=None, qpoints_per_dir: Optional[int] = None) -> Form:
    """kappa grad u . grad v - k^2 u v on a scalar Q_k hexahedral space."""
    if V.mesh.cell_type != "hexahedron":
        raise ValueError("helmholtz() is implemented for hexahedral meshes")
    if V.bs != 1:
        raise NotImplementedError("scalar spaces only")
    el = V.element
    nd = el.ndofs
    m = (el.degree + 1) if qpoints_per_dir is None else qpoints_per_dir
    q = E.hex_gauss(m)
    nq = q.weights.size
    tab = el.tabulate(q.points)
    phi = tab[0]
    dphi = np.stack([tab[1], tab[2], tab[3]], axis=-1)
    dpsi = q1_derivatives(q.points)
    coeffs, consts = [], []
    kexpr = _coef_expr(kappa, 0, "kappa", [], coeffs, consts, [0])
    if kexpr is None:
        kexpr = "w[0]"
    k2 = float(k) ** 2
    uid = _uid("helmholtz", el, m, k2, kexpr)
    name = f"tabulate_tensor_integral_{uid}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phi) + _carr("FE_dphi", dphi)
            + _carr("FE_dpsi", dpsi) + f"""
  const double kap = {kexpr};
  for (int iq = 0; iq < {nq}; ++iq) {{
{_HEX_GEOM}
    double g[{nd}][3];
    for (int i = 0; i < {nd}; ++i) {{
      const double d0 = FE_dphi[iq][i][0], d1 = FE_dphi[iq][i][1], d2 = FE_dphi[iq][i][2];
      g[i][0] = K00 * d0 + K10 * d1 + K20 * d2;
      g[i][1] = K01 * d0 + K11 * d1 + K21 * d2;
      g[i][2] = K02 * d0 + K12 * d1 + K22 * d2;
    }}
    const double fw = FE_w[iq] * adet;
    const double fs = fw * kap, fm = -{k2!r} * fw;
    for (int i = 0; i < {nd}; ++i)
      for (int j = 0; j < {nd}; ++j)
        A[i * {nd} + j] += fs * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2])
                         + fm * FE_phi[iq][i] * FE_phi[iq][j];
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="stiffness", cell=1, degree=el.degree, bs=1, rank=2, qm=m,
                kappa=_src_of(kexpr), k2=k2)
    staged = None
    if not coeffs and not consts:
        staged = HexStaged(el.degree, m, k2, float(kappa) if isinstance(kappa, (int, float)) else 1.0)
    form = Form(2, [V, V], [Integral(CELL, -1, cells, name, src, staged, spec)], coeffs, consts,
                "helmholtz")
    form.hex_contraction = HexContraction(nd, nq, q.weights, phi, dphi, dpsi, k2,
                                          float(kappa) if isinstance(kappa, (int, float)) else 1.0)
    return form


def hex_source(V: FunctionSpace, f: Constant, qpoints_per_dir: Optional[int] = None) -> Form:
    """L(v) = integral f v dx on a scalar Q_k hexahedral space, f constant."""
    if V.mesh.cell_type != "hexahedron" or V.bs != 1:
        raise ValueError("scalar hexahedral space expected")
    el = V.element
    nd = el.ndofs
    m = (el.degree + 1) if qpoints_per_dir is None else qpoints_per_dir
    q = E.hex_gauss(m)
    nq = q.weights.size
    phi = el.tabulate(q.points, 0)[0]
    dpsi = q1_derivatives(q.points)
    uid = _uid("hex_source", el, m)
    name = f"tabulate_tensor_integral_{uid}"
    body = (_carr("FE_w", q.weights) + _carr("FE_phi", phi) + _carr("FE_dpsi", dpsi) + f"""
  for (int iq = 0; iq < {nq}; ++iq) {{
{_HEX_GEOM}
    const double fw = FE_w[iq] * adet * c[0];
    for (int i = 0; i < {nd}; ++i) A[i] += fw * FE_phi[iq][i];
  }}
""")
    src = _SIG.format(name=name) + "\n{\n" + body + "}\n"
    cells = np.arange(V.mesh.num_cells, dtype=np.int32)
    spec = dict(kind="source", cell=1, degree=el.degree, bs=1, rank=1, qm=m, f=("c", 0),
                f_degree=0)
    return Form(1, [V], [Integral(CELL, -1, cells, name, src, None, spec)], [], [f], "hex_source")
