"""Per-cell-type pieces of the generated element bodies on affine simplices:
quadrature rules, facet tables, and the C snippets for the affine map, the
physical basis gradients and the facet normal, for tetrahedra (tdim 3) and
triangles (tdim 2).

coordinate_dofs always holds 3 components per vertex (DOLFINx's geometry
layout; z = 0 on a 2D mesh), so the snippets read it at stride 3.  Facet i
of a cell is opposite its vertex i in both cases.  The snippets define:

  geom(s, xd):  K{s}[t][t] = J^-1, adet{s} = |det J| (and h{s}, the cell
                diameter, on tetrahedra)
  grad(s, tab, nd):  g{s}[nd][t] = sum_r tab[i][r] K{s}[r][d]
  FACET:  lf0 = entity_local_index[0], the facet's size ds (edge length /
          twice the triangle area) and its unit normal nx, ny, nz pointing
          away from the cell's vertex opposite the facet (nz = 0 in 2D)
"""
from __future__ import annotations

import numpy as np

from . import element as E
from .dgforms import _FACET as _FACET3, _geom as _geom3, _grad as _grad3

_FACET2 = """
  /* edge lf0 of the first cell has vertices (1,2), (0,2) or (0,1): picked by
     selects, so coordinate_dofs is only read at constant offsets */
  const int lf0 = entity_local_index[0];
  const double* xa = coordinate_dofs;
  double P0[2], P1[2], PO[2];
  for (int d = 0; d < 2; ++d) {
    const double v0 = xa[d], v1 = xa[3 + d], v2 = xa[6 + d];
    P0[d] = lf0 == 0 ? v1 : v0;
    P1[d] = lf0 <= 1 ? v2 : v1;
    PO[d] = lf0 == 0 ? v0 : (lf0 == 1 ? v1 : v2);
  }
  const double tx = P1[0] - P0[0], ty = P1[1] - P0[1];
  const double ds = sqrt(tx * tx + ty * ty);
  double nx = ty / ds, ny = -tx / ds;
  /* outward from the first cell: away from the vertex opposite the edge */
  if (nx * (PO[0] - P0[0]) + ny * (PO[1] - P0[1]) > 0.0) { nx = -nx; ny = -ny; }
  const double nz = 0.0;
  (void)nz;
"""


def _geom2(s: str, xd: str) -> str:
    return f"""
  const double J00{s} = {xd}[3] - {xd}[0], J01{s} = {xd}[6] - {xd}[0];
  const double J10{s} = {xd}[4] - {xd}[1], J11{s} = {xd}[7] - {xd}[1];
  const double det{s} = J00{s} * J11{s} - J01{s} * J10{s};
  const double id{s} = 1.0 / det{s};
  const double K{s}[2][2] = {{{{J11{s} * id{s}, -J01{s} * id{s}}}, {{-J10{s} * id{s}, J00{s} * id{s}}}}};
  const double adet{s} = fabs(det{s});
  const double h{s} = 0.0;
"""


def _grad2(s: str, tab: str, nd: int) -> str:
    return f"""    double g{s}[{nd}][2];
    for (int i = 0; i < {nd}; ++i)
      for (int d = 0; d < 2; ++d)
        g{s}[i][d] = {tab}[i][0] * K{s}[0][d] + {tab}[i][1] * K{s}[1][d];
"""


class Kit:
    """The generated-body pieces of one simplex cell type."""

    def __init__(self, cell_type: str):
        if cell_type not in ("tetrahedron", "triangle"):
            raise NotImplementedError(f"affine simplex bodies need tetrahedra or triangles, not {cell_type}")
        self.cell = cell_type
        self.tdim = 3 if cell_type == "tetrahedron" else 2
        self.nf = self.tdim + 1
        self.oracle_cell = 0 if cell_type == "tetrahedron" else 2
        self.FACET = _FACET3 if self.tdim == 3 else _FACET2

    def geom(self, s: str, xd: str) -> str:
        return _geom3(s, xd) if self.tdim == 3 else _geom2(s, xd)

    def grad(self, s: str, tab: str, nd: int) -> str:
        return _grad3(s, tab, nd) if self.tdim == 3 else _grad2(s, tab, nd)

    def cell_rule(self, m: int) -> E.Quadrature:
        return E.tet_collapsed(m) if self.tdim == 3 else E.tri_collapsed(m)

    def facet_rule(self, m: int) -> E.Quadrature:
        return E.tri_collapsed(m) if self.tdim == 3 else E.line_gauss(m)

    def dot(self, a: str, b: str) -> str:
        """a[0] * b[0] + ... over the tdim components ({} marks the index)."""
        return " + ".join(f"{a.format(d)} * {b.format(d)}" for d in range(self.tdim))

    def normal_dot(self, a: str) -> str:
        nrm = ("nx", "ny", "nz")
        return " + ".join(f"{a.format(d)} * {nrm[d]}" for d in range(self.tdim))

    def cell_tables(self, el, q):
        """(nq, nd) values and (nq, nd, tdim) reference gradients at cell points."""
        t = el.tabulate(q.points)
        return t[0], np.stack([t[1 + d] for d in range(self.tdim)], axis=-1)

    def facet_tables(self, el, q):
        """(nf, nq, nd) values and (nf, nq, nd, tdim) reference gradients on the facets."""
        phi, dphi = [], []
        for lf in range(self.nf):
            X = (E.tet_facet_points(lf, q.points) if self.tdim == 3
                 else E.tri_facet_points(lf, q.points))
            t = el.tabulate(X)
            phi.append(t[0])
            dphi.append(np.stack([t[1 + d] for d in range(self.tdim)], axis=-1))
        return np.stack(phi), np.stack(dphi)

    def facet_tables_permuted(self, el, q):
        """[lf][perm][iq] tables of interior-facet bodies (values, gradients)."""
        if self.tdim == 3:
            return E.facet_tables_permuted(el, q.points)
        return E.tri_facet_tables_permuted(el, q.points)
