/*
 * oracle.c -- FP64 CPU restatement of the cuDOLFINx / DOLFINx assembly path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * assembly engine: tests, the smoke check and the CPU-baseline leg of the
 * benchmark call it; the product never links or calls it.
 *
 * It restates, in plain serial C (plus a coloured multi-threaded copy of
 * the matrix and vector loops, timed as the multi-core CPU baseline):
 *   - quadrature (Gauss-Jacobi / Gauss-Legendre by Newton on the Jacobi
 *     recurrence, collapsed tetrahedron and triangle rules);
 *   - Lagrange bases (P0-P3 tetrahedra by closed-form barycentric
 *     polynomials, Q1-Q3 hexahedra as GLL tensor products) in the local dof
 *     order the engine uses (vertices, edges e0=(2,3) e1=(1,3) e2=(1,2)
 *     e3=(0,3) e4=(0,2) e5=(0,1), faces opposite vertex i; hexahedra
 *     lexicographic);
 *   - the element integrals the engine generates (stiffness / Helmholtz,
 *     mass between two Lagrange spaces, linear elasticity, the nonlinear
 *     residuals / Jacobians of the reference's assembly tests and the
 *     diagonal part of a bilinear form, volume and facet sources, SIPG + p-Laplacian DG
 *     with its energy functional, a functional over cells, exterior and
 *     interior facets),
 *     evaluated directly at quadrature points, in the UFCx argument layout;
 *   - same-map interpolation between two spaces on one mesh;
 *   - the sparsity build (DOLFINx create_sparsity_pattern + the bs-unrolling
 *     of la/petsc.cpp:14-65);
 *   - the assembly loops of cpp/cudolfinx/fem/CUDAFormIntegral.cpp
 *     (matrix cells :1103-1226, exterior facets :1667-1793, interior facets
 *     :1988-2103 with the joint dofmaps :1948-1986 and coefficient
 *     interlacing :1855-1874; vectors :270-522; lifting :603-947), the
 *     set_bc / set_diagonal / pack_coefficients kernels of
 *     cpp/cudolfinx/fem/CUDAAssembler.cpp:59-164, in the integral order of
 *     CUDAAssembler.h:683-740, with the DOLFINx semantics of lifting over
 *     every integral (cells, exterior and interior facets).
 *
 * Everything is FP64; entries are accumulated in entity order, so results are
 * deterministic.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAXQ 4096
#define OR_MAXND 64

/* ------------------------------------------------------------------ */
/* quadrature                                                          */
/* ------------------------------------------------------------------ */

/* P_n^{(a,b)}(x) by the three-term recurrence. */
static double jacobi_p(int n, double a, double b, double x)
{
  if (n == 0) return 1.0;
  double p0 = 1.0, p1 = 0.5 * ((a + b + 2.0) * x + (a - b));
  for (int k = 1; k < n; ++k) {
    double s = 2.0 * k + a + b;
    double a1 = 2.0 * (k + 1) * (k + a + b + 1.0) * s;
    double a2 = (s + 1.0) * (a * a - b * b);
    double a3 = s * (s + 1.0) * (s + 2.0);
    double a4 = 2.0 * (k + a) * (k + b) * (s + 2.0);
    double p2 = ((a2 + a3 * x) * p1 - a4 * p0) / a1;
    p0 = p1;
    p1 = p2;
  }
  return p1;
}

static double jacobi_dp(int n, double a, double b, double x)
{
  if (n == 0) return 0.0;
  return 0.5 * (n + a + b + 1.0) * jacobi_p(n - 1, a + 1.0, b + 1.0, x);
}

/* m-point Gauss-Jacobi rule on [0,1] for the weight (1-t)^alpha. */
void or_gauss_jacobi(int m, double alpha, double* t, double* w)
{
  double x[OR_MAXND];
  const double b = 0.0;
  for (int k = 0; k < m; ++k) {
    double r = -cos((2.0 * k + 1.0) * M_PI / (2.0 * m));
    if (k > 0) r = 0.5 * (r + x[k - 1]);
    for (int it = 0; it < 200; ++it) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += 1.0 / (r - x[j]);
      double p = jacobi_p(m, alpha, b, r), dp = jacobi_dp(m, alpha, b, r);
      double delta = -p / (dp - s * p);
      r += delta;
      if (fabs(delta) < 1e-17) break;
    }
    x[k] = r;
  }
  double c = pow(2.0, alpha + b + 1.0) * tgamma(m + alpha + 1.0) * tgamma(m + b + 1.0) /
             (tgamma(m + alpha + b + 1.0) * tgamma(m + 1.0));
  for (int k = 0; k < m; ++k) {
    double dp = jacobi_dp(m, alpha, b, x[k]);
    double wk = c / ((1.0 - x[k] * x[k]) * dp * dp);
    t[k] = 0.5 * (x[k] + 1.0);
    w[k] = wk / pow(2.0, alpha + 1.0);
  }
}

/* Collapsed (Stroud) m^3 rule on the unit tetrahedron; returns count. */
int or_tet_rule(int m, double* pts, double* wts)
{
  double a[OR_MAXND], wa[OR_MAXND], b[OR_MAXND], wb[OR_MAXND], c[OR_MAXND], wc[OR_MAXND];
  or_gauss_jacobi(m, 2.0, a, wa);
  or_gauss_jacobi(m, 1.0, b, wb);
  or_gauss_jacobi(m, 0.0, c, wc);
  int n = 0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j)
      for (int l = 0; l < m; ++l) {
        pts[3 * n] = a[i];
        pts[3 * n + 1] = b[j] * (1.0 - a[i]);
        pts[3 * n + 2] = c[l] * (1.0 - a[i]) * (1.0 - b[j]);
        wts[n] = wa[i] * wb[j] * wc[l];
        ++n;
      }
  return n;
}

/* Collapsed m^2 rule on the unit triangle. */
int or_tri_rule(int m, double* pts, double* wts)
{
  double a[OR_MAXND], wa[OR_MAXND], b[OR_MAXND], wb[OR_MAXND];
  or_gauss_jacobi(m, 1.0, a, wa);
  or_gauss_jacobi(m, 0.0, b, wb);
  int n = 0;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      pts[2 * n] = a[i];
      pts[2 * n + 1] = b[j] * (1.0 - a[i]);
      wts[n] = wa[i] * wb[j];
      ++n;
    }
  return n;
}

/* Gauss-Legendre m^3 rule on the unit cube (x fastest). */
int or_hex_rule(int m, double* pts, double* wts)
{
  double g[OR_MAXND], wg[OR_MAXND];
  or_gauss_jacobi(m, 0.0, g, wg);
  int n = 0;
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) {
        pts[3 * n] = g[i];
        pts[3 * n + 1] = g[j];
        pts[3 * n + 2] = g[k];
        wts[n] = wg[i] * wg[j] * wg[k];
        ++n;
      }
  return n;
}

/* ------------------------------------------------------------------ */
/* bases                                                               */
/* ------------------------------------------------------------------ */

static const int TET_E[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};
static const int TET_F[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const double TET_V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const double DLAM[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

/* cell types: 0 tetrahedron, 1 hexahedron, 2 triangle (facet i opposite vertex i) */
static const int TRI_E[3][2] = {{1, 2}, {0, 2}, {0, 1}};
static const double TRI_V[3][2] = {{0, 0}, {1, 0}, {0, 1}};
static const double DLAM2[3][2] = {{-1, -1}, {1, 0}, {0, 1}};

int or_ndofs(int cell, int k)
{
  if (cell == 0) return (k + 1) * (k + 2) * (k + 3) / 6;
  if (cell == 2) return (k + 1) * (k + 2) / 2;
  return (k + 1) * (k + 1) * (k + 1);
}

/* P_k on the reference triangle (Basix node order: vertices, edges (1,2),
 * (0,2), (0,1) from the first to the second vertex, interior): values
 * phi[nd], gradients dphi[nd][3] (third component 0). */
static void tri_basis(int k, const double* X, double* phi, double* dphi)
{
  double l[3] = {1.0 - X[0] - X[1], X[0], X[1]};
  int n = 0;
#define SETG2(idx, c0, v0, c1, v1)                                                     \
  do {                                                                                 \
    for (int d = 0; d < 2; ++d) dphi[3 * (idx) + d] = (c0) * DLAM2[v0][d] + (c1) * DLAM2[v1][d]; \
    dphi[3 * (idx) + 2] = 0.0;                                                         \
  } while (0)
  if (k == 0) {
    phi[0] = 1.0;
    dphi[0] = dphi[1] = dphi[2] = 0.0;
    return;
  }
  if (k == 1) {
    for (int v = 0; v < 3; ++v) {
      phi[v] = l[v];
      SETG2(v, 1.0, v, 0.0, v);
    }
    return;
  }
  if (k == 2) {
    for (int v = 0; v < 3; ++v, ++n) {
      phi[n] = l[v] * (2.0 * l[v] - 1.0);
      SETG2(n, 4.0 * l[v] - 1.0, v, 0.0, v);
    }
    for (int e = 0; e < 3; ++e, ++n) {
      int a = TRI_E[e][0], b = TRI_E[e][1];
      phi[n] = 4.0 * l[a] * l[b];
      SETG2(n, 4.0 * l[b], a, 4.0 * l[a], b);
    }
    return;
  }
  /* k == 3, equispaced */
  for (int v = 0; v < 3; ++v, ++n) {
    double L = l[v];
    phi[n] = 0.5 * L * (3.0 * L - 1.0) * (3.0 * L - 2.0);
    SETG2(n, 0.5 * (27.0 * L * L - 18.0 * L + 2.0), v, 0.0, v);
  }
  for (int e = 0; e < 3; ++e) {
    int a = TRI_E[e][0], b = TRI_E[e][1];
    double la = l[a], lb = l[b];
    phi[n] = 4.5 * la * lb * (3.0 * la - 1.0);
    SETG2(n, 4.5 * lb * (6.0 * la - 1.0), a, 4.5 * la * (3.0 * la - 1.0), b);
    ++n;
    phi[n] = 4.5 * la * lb * (3.0 * lb - 1.0);
    SETG2(n, 4.5 * lb * (3.0 * lb - 1.0), a, 4.5 * la * (6.0 * lb - 1.0), b);
    ++n;
  }
  phi[n] = 27.0 * l[0] * l[1] * l[2];
  for (int d = 0; d < 2; ++d)
    dphi[3 * n + d] = 27.0 * (l[1] * l[2] * DLAM2[0][d] + l[0] * l[2] * DLAM2[1][d] +
                              l[0] * l[1] * DLAM2[2][d]);
  dphi[3 * n + 2] = 0.0;
#undef SETG2
}

/* P_k on the reference tetrahedron: values phi[nd], gradients dphi[nd][3]. */
static void tet_basis(int k, const double* X, double* phi, double* dphi)
{
  double l[4] = {1.0 - X[0] - X[1] - X[2], X[0], X[1], X[2]};
  int n = 0;
#define SETG(idx, c0, v0, c1, v1)                                   \
  for (int d = 0; d < 3; ++d) dphi[3 * (idx) + d] = (c0) * DLAM[v0][d] + (c1) * DLAM[v1][d]
  if (k == 0) {
    phi[0] = 1.0;
    dphi[0] = dphi[1] = dphi[2] = 0.0;
    return;
  }
  if (k == 1) {
    for (int v = 0; v < 4; ++v) {
      phi[v] = l[v];
      for (int d = 0; d < 3; ++d) dphi[3 * v + d] = DLAM[v][d];
    }
    return;
  }
  if (k == 2) {
    for (int v = 0; v < 4; ++v, ++n) {
      phi[n] = l[v] * (2.0 * l[v] - 1.0);
      SETG(n, 4.0 * l[v] - 1.0, v, 0.0, v);
    }
    for (int e = 0; e < 6; ++e, ++n) {
      int a = TET_E[e][0], b = TET_E[e][1];
      phi[n] = 4.0 * l[a] * l[b];
      SETG(n, 4.0 * l[b], a, 4.0 * l[a], b);
    }
    return;
  }
  /* k == 3, equispaced */
  for (int v = 0; v < 4; ++v, ++n) {
    double L = l[v];
    phi[n] = 0.5 * L * (3.0 * L - 1.0) * (3.0 * L - 2.0);
    SETG(n, 0.5 * (27.0 * L * L - 18.0 * L + 2.0), v, 0.0, v);
  }
  for (int e = 0; e < 6; ++e) {
    int a = TET_E[e][0], b = TET_E[e][1];
    double la = l[a], lb = l[b];
    /* node nearer a: 9/2 la lb (3 la - 1) */
    phi[n] = 4.5 * la * lb * (3.0 * la - 1.0);
    SETG(n, 4.5 * lb * (6.0 * la - 1.0), a, 4.5 * la * (3.0 * la - 1.0), b);
    ++n;
    /* node nearer b: 9/2 la lb (3 lb - 1) */
    phi[n] = 4.5 * la * lb * (3.0 * lb - 1.0);
    SETG(n, 4.5 * lb * (3.0 * lb - 1.0), a, 4.5 * la * (6.0 * lb - 1.0), b);
    ++n;
  }
  for (int f = 0; f < 4; ++f, ++n) {
    int a = TET_F[f][0], b = TET_F[f][1], c = TET_F[f][2];
    phi[n] = 27.0 * l[a] * l[b] * l[c];
    for (int d = 0; d < 3; ++d)
      dphi[3 * n + d] = 27.0 * (l[b] * l[c] * DLAM[a][d] + l[a] * l[c] * DLAM[b][d] +
                                l[a] * l[b] * DLAM[c][d]);
  }
#undef SETG
}

static void gll_nodes(int k, double* t)
{
  if (k == 1) { t[0] = 0.0; t[1] = 1.0; }
  else if (k == 2) { t[0] = 0.0; t[1] = 0.5; t[2] = 1.0; }
  else { double s = 0.5 / sqrt(5.0); t[0] = 0.0; t[1] = 0.5 - s; t[2] = 0.5 + s; t[3] = 1.0; }
}

/* 1D Lagrange polynomials on the GLL nodes and their derivatives. */
static void lagrange_1d(int k, double x, double* L, double* dL)
{
  double t[4];
  gll_nodes(k, t);
  for (int a = 0; a <= k; ++a) {
    double v = 1.0, dv = 0.0;
    for (int b = 0; b <= k; ++b) {
      if (b == a) continue;
      double den = t[a] - t[b];
      dv = dv * (x - t[b]) / den + v / den;
      v *= (x - t[b]) / den;
    }
    L[a] = v;
    dL[a] = dv;
  }
}

/* Q_k on the reference hexahedron, lexicographic nodes. */
static void hex_basis(int k, const double* X, double* phi, double* dphi)
{
  double Lx[4], Ly[4], Lz[4], dx[4], dy[4], dz[4];
  lagrange_1d(k, X[0], Lx, dx);
  lagrange_1d(k, X[1], Ly, dy);
  lagrange_1d(k, X[2], Lz, dz);
  int n = 0;
  for (int c = 0; c <= k; ++c)
    for (int b = 0; b <= k; ++b)
      for (int a = 0; a <= k; ++a, ++n) {
        phi[n] = Lx[a] * Ly[b] * Lz[c];
        dphi[3 * n] = dx[a] * Ly[b] * Lz[c];
        dphi[3 * n + 1] = Lx[a] * dy[b] * Lz[c];
        dphi[3 * n + 2] = Lx[a] * Ly[b] * dz[c];
      }
}

void or_basis(int cell, int k, const double* X, double* phi, double* dphi)
{
  if (cell == 0) tet_basis(k, X, phi, dphi);
  else if (cell == 2) tri_basis(k, X, phi, dphi);
  else hex_basis(k, X, phi, dphi);
}

/* ------------------------------------------------------------------ */
/* geometry                                                            */
/* ------------------------------------------------------------------ */

static void facet_point(int lf, double s, double t, double* X);

/* Jacobian J[i][j] = dx_i/dX_j at X; returns det J, fills K = J^{-1}. */
static double jacobian(int cell, const double* xc, const double* X, double K[3][3])
{
  double J[3][3] = {{0}};
  if (cell == 0) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) J[i][j] = xc[3 * (j + 1) + i] - xc[i];
  } else if (cell == 2) {
    /* the 2 x 2 map in the (x, y) plane, with J[2][2] = 1: det J is the 2D
     * determinant and K = J^{-1} keeps the third direction unchanged */
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) J[i][j] = xc[3 * (j + 1) + i] - xc[i];
    J[2][2] = 1.0;
  } else {
    double p[8], dp[24];
    hex_basis(1, X, p, dp);
    for (int v = 0; v < 8; ++v)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) J[i][j] += xc[3 * v + i] * dp[3 * v + j];
  }
  double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  /* inverse by cofactors: K[r][c] = C[c][r] / det */
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      int r1 = (c + 1) % 3, r2 = (c + 2) % 3, c1 = (r + 1) % 3, c2 = (r + 2) % 3;
      K[r][c] = (J[r1][c1] * J[r2][c2] - J[r1][c2] * J[r2][c1]) / det;
    }
  return det;
}

/* ------------------------------------------------------------------ */
/* element kernels                                                     */
/* ------------------------------------------------------------------ */

enum {
  OR_STIFFNESS = 0,   /* kappa grad u.grad v - k2 u v            */
  OR_ELASTICITY = 1,  /* 2 mu eps:eps + lam div div               */
  OR_SOURCE = 2,      /* f . v                                    */
  OR_FACET_SOURCE = 3,/* g . v on exterior facets                 */
  OR_DG_CELL = 4,     /* p-Laplacian cell Jacobian / residual      */
  OR_DG_EXT = 5,      /* Robin exterior facet                     */
  OR_DG_INT = 6,      /* SIPG interior facet                      */
  OR_FN_CELL = 7,     /* functional: exp(u) kappa dx              */
  OR_FN_EXT = 8,      /* functional: u kappa ds                   */
  OR_FN_INT = 9,      /* functional: avg(v^2) avg(kappa) dS       */
  OR_MASS = 10,       /* kappa u v, P_degree test x P_f_degree trial */
  OR_NL_CELL = 11,    /* (exp(u) kappa p + grad u . grad p), its Jacobian */
  OR_NL_EXT = 12,     /* u kappa p (grad u . n), its Jacobian         */
  OR_NL_INT = 13      /* avg(q) avg(kappa) avg(v^2), its Jacobian     */
};

typedef struct {
  int kind;
  int cell;        /* 0 tetrahedron, 1 hexahedron */
  int degree;
  int bs;
  int rank;        /* 0 functional, 1 vector, 2 matrix */
  int qm;          /* quadrature points per direction */
  int kappa_src;   /* 0 value, 1 w[off], 2 c[off] */
  int kappa_off;
  double kappa_val;
  int mu_src, mu_off;
  double mu_val;
  int lam_src, lam_off;
  double lam_val;
  int f_src;       /* 2: constant c[f_off + a]; 1: coefficient w[f_off + l*bs + a] of degree f_degree */
  int f_off;
  int f_degree;
  double k2;
  /* DG / p-Laplacian parameters */
  double p, eps, alpha, penalty, fval;
  int u_off;       /* offset of the DG1 state u in w (per cell) */
  int coeff_width; /* per-cell packed width (for interior-facet interlacing) */
  double mu0;      /* constant part of the p-Laplacian coefficient */
  int diag_part;   /* rank 1 from a bilinear kind: the element matrix's diagonal */
  int kappa_degree;/* degree of kappa where it differs from the test space's */
} or_kernel;

static double param(int src, int off, double val, const double* w, const double* c)
{
  if (src == 1) return w[off];
  if (src == 2) return c[off];
  return val;
}

static int quad_cell(const or_kernel* K, double* pts, double* wts)
{
  return K->cell == 0 ? or_tet_rule(K->qm, pts, wts) : or_hex_rule(K->qm, pts, wts);
}

static void tab_cell(const or_kernel* K, double* A, const double* w, const double* c,
                     const double* xc)
{
  static _Thread_local double pts[3 * OR_MAXQ], wts[OR_MAXQ];
  const int nd = or_ndofs(K->cell, K->degree), bs = K->bs, n = nd * bs;
  const int nq = quad_cell(K, pts, wts);
  double phi[OR_MAXND], dref[3 * OR_MAXND], g[OR_MAXND][3], Kinv[3][3];
  for (int q = 0; q < nq; ++q) {
    const double* X = &pts[3 * q];
    double det = jacobian(K->cell, xc, X, Kinv);
    double dx = wts[q] * fabs(det);
    or_basis(K->cell, K->degree, X, phi, dref);
    for (int i = 0; i < nd; ++i)
      for (int d = 0; d < 3; ++d)
        g[i][d] = dref[3 * i] * Kinv[0][d] + dref[3 * i + 1] * Kinv[1][d] +
                  dref[3 * i + 2] * Kinv[2][d];
    if (K->kind == OR_STIFFNESS) {
      double kap = param(K->kappa_src, K->kappa_off, K->kappa_val, w, c);
      for (int i = 0; i < nd; ++i)
        for (int j = 0; j < nd; ++j)
          A[i * nd + j] += dx * (kap * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]) -
                                 K->k2 * phi[i] * phi[j]);
    } else if (K->kind == OR_ELASTICITY) {
      double mu = param(K->mu_src, K->mu_off, K->mu_val, w, c);
      double lam = param(K->lam_src, K->lam_off, K->lam_val, w, c);
      for (int i = 0; i < nd; ++i)
        for (int a = 0; a < 3; ++a)
          for (int j = 0; j < nd; ++j)
            for (int b = 0; b < 3; ++b) {
              double v = lam * g[i][a] * g[j][b] + mu * g[i][b] * g[j][a];
              if (a == b) v += mu * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
              A[(3 * i + a) * n + 3 * j + b] += dx * v;
            }
    } else if (K->kind == OR_SOURCE) {
      double f[3] = {0, 0, 0};
      if (K->f_src == 2) {
        for (int a = 0; a < bs; ++a) f[a] = c[K->f_off + a];
      } else {
        double psi[OR_MAXND], dpsi[3 * OR_MAXND];
        or_basis(K->cell, K->f_degree, X, psi, dpsi);
        int nf = or_ndofs(K->cell, K->f_degree);
        for (int l = 0; l < nf; ++l)
          for (int a = 0; a < bs; ++a) f[a] += psi[l] * w[K->f_off + l * bs + a];
      }
      for (int i = 0; i < nd; ++i)
        for (int a = 0; a < bs; ++a) A[i * bs + a] += dx * phi[i] * f[a];
    }
  }
}

static void tab_ext_facet(const or_kernel* K, double* A, const double* w, const double* c,
                          const double* xc, int lf)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND];
  const int nd = or_ndofs(K->cell, K->degree), bs = K->bs;
  const int nq = or_tri_rule(K->qm, pts, wts);
  const int* f = TET_F[lf];
  double e1[3], e2[3];
  for (int d = 0; d < 3; ++d) {
    e1[d] = xc[3 * f[1] + d] - xc[3 * f[0] + d];
    e2[d] = xc[3 * f[2] + d] - xc[3 * f[0] + d];
  }
  double nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
  double ds = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  for (int q = 0; q < nq; ++q) {
    double X[3], phi[OR_MAXND], dref[3 * OR_MAXND];
    for (int d = 0; d < 3; ++d)
      X[d] = TET_V[f[0]][d] + pts[2 * q] * (TET_V[f[1]][d] - TET_V[f[0]][d]) +
             pts[2 * q + 1] * (TET_V[f[2]][d] - TET_V[f[0]][d]);
    tet_basis(K->degree, X, phi, dref);
    if (K->kind == OR_FACET_SOURCE)
      for (int i = 0; i < nd; ++i)
        for (int a = 0; a < bs; ++a) A[i * bs + a] += wts[q] * ds * phi[i] * c[K->f_off + a];
  }
}

/* ------------------------------------------------------------------ */
/* SIPG p-Laplacian with Robin boundary (DG_k on affine tetrahedra)     */
/* ------------------------------------------------------------------ */

static double diameter(const double* xc)
{
  double h = 0.0;
  for (int a = 0; a < 4; ++a)
    for (int b = a + 1; b < 4; ++b) {
      double s = 0.0;
      for (int d = 0; d < 3; ++d) s += (xc[3 * a + d] - xc[3 * b + d]) * (xc[3 * a + d] - xc[3 * b + d]);
      if (sqrt(s) > h) h = sqrt(s);
    }
  return h;
}

/* Size of facet lf (twice the triangle's area on a tetrahedron, the edge's
 * length on a triangle) and its unit normal pointing away from the opposite
 * vertex. */
static double facet_normal_c(int cell, const double* xc, int lf, double* n)
{
  if (cell == 2) {
    const int a = TRI_E[lf][0], b = TRI_E[lf][1];
    const double tx = xc[3 * b] - xc[3 * a], ty = xc[3 * b + 1] - xc[3 * a + 1];
    const double ox = xc[3 * lf] - xc[3 * a], oy = xc[3 * lf + 1] - xc[3 * a + 1];
    const double ds = sqrt(tx * tx + ty * ty);
    double sgn = (ty * ox - tx * oy) > 0.0 ? -1.0 : 1.0;
    n[0] = sgn * ty / ds;
    n[1] = -sgn * tx / ds;
    n[2] = 0.0;
    return ds;
  }
  const int* f = TET_F[lf];
  double e1[3], e2[3], o[3];
  for (int d = 0; d < 3; ++d) {
    e1[d] = xc[3 * f[1] + d] - xc[3 * f[0] + d];
    e2[d] = xc[3 * f[2] + d] - xc[3 * f[0] + d];
    o[d] = xc[3 * lf + d] - xc[3 * f[0] + d];
  }
  n[0] = e1[1] * e2[2] - e1[2] * e2[1];
  n[1] = e1[2] * e2[0] - e1[0] * e2[2];
  n[2] = e1[0] * e2[1] - e1[1] * e2[0];
  double ds = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  double sgn = (n[0] * o[0] + n[1] * o[1] + n[2] * o[2]) > 0.0 ? -1.0 : 1.0;
  for (int d = 0; d < 3; ++d) n[d] *= sgn / ds;
  return ds;
}

static double facet_normal(const double* xc, int lf, double* n) { return facet_normal_c(0, xc, lf, n); }

/* reference point of facet lf at facet coordinates (s, t) (t unused on a
 * triangle's edge, which runs from its first to its second vertex) */
static void facet_point_c(int cell, int lf, double s, double t, double* X)
{
  if (cell == 2) {
    const int a = TRI_E[lf][0], b = TRI_E[lf][1];
    for (int d = 0; d < 2; ++d) X[d] = TRI_V[a][d] + s * (TRI_V[b][d] - TRI_V[a][d]);
    X[2] = 0.0;
    return;
  }
  facet_point(lf, s, t, X);
}

/* reference point of facet lf at triangle coordinates (s, t) */
static void facet_point(int lf, double s, double t, double* X)
{
  const int* f = TET_F[lf];
  for (int d = 0; d < 3; ++d)
    X[d] = TET_V[f[0]][d] + s * (TET_V[f[1]][d] - TET_V[f[0]][d]) + t * (TET_V[f[2]][d] - TET_V[f[0]][d]);
}

/* Interior facets: the reference point of the second cell (vertices xc1) at
 * the physical position of reference point X0 of the first cell (xc0), by the
 * affine maps x = x0 + J X: X1 = J1^{-1} (x - x1_0).  The two sides of a
 * facet are paired through physical space, whatever order either cell lists
 * the facet's vertices in (the engine's bodies pair them through facet
 * quadrature permutations instead; CUDAFormIntegral.cpp:1913-1920 gathers
 * those per facet). */
static void other_side_point_c(int cell, const double* xc0, const double* X0, const double* xc1,
                               double* X1)
{
  const int dim = cell == 2 ? 2 : 3;
  double x[3], Kinv[3][3];
  for (int i = 0; i < 3; ++i) {
    x[i] = xc0[i];
    for (int j = 0; j < dim; ++j) x[i] += (xc0[3 * (j + 1) + i] - xc0[i]) * X0[j];
  }
  jacobian(cell, xc1, X0, Kinv);
  for (int r = 0; r < 3; ++r) {
    X1[r] = 0.0;
    for (int c = 0; c < 3; ++c) X1[r] += Kinv[r][c] * (x[c] - xc1[c]);
  }
  if (dim == 2) X1[2] = 0.0;
}

static void other_side_point(const double* xc0, const double* X0, const double* xc1, double* X1)
{
  other_side_point_c(0, xc0, X0, xc1, X1);
}

/* physical basis gradients of a P_k tetrahedron at X */
static void phys_grad(const or_kernel* K, const double* xc, const double* X, double* phi,
                      double g[][3], double* adet)
{
  double dref[3 * OR_MAXND], Kinv[3][3];
  double det = jacobian(K->cell, xc, X, Kinv);
  if (adet) *adet = fabs(det);
  or_basis(K->cell, K->degree, X, phi, dref);
  int nd = or_ndofs(K->cell, K->degree);
  for (int i = 0; i < nd; ++i)
    for (int d = 0; d < 3; ++d)
      g[i][d] = dref[3 * i] * Kinv[0][d] + dref[3 * i + 1] * Kinv[1][d] + dref[3 * i + 2] * Kinv[2][d];
}

static void tab_dg_cell(const or_kernel* K, double* A, const double* w, const double* xc)
{
  static _Thread_local double pts[3 * OR_MAXQ], wts[OR_MAXQ];
  const int nd = or_ndofs(0, K->degree);
  const int nq = or_tet_rule(K->qm, pts, wts);
  double phi[OR_MAXND], g[OR_MAXND][3], adet;
  for (int q = 0; q < nq; ++q) {
    phys_grad(K, xc, &pts[3 * q], phi, g, &adet);
    double dx = wts[q] * adet, gu[3] = {0, 0, 0};
    for (int i = 0; i < nd; ++i)
      for (int d = 0; d < 3; ++d) gu[d] += w[K->u_off + i] * g[i][d];
    double s = K->eps * K->eps + gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2];
    double mu = K->mu0 + pow(s, 0.5 * (K->p - 2.0)), dmu = (K->p - 2.0) * pow(s, 0.5 * (K->p - 4.0));
    if (K->rank == 0) {
      /* energy density whose first variation is the residual below */
      double uq = 0.0, g2 = gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2];
      for (int i = 0; i < nd; ++i) uq += w[K->u_off + i] * phi[i];
      A[0] += dx * (pow(s, 0.5 * K->p) / K->p + 0.5 * K->mu0 * g2 - K->fval * uq);
      continue;
    }
    for (int i = 0; i < nd; ++i) {
      double gi = gu[0] * g[i][0] + gu[1] * g[i][1] + gu[2] * g[i][2];
      if (K->rank == 1) {
        A[i] += dx * (mu * gi - K->fval * phi[i]);
        continue;
      }
      for (int j = 0; j < nd; ++j) {
        double gj = gu[0] * g[j][0] + gu[1] * g[j][1] + gu[2] * g[j][2];
        A[i * nd + j] += dx * (mu * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]) +
                               dmu * gi * gj);
      }
    }
  }
}

static void tab_dg_ext(const or_kernel* K, double* A, const double* w, const double* xc, int lf)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], nrm[3];
  const int nd = or_ndofs(0, K->degree);
  const int nq = or_tri_rule(K->qm, pts, wts);
  double ds = facet_normal(xc, lf, nrm);
  for (int q = 0; q < nq; ++q) {
    double X[3], phi[OR_MAXND], dref[3 * OR_MAXND];
    facet_point(lf, pts[2 * q], pts[2 * q + 1], X);
    tet_basis(K->degree, X, phi, dref);
    double fw = wts[q] * ds * K->alpha, uq = 0.0;
    for (int j = 0; j < nd; ++j) uq += w[K->u_off + j] * phi[j];
    if (K->rank == 0) { A[0] += 0.5 * fw * uq * uq; continue; }
    for (int i = 0; i < nd; ++i) {
      if (K->rank == 1) { A[i] += fw * uq * phi[i]; continue; }
      for (int j = 0; j < nd; ++j) A[i * nd + j] += fw * phi[i] * phi[j];
    }
  }
}

/* interior facet: rows/columns [cell 0 dofs, cell 1 dofs]; w interlaced (u0, u1) */
static void tab_dg_int(const or_kernel* K, double* A, const double* w, const double* xc,
                       const int* eli)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], nrm[3];
  const int nd = or_ndofs(0, K->degree), n2 = 2 * nd;
  const int nq = or_tri_rule(K->qm, pts, wts);
  const double* xc1 = xc + 12;
  double ds = facet_normal(xc, eli[0], nrm);
  double sig = K->penalty / (0.5 * (diameter(xc) + diameter(xc1)));
  for (int q = 0; q < nq; ++q) {
    double X0[3], X1[3], phi0[OR_MAXND], phi1[OR_MAXND], g0[OR_MAXND][3], g1[OR_MAXND][3];
    double jmp[2 * OR_MAXND], dn[2 * OR_MAXND];
    facet_point(eli[0], pts[2 * q], pts[2 * q + 1], X0);
    other_side_point(xc, X0, xc + 12, X1);
    phys_grad(K, xc, X0, phi0, g0, NULL);
    phys_grad(K, xc1, X1, phi1, g1, NULL);
    for (int i = 0; i < nd; ++i) {
      jmp[i] = phi0[i];
      jmp[nd + i] = -phi1[i];
      dn[i] = 0.5 * (g0[i][0] * nrm[0] + g0[i][1] * nrm[1] + g0[i][2] * nrm[2]);
      dn[nd + i] = 0.5 * (g1[i][0] * nrm[0] + g1[i][1] * nrm[1] + g1[i][2] * nrm[2]);
    }
    double fw = wts[q] * ds;
    if (K->rank == 0) {
      const double* u = w + 2 * K->u_off;
      double ju = 0.0, du = 0.0;
      for (int j = 0; j < n2; ++j) { ju += jmp[j] * u[j]; du += dn[j] * u[j]; }
      A[0] += fw * (0.5 * sig * ju * ju - du * ju);
      continue;
    }
    if (K->rank == 1) {
      const double* u = w + 2 * K->u_off;
      double ju = 0.0, du = 0.0;
      for (int j = 0; j < n2; ++j) { ju += jmp[j] * u[j]; du += dn[j] * u[j]; }
      for (int i = 0; i < n2; ++i) A[i] += fw * (-du * jmp[i] - dn[i] * ju + sig * ju * jmp[i]);
      continue;
    }
    for (int i = 0; i < n2; ++i)
      for (int j = 0; j < n2; ++j)
        A[i * n2 + j] += fw * (-dn[j] * jmp[i] - dn[i] * jmp[j] + sig * jmp[i] * jmp[j]);
  }
}

/* ------------------------------------------------------------------ */
/* a functional over all three integral types (the mixed functional of  */
/* python/test/test_cuda_assembly.py:66-71): u, kappa in P_degree       */
/* (offsets u_off, kappa_off), v in DG_f_degree (offset f_off)          */
/* ------------------------------------------------------------------ */

static double eval_at_c(int cell, int k, const double* X, const double* coef)
{
  double phi[OR_MAXND], dref[3 * OR_MAXND];
  or_basis(cell, k, X, phi, dref);
  double v = 0.0;
  for (int i = 0; i < or_ndofs(cell, k); ++i) v += coef[i] * phi[i];
  return v;
}

/* cell rule (points at stride 3) and facet rule (points (s, t) at stride 2)
 * of a simplex cell: tetrahedron, or triangle with Gauss-Legendre on edges */
static int cell_rule(int cell, int m, double* pts, double* wts)
{
  if (cell != 2) return or_tet_rule(m, pts, wts);
  double p2[2 * OR_MAXQ];
  const int n = or_tri_rule(m, p2, wts);
  for (int q = n - 1; q >= 0; --q) {
    pts[3 * q] = p2[2 * q];
    pts[3 * q + 1] = p2[2 * q + 1];
    pts[3 * q + 2] = 0.0;
  }
  return n;
}

static int facet_rule(int cell, int m, double* pts, double* wts)
{
  if (cell != 2) return or_tri_rule(m, pts, wts);
  double t[OR_MAXND];
  or_gauss_jacobi(m, 0.0, t, wts);
  for (int q = 0; q < m; ++q) {
    pts[2 * q] = t[q];
    pts[2 * q + 1] = 0.0;
  }
  return m;
}

static int cell_nv(int cell) { return cell == 2 ? 3 : (cell == 0 ? 4 : 8); }

static void tab_fn_cell(const or_kernel* K, double* A, const double* w, const double* xc)
{
  static _Thread_local double pts[3 * OR_MAXQ], wts[OR_MAXQ];
  const int nq = cell_rule(K->cell, K->qm, pts, wts);
  double Kinv[3][3];
  for (int q = 0; q < nq; ++q) {
    const double* X = &pts[3 * q];
    double dx = wts[q] * fabs(jacobian(K->cell, xc, X, Kinv));
    A[0] += dx * exp(eval_at_c(K->cell, K->degree, X, w + K->u_off)) * eval_at_c(K->cell, K->degree, X, w + K->kappa_off);
  }
}

static void tab_fn_ext(const or_kernel* K, double* A, const double* w, const double* xc, int lf)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], nrm[3];
  const int nq = facet_rule(K->cell, K->qm, pts, wts);
  double ds = facet_normal_c(K->cell, xc, lf, nrm);
  for (int q = 0; q < nq; ++q) {
    double X[3];
    facet_point_c(K->cell, lf, pts[2 * q], pts[2 * q + 1], X);
    A[0] += wts[q] * ds * eval_at_c(K->cell, K->degree, X, w + K->u_off) * eval_at_c(K->cell, K->degree, X, w + K->kappa_off);
  }
}

/* w interlaced per coefficient: (cell 0 values, cell 1 values) at twice the offset */
static void tab_fn_int(const or_kernel* K, double* A, const double* w, const double* xc,
                       const int* eli)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], nrm[3];
  const int nq = facet_rule(K->cell, K->qm, pts, wts);
  const int np = or_ndofs(K->cell, K->degree), nv = or_ndofs(K->cell, K->f_degree);
  double ds = facet_normal_c(K->cell, xc, eli[0], nrm);
  const double* k0 = w + 2 * K->kappa_off;
  const double* v0 = w + 2 * K->f_off;
  for (int q = 0; q < nq; ++q) {
    double X0[3], X1[3];
    facet_point_c(K->cell, eli[0], pts[2 * q], pts[2 * q + 1], X0);
    other_side_point_c(K->cell, xc, X0, xc + 3 * cell_nv(K->cell), X1);
    double va = eval_at_c(K->cell, K->f_degree, X0, v0), vb = eval_at_c(K->cell, K->f_degree, X1, v0 + nv);
    double ka = eval_at_c(K->cell, K->degree, X0, k0), kb = eval_at_c(K->cell, K->degree, X1, k0 + np);
    A[0] += wts[q] * ds * 0.5 * (va * va + vb * vb) * 0.5 * (ka + kb);
  }
}

/* kappa u v with v in P_degree (rows) and u in P_f_degree (columns) */
static void tab_mass(const or_kernel* K, double* A, const double* w, const double* c,
                     const double* xc)
{
  static _Thread_local double pts[3 * OR_MAXQ], wts[OR_MAXQ];
  const int nq = or_tet_rule(K->qm, pts, wts);
  const int nv = or_ndofs(0, K->degree), nw = or_ndofs(0, K->f_degree);
  double Kinv[3][3], phi[OR_MAXND], psi[OR_MAXND], d[3 * OR_MAXND];
  const double kap = param(K->kappa_src, K->kappa_off, K->kappa_val, w, c);
  for (int q = 0; q < nq; ++q) {
    const double* X = &pts[3 * q];
    double dx = wts[q] * fabs(jacobian(0, xc, X, Kinv)) * kap;
    tet_basis(K->degree, X, phi, d);
    tet_basis(K->f_degree, X, psi, d);
    for (int i = 0; i < nv; ++i)
      for (int j = 0; j < nw; ++j) A[i * nw + j] += dx * phi[i] * psi[j];
  }
}

/* ------------------------------------------------------------------ */
/* the nonlinear forms of python/test/test_cuda_assembly.py:66-81       */
/* (u at u_off, kappa at kappa_off, both P_degree; v in DG_f_degree)     */
/* ------------------------------------------------------------------ */

static void tab_nl_cell(const or_kernel* K, double* A, const double* w, const double* xc)
{
  static _Thread_local double pts[3 * OR_MAXQ], wts[OR_MAXQ];
  const int nd = or_ndofs(K->cell, K->degree);
  const int nq = cell_rule(K->cell, K->qm, pts, wts);
  double phi[OR_MAXND], g[OR_MAXND][3], adet;
  for (int q = 0; q < nq; ++q) {
    phys_grad(K, xc, &pts[3 * q], phi, g, &adet);
    double uq = 0.0, kq = 0.0, gu[3] = {0, 0, 0};
    for (int i = 0; i < nd; ++i) {
      uq += w[K->u_off + i] * phi[i];
      kq += w[K->kappa_off + i] * phi[i];
      for (int d = 0; d < 3; ++d) gu[d] += w[K->u_off + i] * g[i][d];
    }
    const double dx = wts[q] * adet, ek = exp(uq) * kq;
    for (int i = 0; i < nd; ++i) {
      if (K->rank == 1) {
        A[i] += dx * (ek * phi[i] + gu[0] * g[i][0] + gu[1] * g[i][1] + gu[2] * g[i][2]);
        continue;
      }
      for (int j = 0; j < nd; ++j)
        A[i * nd + j] += dx * (ek * phi[j] * phi[i] + g[j][0] * g[i][0] + g[j][1] * g[i][1] +
                               g[j][2] * g[i][2]);
    }
  }
}

static void tab_nl_ext(const or_kernel* K, double* A, const double* w, const double* xc, int lf)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], n[3];
  const int nd = or_ndofs(K->cell, K->degree);
  const int nq = facet_rule(K->cell, K->qm, pts, wts);
  const double ds = facet_normal_c(K->cell, xc, lf, n);
  for (int q = 0; q < nq; ++q) {
    double X[3], phi[OR_MAXND], g[OR_MAXND][3];
    facet_point_c(K->cell, lf, pts[2 * q], pts[2 * q + 1], X);
    phys_grad(K, xc, X, phi, g, NULL);
    double uq = 0.0, kq = 0.0, gun = 0.0;
    for (int i = 0; i < nd; ++i) {
      uq += w[K->u_off + i] * phi[i];
      kq += w[K->kappa_off + i] * phi[i];
      gun += w[K->u_off + i] * (g[i][0] * n[0] + g[i][1] * n[1] + g[i][2] * n[2]);
    }
    const double fw = wts[q] * ds;
    for (int i = 0; i < nd; ++i) {
      if (K->rank == 1) { A[i] += fw * uq * kq * phi[i] * gun; continue; }
      for (int j = 0; j < nd; ++j) {
        const double gjn = g[j][0] * n[0] + g[j][1] * n[1] + g[j][2] * n[2];
        A[i * nd + j] += fw * kq * phi[i] * (phi[j] * gun + uq * gjn);
      }
    }
  }
}

/* interlaced w: kappa (P_kappa_degree) at 2 kappa_off, v (DG_degree) at 2 f_off */
static void tab_nl_int(const or_kernel* K, double* A, const double* w, const double* xc,
                       const int* eli)
{
  double pts[2 * OR_MAXND * OR_MAXND], wts[OR_MAXND * OR_MAXND], n[3];
  const int nd = or_ndofs(K->cell, K->degree), n2 = 2 * nd;
  const int nc = or_ndofs(K->cell, K->kappa_degree);
  const int nq = facet_rule(K->cell, K->qm, pts, wts);
  const double ds = facet_normal_c(K->cell, xc, eli[0], n);
  const double* k0 = w + 2 * K->kappa_off;
  const double* v0 = w + 2 * K->f_off;
  for (int q = 0; q < nq; ++q) {
    double X0[3], X1[3], pa[OR_MAXND], pb[OR_MAXND], d[3 * OR_MAXND];
    facet_point_c(K->cell, eli[0], pts[2 * q], pts[2 * q + 1], X0);
    other_side_point_c(K->cell, xc, X0, xc + 3 * cell_nv(K->cell), X1);
    const double kav = 0.5 * (eval_at_c(K->cell, K->kappa_degree, X0, k0) + eval_at_c(K->cell, K->kappa_degree, X1, k0 + nc));
    const double ua = eval_at_c(K->cell, K->degree, X0, v0), ub = eval_at_c(K->cell, K->degree, X1, v0 + nd);
    or_basis(K->cell, K->degree, X0, pa, d);
    or_basis(K->cell, K->degree, X1, pb, d);
    const double fw = wts[q] * ds;
    for (int I = 0; I < n2; ++I) {
      const double pI = 0.5 * (I < nd ? pa[I] : pb[I - nd]);
      if (K->rank == 1) { A[I] += fw * pI * kav * 0.5 * (ua * ua + ub * ub); continue; }
      for (int J = 0; J < n2; ++J)
        A[I * n2 + J] += fw * pI * kav * (J < nd ? ua * pa[J] : ub * pb[J - nd]);
    }
  }
}

void or_tabulate(const or_kernel* K, double* A, const double* w, const double* c,
                 const double* coordinate_dofs, const int* entity_local_index);

/* UFCx-shaped entry point: A += element tensor. */
void or_tabulate(const or_kernel* K, double* A, const double* w, const double* c,
                 const double* coordinate_dofs, const int* entity_local_index)
{
  if (K->diag_part) {
    /* the diagonal of the kind's element matrix */
    static _Thread_local double T[4 * OR_MAXND * OR_MAXND];
    or_kernel M = *K;
    M.diag_part = 0;
    M.rank = 2;
    const int n = or_ndofs(K->cell, K->degree) * K->bs * (K->kind == OR_NL_INT || K->kind == OR_DG_INT ? 2 : 1);
    memset(T, 0, sizeof(double) * n * n);
    or_tabulate(&M, T, w, c, coordinate_dofs, entity_local_index);
    for (int i = 0; i < n; ++i) A[i] += T[i * n + i];
    return;
  }
  switch (K->kind) {
  case OR_STIFFNESS:
  case OR_ELASTICITY:
  case OR_SOURCE: tab_cell(K, A, w, c, coordinate_dofs); break;
  case OR_FACET_SOURCE: tab_ext_facet(K, A, w, c, coordinate_dofs, entity_local_index[0]); break;
  case OR_DG_CELL: tab_dg_cell(K, A, w, coordinate_dofs); break;
  case OR_DG_EXT: tab_dg_ext(K, A, w, coordinate_dofs, entity_local_index[0]); break;
  case OR_DG_INT: tab_dg_int(K, A, w, coordinate_dofs, entity_local_index); break;
  case OR_FN_CELL: tab_fn_cell(K, A, w, coordinate_dofs); break;
  case OR_MASS: tab_mass(K, A, w, c, coordinate_dofs); break;
  case OR_NL_CELL: tab_nl_cell(K, A, w, coordinate_dofs); break;
  case OR_NL_EXT: tab_nl_ext(K, A, w, coordinate_dofs, entity_local_index[0]); break;
  case OR_NL_INT: tab_nl_int(K, A, w, coordinate_dofs, entity_local_index); break;
  case OR_FN_EXT: tab_fn_ext(K, A, w, coordinate_dofs, entity_local_index[0]); break;
  case OR_FN_INT: tab_fn_int(K, A, w, coordinate_dofs, entity_local_index); break;
  default: break;
  }
}

/* ------------------------------------------------------------------ */
/* coefficient packing (CUDAAssembler.cpp:59-96)                       */
/* ------------------------------------------------------------------ */

/* packed[c*W + off_j + k] = values_j[dofs_j[c*n_j + k]] for the unrolled dofmaps. */
void or_pack_coefficients(int ncoef, const int* n_per_cell, const int32_t* const* dofs,
                          const double* const* values, const int* offsets, int64_t ncells, int W,
                          double* packed)
{
  for (int j = 0; j < ncoef; ++j)
    for (int64_t cc = 0; cc < ncells; ++cc)
      for (int k = 0; k < n_per_cell[j]; ++k)
        packed[cc * W + offsets[j] + k] = values[j][dofs[j][cc * n_per_cell[j] + k]];
}

/* ------------------------------------------------------------------ */
/* sparsity                                                            */
/* ------------------------------------------------------------------ */

static int cmp_i32(const void* a, const void* b)
{
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

typedef struct {
  int type;                /* 0 cell, 1 exterior facet, 2 interior facet */
  const int32_t* entities; /* 1, 2 or 4 ints per entity */
  int64_t n;
} or_entities;

static int stride_of(int type) { return type == 0 ? 1 : (type == 1 ? 2 : 4); }

/*
 * Blocked sparsity of the integrals (every (row node, col node) pair of the
 * entity's cell(s); interior facets couple the union of both cells), sorted
 * and unique per row, then unrolled to scalar rows bs0*r + a with columns
 * bs1*c + b (la/petsc.cpp:38-65).  Outputs are malloc'd; free with or_free.
 */
int or_sparsity(int nint, const or_entities* ints, const int32_t* dofmap0, int nd0, int bs0,
                const int32_t* dofmap1, int nd1, int bs1, int64_t nrows_b, int32_t** row_ptr_out,
                int32_t** cols_out, int64_t* nnz_out, int32_t** rpb_out, int32_t** colsb_out)
{
  int64_t* cnt = calloc(nrows_b + 1, sizeof(int64_t));
  for (int it = 0; it < nint; ++it) {
    int s = stride_of(ints[it].type);
    int ncell = ints[it].type == 2 ? 2 : 1;
    for (int64_t e = 0; e < ints[it].n; ++e)
      for (int ci = 0; ci < ncell; ++ci) {
        int32_t cc = ints[it].entities[e * s + 2 * ci];
        for (int i = 0; i < nd0; ++i) cnt[dofmap0[(int64_t)cc * nd0 + i] + 1] += (int64_t)ncell * nd1;
      }
  }
  for (int64_t r = 0; r < nrows_b; ++r) cnt[r + 1] += cnt[r];
  int32_t* buf = malloc((cnt[nrows_b] + 1) * sizeof(int32_t));
  int64_t* pos = malloc(nrows_b * sizeof(int64_t));
  memcpy(pos, cnt, nrows_b * sizeof(int64_t));
  for (int it = 0; it < nint; ++it) {
    int s = stride_of(ints[it].type);
    int ncell = ints[it].type == 2 ? 2 : 1;
    for (int64_t e = 0; e < ints[it].n; ++e)
      for (int ci = 0; ci < ncell; ++ci) {
        int32_t cr = ints[it].entities[e * s + 2 * ci];
        for (int i = 0; i < nd0; ++i) {
          int32_t row = dofmap0[(int64_t)cr * nd0 + i];
          for (int cj = 0; cj < ncell; ++cj) {
            int32_t ccol = ints[it].entities[e * s + 2 * cj];
            for (int j = 0; j < nd1; ++j) buf[pos[row]++] = dofmap1[(int64_t)ccol * nd1 + j];
          }
        }
      }
  }
  int32_t* rpb = malloc((nrows_b + 1) * sizeof(int32_t));
  int64_t out = 0;
  rpb[0] = 0;
  for (int64_t r = 0; r < nrows_b; ++r) {
    int64_t b0 = cnt[r], b1 = cnt[r + 1];
    qsort(buf + b0, b1 - b0, sizeof(int32_t), cmp_i32);
    int64_t start = out;
    for (int64_t k = b0; k < b1; ++k)
      if (k == b0 || buf[k] != buf[k - 1]) buf[out++] = buf[k];
    (void)start;
    rpb[r + 1] = (int32_t)out;
  }
  int64_t nnzb = out;
  int64_t nrows = nrows_b * bs0;
  int64_t nnz = nnzb * bs0 * bs1;
  int32_t* rp = malloc((nrows + 1) * sizeof(int32_t));
  int32_t* cols = malloc((nnz > 0 ? nnz : 1) * sizeof(int32_t));
  rp[0] = 0;
  int64_t w = 0;
  for (int64_t r = 0; r < nrows_b; ++r) {
    int64_t rn = rpb[r + 1] - rpb[r];
    for (int a = 0; a < bs0; ++a) {
      for (int64_t k = 0; k < rn; ++k)
        for (int b = 0; b < bs1; ++b) cols[w++] = bs1 * buf[rpb[r] + k] + b;
      rp[r * bs0 + a + 1] = (int32_t)w;
    }
  }
  int32_t* colsb = malloc((nnzb > 0 ? nnzb : 1) * sizeof(int32_t));
  memcpy(colsb, buf, nnzb * sizeof(int32_t));
  free(buf);
  free(pos);
  free(cnt);
  *row_ptr_out = rp;
  *cols_out = cols;
  *nnz_out = nnz;
  *rpb_out = rpb;
  *colsb_out = colsb;
  return 0;
}

void or_free(void* p) { free(p); }

/* ------------------------------------------------------------------ */
/* assembly                                                            */
/* ------------------------------------------------------------------ */

typedef struct {
  const double* x;          /* [Nv x 3] */
  const int32_t* cells;     /* [Nc x nv] geometry dofmap */
  int nv;
  const int32_t* dofmap0;   /* blocked [Nc x nd0] */
  int nd0, bs0;
  const int32_t* dofmap1;
  int nd1, bs1;
  const double* coeffs;     /* packed [Nc x W] */
  int W;
  const int* coeff_offsets; /* ncoef + 1 */
  int ncoef;
  const double* constants;
} or_problem;

/* binary search of key in a sorted row (CUDAFormIntegral.cpp:2572-2602) */
static int64_t find(const int32_t* row_ptr, const int32_t* cols, int32_t row, int32_t key)
{
  int64_t lo = row_ptr[row], hi = row_ptr[row + 1] - 1;
  while (lo <= hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (cols[mid] == key) return mid;
    if (cols[mid] < key) lo = mid + 1;
    else hi = mid - 1;
  }
  return -1;
}

/* Gather the element data of one entity: coordinate dofs, coefficients
 * (interlaced for interior facets), joint unrolled dofs, local indices. */
static int gather(const or_problem* P, int type, const int32_t* ent, double* xc, double* w,
                  int32_t* dofs0, int32_t* dofs1, int* eli)
{
  int ncell = type == 2 ? 2 : 1;
  for (int ci = 0; ci < ncell; ++ci) {
    int32_t cc = ent[2 * ci];
    for (int v = 0; v < P->nv; ++v)
      for (int d = 0; d < 3; ++d)
        xc[(ci * P->nv + v) * 3 + d] = P->x[3 * (int64_t)P->cells[(int64_t)cc * P->nv + v] + d];
    for (int i = 0; i < P->nd0; ++i)
      for (int a = 0; a < P->bs0; ++a)
        dofs0[(ci * P->nd0 + i) * P->bs0 + a] = P->bs0 * P->dofmap0[(int64_t)cc * P->nd0 + i] + a;
    if (P->dofmap1)
      for (int j = 0; j < P->nd1; ++j)
        for (int b = 0; b < P->bs1; ++b)
          dofs1[(ci * P->nd1 + j) * P->bs1 + b] = P->bs1 * P->dofmap1[(int64_t)cc * P->nd1 + j] + b;
  }
  if (type == 0) {
    if (P->W) memcpy(w, &P->coeffs[(int64_t)ent[0] * P->W], P->W * sizeof(double));
  } else if (type == 1) {
    if (P->W) memcpy(w, &P->coeffs[(int64_t)ent[0] * P->W], P->W * sizeof(double));
    eli[0] = ent[1];
  } else {
    const double* w0 = &P->coeffs[(int64_t)ent[0] * P->W];
    const double* w1 = &P->coeffs[(int64_t)ent[2] * P->W];
    for (int j = 0; j < P->ncoef; ++j) {
      int off = P->coeff_offsets[j], sz = P->coeff_offsets[j + 1] - off;
      for (int k = 0; k < sz; ++k) {
        w[2 * off + k] = w0[off + k];
        w[2 * off + sz + k] = w1[off + k];
      }
    }
    eli[0] = ent[1];
    eli[1] = ent[3];
  }
  return ncell;
}

#define OR_MAXT (2 * OR_MAXND * 3)

/*
 * values += A over every integral; rows with bc0 / columns with bc1 are
 * skipped; afterwards, if insert_diag, values[d,d] = diagonal for bc0 rows.
 */
int or_assemble_matrix(const or_problem* P, int nint, const or_kernel* kernels,
                       const or_entities* ints, const int8_t* bc0, const int8_t* bc1,
                       const int32_t* row_ptr, const int32_t* cols, double* values,
                       int insert_diag, double diagonal, int64_t nrows)
{
  static double A[OR_MAXT * OR_MAXT];
  double xc[2 * 27 * 3], w[1024];
  int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
  int eli[2];
  int missing = 0;
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int64_t e = 0; e < ints[it].n; ++e) {
      const int32_t* ent = &ints[it].entities[e * s];
      int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
      int nr = ncell * P->nd0 * P->bs0, nc = ncell * P->nd1 * P->bs1;
      memset(A, 0, sizeof(double) * nr * nc);
      or_tabulate(&kernels[it], A, w, P->constants, xc, eli);
      for (int i = 0; i < nr; ++i) {
        int32_t row = dofs0[i];
        if (bc0 && bc0[row]) continue;
        for (int j = 0; j < nc; ++j) {
          int32_t col = dofs1[j];
          if (bc1 && bc1[col]) continue;
          int64_t p = find(row_ptr, cols, row, col);
          if (p < 0) { ++missing; continue; }
          values[p] += A[i * nc + j];
        }
      }
    }
  }
  if (insert_diag && bc0)
    for (int64_t r = 0; r < nrows; ++r)
      if (bc0[r]) {
        int64_t p = find(row_ptr, cols, (int32_t)r, (int32_t)r);
        if (p >= 0) values[p] = diagonal;
      }
  return missing;
}

/* b += element vectors over every integral (no masking). */
int or_assemble_vector(const or_problem* P, int nint, const or_kernel* kernels,
                       const or_entities* ints, double* b)
{
  double be[OR_MAXT], xc[2 * 27 * 3], w[1024];
  int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
  int eli[2];
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int64_t e = 0; e < ints[it].n; ++e) {
      const int32_t* ent = &ints[it].entities[e * s];
      int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
      int nr = ncell * P->nd0 * P->bs0;
      memset(be, 0, sizeof(double) * nr);
      or_tabulate(&kernels[it], be, w, P->constants, xc, eli);
      for (int i = 0; i < nr; ++i) b[dofs0[i]] += be[i];
    }
  }
  return 0;
}

/* Value of a functional: the sum over every integral's entities, in entity
 * order (CUDAAssembler.h:296-351 sums cells, exterior, interior facets). */
double or_assemble_scalar(const or_problem* P, int nint, const or_kernel* kernels,
                          const or_entities* ints)
{
  double total = 0.0, xc[2 * 27 * 3], w[1024];
  int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
  int eli[2];
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int64_t e = 0; e < ints[it].n; ++e) {
      const int32_t* ent = &ints[it].entities[e * s];
      gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
      double a = 0.0;
      or_tabulate(&kernels[it], &a, w, P->constants, xc, eli);
      total += a;
    }
  }
  return total;
}

/* ---------------------------------------------------------------- multi-threaded
 * The same matrix and vector assembly on several cores, for the CPU
 * baseline (DOLFINx itself assembles serially per MPI rank).  Entities are
 * coloured so that no two entities of one colour share a row dof; each colour
 * is one parallel loop, so every value is written by one thread at a time.
 * Within a value, contributions arrive colour by colour instead of entity by
 * entity, so results match the serial functions to rounding. */

/* Greedy colouring of integral `it`'s entities by their row dofs (dofs0):
 * colours[e] in [0, 64).  Returns the number of colours, -1 if 64 do not
 * suffice, -2 if out of memory. */
int or_colour_entities(const or_problem* P, const or_entities* ent, int64_t nrows,
                       int8_t* colours)
{
  uint64_t* used = (uint64_t*)calloc((size_t)(nrows > 0 ? nrows : 1), sizeof(uint64_t));
  if (!used) return -2;
  double xc[2 * 27 * 3], w[1024];
  int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
  int eli[2], ncol = 0;
  const int s = stride_of(ent->type);
  for (int64_t e = 0; e < ent->n; ++e) {
    int ncell = gather(P, ent->type, &ent->entities[e * s], xc, w, dofs0, dofs1, eli);
    int nr = ncell * P->nd0 * P->bs0;
    uint64_t mask = 0;
    for (int i = 0; i < nr; ++i) mask |= used[dofs0[i]];
    if (mask == ~(uint64_t)0) { free(used); return -1; }
    int c = 0;
    while (mask & ((uint64_t)1 << c)) ++c;
    for (int i = 0; i < nr; ++i) used[dofs0[i]] |= (uint64_t)1 << c;
    colours[e] = (int8_t)c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(used);
  return ncol;
}

/* or_assemble_matrix with `nthreads` threads: order[it] lists integral it's
 * entities colour by colour, cstart[it][c] .. cstart[it][c + 1] those of
 * colour c (ncol[it] colours). */
int or_assemble_matrix_mt(const or_problem* P, int nint, const or_kernel* kernels,
                          const or_entities* ints, const int64_t* const* order,
                          const int64_t* const* cstart, const int* ncol, int nthreads,
                          const int8_t* bc0, const int8_t* bc1, const int32_t* row_ptr,
                          const int32_t* cols, double* values, int insert_diag, double diagonal,
                          int64_t nrows)
{
  int missing = 0;
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int c = 0; c < ncol[it]; ++c) {
#pragma omp parallel num_threads(nthreads) reduction(+ : missing)
      {
        double* A = (double*)malloc(sizeof(double) * OR_MAXT * OR_MAXT);
        double xc[2 * 27 * 3], w[1024];
        int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
        int eli[2];
#pragma omp for schedule(dynamic, 256)
        for (int64_t k = cstart[it][c]; k < cstart[it][c + 1]; ++k) {
          const int32_t* ent = &ints[it].entities[order[it][k] * s];
          int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
          int nr = ncell * P->nd0 * P->bs0, nc = ncell * P->nd1 * P->bs1;
          memset(A, 0, sizeof(double) * nr * nc);
          or_tabulate(&kernels[it], A, w, P->constants, xc, eli);
          for (int i = 0; i < nr; ++i) {
            int32_t row = dofs0[i];
            if (bc0 && bc0[row]) continue;
            for (int j = 0; j < nc; ++j) {
              int32_t col = dofs1[j];
              if (bc1 && bc1[col]) continue;
              int64_t p = find(row_ptr, cols, row, col);
              if (p < 0) { ++missing; continue; }
              values[p] += A[i * nc + j];
            }
          }
        }
        free(A);
      }
    }
  }
  if (insert_diag && bc0)
    for (int64_t r = 0; r < nrows; ++r)
      if (bc0[r]) {
        int64_t p = find(row_ptr, cols, (int32_t)r, (int32_t)r);
        if (p >= 0) values[p] = diagonal;
      }
  return missing;
}

/* or_assemble_vector with `nthreads` threads, colours as above. */
int or_assemble_vector_mt(const or_problem* P, int nint, const or_kernel* kernels,
                          const or_entities* ints, const int64_t* const* order,
                          const int64_t* const* cstart, const int* ncol, int nthreads, double* b)
{
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int c = 0; c < ncol[it]; ++c) {
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 256)
      for (int64_t k = cstart[it][c]; k < cstart[it][c + 1]; ++k) {
        double be[OR_MAXT], xc[2 * 27 * 3], w[1024];
        int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
        int eli[2];
        const int32_t* ent = &ints[it].entities[order[it][k] * s];
        int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
        int nr = ncell * P->nd0 * P->bs0;
        memset(be, 0, sizeof(double) * nr);
        or_tabulate(&kernels[it], be, w, P->constants, xc, eli);
        for (int i = 0; i < nr; ++i) b[dofs0[i]] += be[i];
      }
    }
  }
  return 0;
}

/* or_apply_lifting with `nthreads` threads, colours as above. */
int or_apply_lifting_mt(const or_problem* P, int nint, const or_kernel* kernels,
                        const or_entities* ints, const int64_t* const* order,
                        const int64_t* const* cstart, const int* ncol, int nthreads,
                        const int8_t* bc1, const double* g, const double* x0, double scale,
                        double* b)
{
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int c = 0; c < ncol[it]; ++c) {
#pragma omp parallel num_threads(nthreads)
      {
        double* A = (double*)malloc(sizeof(double) * OR_MAXT * OR_MAXT);
        double be[OR_MAXT], xc[2 * 27 * 3], w[1024];
        int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
        int eli[2];
#pragma omp for schedule(dynamic, 256)
        for (int64_t k = cstart[it][c]; k < cstart[it][c + 1]; ++k) {
          const int32_t* ent = &ints[it].entities[order[it][k] * s];
          int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
          int nr = ncell * P->nd0 * P->bs0, nc = ncell * P->nd1 * P->bs1;
          int has = 0;
          for (int j = 0; j < nc; ++j) has |= bc1[dofs1[j]] != 0;
          if (!has) continue;
          memset(A, 0, sizeof(double) * nr * nc);
          memset(be, 0, sizeof(double) * nr);
          or_tabulate(&kernels[it], A, w, P->constants, xc, eli);
          for (int j = 0; j < nc; ++j) {
            int32_t col = dofs1[j];
            if (!bc1[col]) continue;
            double v = scale * (g[col] - (x0 ? x0[col] : 0.0));
            for (int i = 0; i < nr; ++i) be[i] -= A[i * nc + j] * v;
          }
          for (int i = 0; i < nr; ++i) b[dofs0[i]] += be[i];
        }
        free(A);
      }
    }
  }
  return 0;
}

/* b -= scale * A[:, bc] (g - x0) over every integral of the bilinear form. */
int or_apply_lifting(const or_problem* P, int nint, const or_kernel* kernels,
                     const or_entities* ints, const int8_t* bc1, const double* g,
                     const double* x0, double scale, double* b)
{
  static double A[OR_MAXT * OR_MAXT];
  double be[OR_MAXT], xc[2 * 27 * 3], w[1024];
  int32_t dofs0[OR_MAXT], dofs1[OR_MAXT];
  int eli[2];
  for (int it = 0; it < nint; ++it) {
    const int s = stride_of(ints[it].type);
    for (int64_t e = 0; e < ints[it].n; ++e) {
      const int32_t* ent = &ints[it].entities[e * s];
      int ncell = gather(P, ints[it].type, ent, xc, w, dofs0, dofs1, eli);
      int nr = ncell * P->nd0 * P->bs0, nc = ncell * P->nd1 * P->bs1;
      int has = 0;
      for (int j = 0; j < nc; ++j) has |= bc1[dofs1[j]] != 0;
      if (!has) continue;
      memset(A, 0, sizeof(double) * nr * nc);
      memset(be, 0, sizeof(double) * nr);
      or_tabulate(&kernels[it], A, w, P->constants, xc, eli);
      for (int j = 0; j < nc; ++j) {
        int32_t col = dofs1[j];
        if (!bc1[col]) continue;
        double v = scale * (g[col] - (x0 ? x0[col] : 0.0));
        for (int i = 0; i < nr; ++i) be[i] -= A[i * nc + j] * v;
      }
      for (int i = 0; i < nr; ++i) b[dofs0[i]] += be[i];
    }
  }
  return 0;
}

/* Same-map interpolation, cell by cell in order (CUDAInterpolate.cu:48-74:
 * gather u0's cell values, multiply by the n1 x n0 operator im (k ascending),
 * scatter; a node shared by several cells keeps the last cell's value). */
void or_interpolate_same_map(int64_t ncells, const int32_t* dm1, int n1, const int32_t* dm0,
                             int n0, int bs, const double* im, const double* u0, double* u1)
{
  for (int64_t c = 0; c < ncells; ++c)
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < bs; ++j) {
        double acc = 0.0;
        for (int k = 0; k < n0; ++k) acc += im[i * n0 + k] * u0[(int64_t)bs * dm0[c * n0 + k] + j];
        u1[(int64_t)bs * dm1[c * n1 + i] + j] = acc;
      }
}

/* b[d] = scale * (g[d] - x0[d]) for the constrained dofs (CUDAAssembler.cpp:99-126). */
void or_set_bc(int64_t nbc, const int32_t* dofs, const double* g, const double* x0, double scale,
               double* b)
{
  for (int64_t i = 0; i < nbc; ++i) {
    int32_t d = dofs[i];
    b[d] = scale * (g[d] - (x0 ? x0[d] : 0.0));
  }
}
