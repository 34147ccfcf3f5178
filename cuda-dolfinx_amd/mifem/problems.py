"""The measured assembly problems, built the way python/examples/poisson.py
builds its forms: mesh, space, bilinear and linear forms, Dirichlet BCs.

* `poisson_p1(n)`: -Laplace u = f, P1 on the n^3 x 6 unit-cube tet mesh,
  f = sin(pi x) sin(pi y) sin(pi z) interpolated into P1, u = 0 on the whole
  boundary (n = 32 is the 196,608-cell oracle configuration).
* `elasticity_p2(n)`: vector P2 linear elasticity on n^3 x 6 tets,
  Lame mu, lambda piecewise constant ~ U[1, 10] per cell from a seed, body
  force (0, 0, -1), clamped x = 0 face, traction (0, 0, -1) on x = 1
  (n = 64 is the 1,572,864-cell configuration).
* `helmholtz_q3(n)`: grad u . grad v - k^2 u v, k = 5, Q3 on n^3 hexahedra
  with interior vertices jittered by +-0.2 h, f = 1, u = 0 on the boundary,
  4 x 4 x 4 Gauss points.
* `dg_plaplace(n)`: SIPG DG1 p-Laplacian (coefficient 1 + (eps^2 + |grad u|^2)^((p-2)/2),
  p = 3, eps = 1e-3) with Robin
  alpha = 2 and f = 1 on n^3 x 6 tets with interior vertices jittered by
  +-0.2 h; `a` is the Newton Jacobian J(u), `L` the residual F(u) at the
  state u = sin(pi x) sin(pi y) sin(pi z) interpolated into DG1.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import forms
from .fem import Constant, DirichletBC, Function, dirichletbc, functionspace, locate_dofs_topological
from .mesh import Mesh, cell_random_uniform, create_unit_cube


@dataclass(eq=False)
class Problem:
    name: str
    mesh: Mesh
    V: object
    a: forms.Form
    L: forms.Form
    bcs: List[DirichletBC]
    extras: dict = field(default_factory=dict)

    @property
    def ndofs(self) -> int:
        return self.V.size


def _boundary_bc(mesh, V, value=0.0):
    facets = mesh.exterior_facets
    return dirichletbc(value, locate_dofs_topological(V, 2, facets), V)


def poisson_p1(n: int = 32) -> Problem:
    mesh = create_unit_cube(n, "tetrahedron")
    V = functionspace(mesh, ("P", 1))
    f = Function(V, "f")
    f.interpolate(lambda x: np.sin(np.pi * x[0]) * np.sin(np.pi * x[1]) * np.sin(np.pi * x[2]))
    a = forms.stiffness(V)
    L = forms.source(V, f)
    return Problem("poisson_p1", mesh, V, a, L, [_boundary_bc(mesh, V)], {"f": f})


def elasticity_p2(n: int = 64, seed: int = 0) -> Problem:
    mesh = create_unit_cube(n, "tetrahedron")
    V = functionspace(mesh, ("P", 2), shape=(3,))
    DG0 = functionspace(mesh, ("DG", 0))
    mu, lam = Function(DG0, "mu"), Function(DG0, "lambda")
    mu.set_cellwise(cell_random_uniform(mesh, seed, 0, 1.0, 10.0))
    lam.set_cellwise(cell_random_uniform(mesh, seed, 1, 1.0, 10.0))
    a = forms.elasticity(V, mu, lam)
    body = forms.source(V, Constant([0.0, 0.0, -1.0]))
    right = mesh.locate_boundary_facets(lambda x: np.isclose(x[0], 1.0))
    traction = forms.facet_source(V, Constant([0.0, 0.0, -1.0]), right)
    L = forms.linear_sum(body, traction)
    left = mesh.locate_boundary_facets(lambda x: np.isclose(x[0], 0.0))
    bc = dirichletbc(np.zeros(3), locate_dofs_topological(V, 2, left), V)
    return Problem("elasticity_p2", mesh, V, a, L, [bc], {"mu": mu, "lambda": lam})


def helmholtz_q3(n: int = 48, seed: int = 1, k: float = 5.0, degree: int = 3) -> Problem:
    mesh = create_unit_cube(n, "hexahedron", jitter=0.2, seed=seed)
    V = functionspace(mesh, ("P", degree))
    a = forms.helmholtz(V, k, qpoints_per_dir=degree + 1)
    L = forms.hex_source(V, Constant(1.0), qpoints_per_dir=degree + 1)
    return Problem("helmholtz_q3", mesh, V, a, L, [_boundary_bc(mesh, V)], {"k": k})


def dg_plaplace(n: int = 48, seed: int = 2, degree: int = 1, p: float = 3.0,
                shuffle_vertices=None) -> Problem:
    """SIPG p-Laplacian on jittered tetrahedra; shuffle_vertices (a seed) lists
    each cell's vertices in random order (non-zero facet permutations)."""
    from .dgforms import sipg_plaplace
    mesh = create_unit_cube(n, "tetrahedron", jitter=0.2, seed=seed,
                            shuffle_vertices=shuffle_vertices)
    V = functionspace(mesh, ("DG", degree))
    u = Function(V, "u")
    u.interpolate(lambda x: np.sin(np.pi * x[0]) * np.sin(np.pi * x[1]) * np.sin(np.pi * x[2]))
    J, F = sipg_plaplace(V, u, f=1.0, alpha=2.0, p=p, eps=1e-3)
    return Problem("dg_plaplace", mesh, V, J, F, [], {"u": u, "p": p})
