"""ctypes front end of the CPU oracle (liboracle.so built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests, by the smoke check and by the
CPU-baseline leg of the benchmark, never by the engine itself.

`OracleAssembler` assembles a `mifem.forms.Form` the way DOLFINx does on the
CPU (reference semantics cited in oracle.c), taking only the mesh, dofmaps,
entity lists, packed coefficients, constants and each integral's integrand
description (`Integral.spec`).  It never looks at the generated C bodies or
the HIP templates, so a mistake in the engine's code generation shows up as
a parity failure.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liboracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(_LIB_PATH))
        _declare(_lib)
    return _lib


class Kernel(C.Structure):
    _fields_ = [("kind", C.c_int), ("cell", C.c_int), ("degree", C.c_int), ("bs", C.c_int),
                ("rank", C.c_int), ("qm", C.c_int),
                ("kappa_src", C.c_int), ("kappa_off", C.c_int), ("kappa_val", C.c_double),
                ("mu_src", C.c_int), ("mu_off", C.c_int), ("mu_val", C.c_double),
                ("lam_src", C.c_int), ("lam_off", C.c_int), ("lam_val", C.c_double),
                ("f_src", C.c_int), ("f_off", C.c_int), ("f_degree", C.c_int),
                ("k2", C.c_double),
                ("p", C.c_double), ("eps", C.c_double), ("alpha", C.c_double),
                ("penalty", C.c_double), ("fval", C.c_double),
                ("u_off", C.c_int), ("coeff_width", C.c_int), ("mu0", C.c_double),
                ("diag_part", C.c_int), ("kappa_degree", C.c_int)]


class Entities(C.Structure):
    _fields_ = [("type", C.c_int), ("entities", C.POINTER(C.c_int32)), ("n", C.c_int64)]


class Problem(C.Structure):
    _fields_ = [("x", C.POINTER(C.c_double)), ("cells", C.POINTER(C.c_int32)), ("nv", C.c_int),
                ("dofmap0", C.POINTER(C.c_int32)), ("nd0", C.c_int), ("bs0", C.c_int),
                ("dofmap1", C.POINTER(C.c_int32)), ("nd1", C.c_int), ("bs1", C.c_int),
                ("coeffs", C.POINTER(C.c_double)), ("W", C.c_int),
                ("coeff_offsets", C.POINTER(C.c_int)), ("ncoef", C.c_int),
                ("constants", C.POINTER(C.c_double))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i8 = C.POINTER(C.c_int8)


def _declare(L):
    L.or_gauss_jacobi.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.or_tet_rule.argtypes = [C.c_int, _dp, _dp]
    L.or_tri_rule.argtypes = [C.c_int, _dp, _dp]
    L.or_hex_rule.argtypes = [C.c_int, _dp, _dp]
    L.or_basis.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp]
    L.or_ndofs.argtypes = [C.c_int, C.c_int]
    L.or_tabulate.argtypes = [C.POINTER(Kernel), _dp, _dp, _dp, _dp, C.POINTER(C.c_int)]
    L.or_sparsity.argtypes = [C.c_int, C.POINTER(Entities), _ip, C.c_int, C.c_int, _ip, C.c_int,
                              C.c_int, C.c_int64, C.POINTER(_ip), C.POINTER(_ip),
                              C.POINTER(C.c_int64), C.POINTER(_ip), C.POINTER(_ip)]
    L.or_free.argtypes = [C.c_void_p]
    _i8 = C.POINTER(C.c_int8)
    _lp = C.POINTER(C.c_int64)
    L.or_colour_entities.argtypes = [C.POINTER(Problem), C.POINTER(Entities), C.c_int64, _i8]
    L.or_assemble_matrix_mt.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                        C.POINTER(Entities), C.POINTER(_lp), C.POINTER(_lp),
                                        C.POINTER(C.c_int), C.c_int, _i8, _i8, _ip, _ip, _dp,
                                        C.c_int, C.c_double, C.c_int64]
    L.or_assemble_vector_mt.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                        C.POINTER(Entities), C.POINTER(_lp), C.POINTER(_lp),
                                        C.POINTER(C.c_int), C.c_int, _dp]
    L.or_assemble_matrix.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                     C.POINTER(Entities), _i8, _i8, _ip, _ip, _dp, C.c_int,
                                     C.c_double, C.c_int64]
    L.or_assemble_matrix.restype = C.c_int
    L.or_assemble_vector.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                     C.POINTER(Entities), _dp]
    L.or_apply_lifting.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                   C.POINTER(Entities), _i8, _dp, _dp, C.c_double, _dp]
    L.or_apply_lifting_mt.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                      C.POINTER(Entities), C.POINTER(_lp), C.POINTER(_lp),
                                      C.POINTER(C.c_int), C.c_int, _i8, _dp, _dp, C.c_double,
                                      _dp]
    L.or_assemble_scalar.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(Kernel),
                                     C.POINTER(Entities)]
    L.or_assemble_scalar.restype = C.c_double
    L.or_set_bc.argtypes = [C.c_int64, _ip, _dp, _dp, C.c_double, _dp]
    L.or_interpolate_same_map.argtypes = [C.c_int64, _ip, C.c_int, _ip, C.c_int, C.c_int, _dp,
                                          _dp, _dp]
    L.or_pack_coefficients.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(_ip),
                                       C.POINTER(_dp), C.POINTER(C.c_int), C.c_int64, C.c_int, _dp]


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


# ------------------------------------------------------------- small pieces
def gauss_jacobi(m: int, alpha: float):
    t, w = np.zeros(m), np.zeros(m)
    lib().or_gauss_jacobi(m, alpha, _p(t, C.c_double), _p(w, C.c_double))
    return t, w


def tet_rule(m: int):
    pts, wts = np.zeros((m ** 3, 3)), np.zeros(m ** 3)
    lib().or_tet_rule(m, _p(pts, C.c_double), _p(wts, C.c_double))
    return pts, wts


def tri_rule(m: int):
    pts, wts = np.zeros((m ** 2, 2)), np.zeros(m ** 2)
    lib().or_tri_rule(m, _p(pts, C.c_double), _p(wts, C.c_double))
    return pts, wts


def hex_rule(m: int):
    pts, wts = np.zeros((m ** 3, 3)), np.zeros(m ** 3)
    lib().or_hex_rule(m, _p(pts, C.c_double), _p(wts, C.c_double))
    return pts, wts


def basis(cell: int, k: int, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    nd = lib().or_ndofs(cell, k)
    phi, dphi = np.zeros(nd), np.zeros((nd, 3))
    lib().or_basis(cell, k, _p(X, C.c_double), _p(phi, C.c_double), _p(dphi, C.c_double))
    return phi, dphi


# ------------------------------------------------------------- kernels
_KINDS = {"stiffness": 0, "elasticity": 1, "source": 2, "facet_source": 3,
          "dg_cell": 4, "dg_ext": 5, "dg_int": 6, "fn_cell": 7, "fn_ext": 8, "fn_int": 9,
          "mass": 10, "nl_cell": 11, "nl_ext": 12, "nl_int": 13}
_TYPES = {"cell": 0, "exterior_facet": 1, "interior_facet": 2}


def _src(v):
    kind, x = v
    if kind == "w":
        return 1, int(x), 0.0
    if kind == "c":
        return 2, int(x), 0.0
    return 0, 0, float(x)


def kernel_from_spec(spec: dict, coeff_width: int = 0) -> Kernel:
    k = Kernel()
    k.kind = _KINDS[spec["kind"]]
    k.cell = spec["cell"]
    k.degree = spec["degree"]
    k.bs = spec["bs"]
    k.rank = spec["rank"]
    if "qm" in spec:
        k.qm = spec["qm"]
    else:  # exact for the integrand's polynomial degree
        k.qm = max(1, (spec["qdeg"] + 2) // 2)
    for name in ("kappa", "mu", "lam"):
        if name in spec:
            s, o, v = _src(spec[name])
            setattr(k, f"{name}_src", s)
            setattr(k, f"{name}_off", o)
            setattr(k, f"{name}_val", v)
    if "f" in spec:
        s, o, _ = _src(spec["f"])
        k.f_src, k.f_off = s, o
        k.f_degree = spec.get("f_degree", 0)
    k.k2 = spec.get("k2", 0.0)
    for name in ("p", "eps", "alpha", "penalty", "fval", "mu0"):
        setattr(k, name, float(spec.get(name, 0.0)))
    k.u_off = int(spec.get("u_off", 0))
    if "kappa_off" in spec:
        k.kappa_off = int(spec["kappa_off"])
    if "v_off" in spec:
        k.f_off, k.f_degree = int(spec["v_off"]), int(spec["v_degree"])
    if "trial_degree" in spec:
        k.f_degree = int(spec["trial_degree"])
    k.diag_part = int(spec.get("diag_part", 0))
    k.kappa_degree = int(spec.get("kappa_degree", spec["degree"]))
    k.coeff_width = coeff_width
    return k


def tabulate(spec: dict, w, c, coordinate_dofs, entity_local_index=None):
    """Element tensor of one entity, as a flat array (rank 2: nr x nc)."""
    k = kernel_from_spec(spec, len(w))
    nd = lib().or_ndofs(spec["cell"], spec["degree"]) * spec["bs"]
    mult = 2 if spec["kind"] in ("dg_int", "nl_int") else 1
    n = nd * mult
    m = (lib().or_ndofs(spec["cell"], spec["trial_degree"]) * spec["bs"] * mult
         if "trial_degree" in spec else n)
    A = np.zeros(n * m if spec["rank"] == 2 else (n if spec["rank"] == 1 else 1))
    w = np.ascontiguousarray(w, dtype=np.float64)
    c = np.ascontiguousarray(c if len(c) else np.zeros(1), dtype=np.float64)
    xc = np.ascontiguousarray(coordinate_dofs, dtype=np.float64)
    eli = (C.c_int * 2)(*(list(entity_local_index or [0, 0]) + [0, 0])[:2])
    lib().or_tabulate(C.byref(k), _p(A, C.c_double), _p(w if w.size else np.zeros(1), C.c_double),
                      _p(c, C.c_double), _p(xc, C.c_double), eli)
    return A.reshape(n, m) if spec["rank"] == 2 else (A if spec["rank"] == 1 else A[0])


# ------------------------------------------------------------- assembly
class OracleAssembler:
    """DOLFINx-semantics FP64 CPU assembly of a mifem Form."""

    def __init__(self, form):
        self.form = form
        V0 = form.function_spaces[0] if form.rank >= 1 else None
        V1 = form.function_spaces[-1] if form.rank >= 1 else None
        mesh = form.mesh
        self._keep = []
        self.x = np.ascontiguousarray(mesh.x)
        self.cells = np.ascontiguousarray(mesh.cells, dtype=np.int32)
        self.dm0 = np.ascontiguousarray(V0.dofmap if V0 else np.zeros(1), dtype=np.int32)
        self.dm1 = np.ascontiguousarray(V1.dofmap if V1 else np.zeros(1), dtype=np.int32)
        self.V0, self.V1 = V0, V1
        self.coeffs = np.ascontiguousarray(self.pack_coefficients())
        self.coff = np.ascontiguousarray(np.array(form.coefficient_offsets, dtype=np.int32))
        self.consts = np.ascontiguousarray(np.concatenate([form.constant_values(), np.zeros(1)]))
        P = Problem()
        P.x, P.cells, P.nv = _p(self.x, C.c_double), _p(self.cells, C.c_int32), mesh.nv_per_cell
        if V0 is not None:
            P.dofmap0, P.nd0, P.bs0 = _p(self.dm0, C.c_int32), V0.element.ndofs, V0.bs
        if form.rank == 2:
            P.dofmap1, P.nd1, P.bs1 = _p(self.dm1, C.c_int32), V1.element.ndofs, V1.bs
        P.coeffs = _p(self.coeffs, C.c_double) if self.coeffs.size else None
        P.W = self.coeffs.shape[1] if self.coeffs.ndim == 2 else 0
        P.coeff_offsets = self.coff.ctypes.data_as(C.POINTER(C.c_int))
        P.ncoef = len(form.coefficients)
        P.constants = _p(self.consts, C.c_double)
        self.P = P
        n = len(form.integrals)
        self.kernels = (Kernel * n)(*[kernel_from_spec(I.spec, P.W) for I in form.integrals])
        ents = []
        for I in form.integrals:
            e = np.ascontiguousarray(I.entities, dtype=np.int32)
            self._keep.append(e)
            ents.append(Entities(_TYPES[I.type], _p(e, C.c_int32), I.num_entities))
        self.ents = (Entities * n)(*ents)

    def pack_coefficients(self) -> np.ndarray:
        """Packed [Nc x W] coefficients via the oracle's own pack loop."""
        f = self.form
        nc = f.mesh.num_cells
        off = f.coefficient_offsets
        W = off[-1]
        out = np.zeros((nc, max(W, 0)))
        if not f.coefficients:
            return out
        n = len(f.coefficients)
        dofs = [np.ascontiguousarray(c.function_space.unrolled_dofmap()) for c in f.coefficients]
        vals = [np.ascontiguousarray(c.x) for c in f.coefficients]
        npc = (C.c_int * n)(*[d.shape[1] for d in dofs])
        dps = (_ip * n)(*[_p(d, C.c_int32) for d in dofs])
        vps = (_dp * n)(*[_p(v, C.c_double) for v in vals])
        offs = (C.c_int * n)(*off[:-1])
        lib().or_pack_coefficients(n, npc, dps, vps, offs, nc, W, _p(out, C.c_double))
        return out

    # sparsity ---------------------------------------------------------
    def sparsity(self):
        """(row_ptr, cols, rpb, colsb): bs-unrolled CSR and its blocked pattern."""
        V0, V1 = self.V0, self.V1
        rp, cols, rpb, colsb = _ip(), _ip(), _ip(), _ip()
        nnz = C.c_int64()
        lib().or_sparsity(len(self.form.integrals), self.ents, self.P.dofmap0, V0.element.ndofs,
                          V0.bs, self.P.dofmap1, V1.element.ndofs, V1.bs, V0.num_nodes,
                          C.byref(rp), C.byref(cols), C.byref(nnz), C.byref(rpb), C.byref(colsb))
        nrows = V0.num_nodes * V0.bs
        row_ptr = np.ctypeslib.as_array(rp, shape=(nrows + 1,)).copy()
        c = np.ctypeslib.as_array(cols, shape=(max(nnz.value, 1),))[:nnz.value].copy()
        row_ptr_b = np.ctypeslib.as_array(rpb, shape=(V0.num_nodes + 1,)).copy()
        nb = int(row_ptr_b[-1])
        cb = np.ctypeslib.as_array(colsb, shape=(max(nb, 1),))[:nb].copy()
        for ptr in (rp, cols, rpb, colsb):
            lib().or_free(ptr)
        return row_ptr, c, row_ptr_b, cb

    # multi-core assembly (CPU baseline) -------------------------------
    def colouring(self):
        """Per integral: (entities ordered colour by colour, colour starts),
        no two entities of a colour sharing a row dof.  Computed once."""
        if getattr(self, "_colouring", None) is None:
            out = []
            nrows = self.V0.num_nodes * self.V0.bs
            for k, I in enumerate(self.form.integrals):
                col = np.zeros(max(I.num_entities, 1), dtype=np.int8)
                nc = lib().or_colour_entities(C.byref(self.P), C.byref(self.ents[k]), nrows,
                                              _p(col, C.c_int8))
                if nc < 0:
                    raise RuntimeError("entity colouring needs more than 64 colours")
                col = col[:I.num_entities]
                order = np.argsort(col, kind="stable").astype(np.int64)
                starts = np.zeros(nc + 1, dtype=np.int64)
                starts[1:] = np.cumsum(np.bincount(col, minlength=nc))
                out.append((order, starts, nc))
            self._colouring = out
        return self._colouring

    def _colour_args(self):
        cl = self.colouring()
        n = len(cl)
        lp = C.POINTER(C.c_int64)
        orders = (lp * n)(*[_p(o, C.c_int64) for o, _, _ in cl])
        starts = (lp * n)(*[_p(st, C.c_int64) for _, st, _ in cl])
        ncol = (C.c_int * n)(*[nc for _, _, nc in cl])
        return orders, starts, ncol

    def assemble_matrix_mt(self, row_ptr, cols, bcs=(), diagonal=1.0, threads=None):
        """assemble_matrix on `threads` cores (default: all), entities in colours."""
        threads = threads or os.cpu_count() or 1
        n0 = self.V0.size
        bc0 = _markers(bcs, self.V0)
        bc1 = _markers(bcs, self.V1)
        vals = np.zeros(cols.size)
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        orders, starts, ncol = self._colour_args()
        missing = lib().or_assemble_matrix_mt(
            C.byref(self.P), len(self.form.integrals), self.kernels, self.ents, orders, starts,
            ncol, int(threads), _p(bc0, C.c_int8), _p(bc1, C.c_int8), _p(row_ptr, C.c_int32),
            _p(cols, C.c_int32), _p(vals, C.c_double),
            int(self.V0 is self.V1 and bc0 is not None), diagonal, n0)
        if missing:
            raise RuntimeError(f"{missing} element entries missing from the sparsity pattern")
        return vals

    def assemble_vector_mt(self, threads=None):
        threads = threads or os.cpu_count() or 1
        b = np.zeros(self.V0.size)
        orders, starts, ncol = self._colour_args()
        lib().or_assemble_vector_mt(C.byref(self.P), len(self.form.integrals), self.kernels,
                                    self.ents, orders, starts, ncol, int(threads),
                                    _p(b, C.c_double))
        return b

    # assembly ---------------------------------------------------------
    def assemble_matrix(self, row_ptr, cols, bcs=(), diagonal=1.0, values=None):
        n0 = self.V0.size
        bc0 = _markers(bcs, self.V0)
        bc1 = _markers(bcs, self.V1)
        vals = np.zeros(cols.size) if values is None else values
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        missing = lib().or_assemble_matrix(
            C.byref(self.P), len(self.form.integrals), self.kernels, self.ents, _p(bc0, C.c_int8),
            _p(bc1, C.c_int8), _p(row_ptr, C.c_int32), _p(cols, C.c_int32),
            _p(vals, C.c_double), int(self.V0 is self.V1 and bc0 is not None), diagonal, n0)
        if missing:
            raise RuntimeError(f"{missing} element entries missing from the sparsity pattern")
        return vals

    def assemble_vector(self, b=None):
        b = np.zeros(self.V0.size) if b is None else b
        lib().or_assemble_vector(C.byref(self.P), len(self.form.integrals), self.kernels,
                                 self.ents, _p(b, C.c_double))
        return b

    def assemble_scalar(self) -> float:
        """Value of a functional: every integral's entity values summed in order."""
        if self.form.rank != 0:
            raise ValueError("assemble_scalar needs a rank-0 form")
        return float(lib().or_assemble_scalar(C.byref(self.P), len(self.form.integrals),
                                              self.kernels, self.ents))

    def apply_lifting(self, b, bcs, x0=None, scale=1.0, threads=1):
        """b -= scale A[:, bc](g - x0); threads > 1 runs the coloured multi-core loop."""
        bc1 = _markers(bcs, self.V1)
        if bc1 is None:
            return b
        g = _dense_values(bcs, self.V1)
        x0 = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        x0p = _p(x0, C.c_double) if x0 is not None else None
        if threads > 1:
            orders, starts, ncol = self._colour_args()
            lib().or_apply_lifting_mt(C.byref(self.P), len(self.form.integrals), self.kernels,
                                      self.ents, orders, starts, ncol, int(threads),
                                      _p(bc1, C.c_int8), _p(g, C.c_double), x0p, scale,
                                      _p(b, C.c_double))
            return b
        lib().or_apply_lifting(C.byref(self.P), len(self.form.integrals), self.kernels, self.ents,
                               _p(bc1, C.c_int8), _p(g, C.c_double), x0p, scale,
                               _p(b, C.c_double))
        return b


def interpolate_same_map(V1, V0, im, u0, u1=None):
    """u1 on V1 from u0 on V0 (same mesh) with the (n1, n0) element operator im."""
    dm1 = np.ascontiguousarray(V1.dofmap, dtype=np.int32)
    dm0 = np.ascontiguousarray(V0.dofmap, dtype=np.int32)
    im = np.ascontiguousarray(im, dtype=np.float64)
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    u1 = np.zeros(V1.size) if u1 is None else u1
    lib().or_interpolate_same_map(dm1.shape[0], _p(dm1, C.c_int32), dm1.shape[1],
                                  _p(dm0, C.c_int32), dm0.shape[1], V1.bs, _p(im, C.c_double),
                                  _p(u0, C.c_double), _p(u1, C.c_double))
    return u1


def set_bc(b, bcs, V, x0=None, scale=1.0):
    for bc in bcs:
        if bc.function_space is not V:
            continue
        d = np.ascontiguousarray(bc.dofs, dtype=np.int32)
        g = np.ascontiguousarray(bc.values)
        x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
        lib().or_set_bc(d.size, _p(d, C.c_int32), _p(g, C.c_double),
                        _p(x0a, C.c_double) if x0a is not None else None, scale, _p(b, C.c_double))
    return b


def _markers(bcs, V):
    sel = [bc for bc in bcs if bc.function_space is V]
    if not sel:
        return None
    m = np.zeros(V.size, dtype=np.int8)
    for bc in sel:
        m[bc.dofs] = 1
    return m


def _dense_values(bcs, V):
    g = np.zeros(V.size)
    for bc in bcs:
        if bc.function_space is V:
            g[bc.dofs] = bc.values[bc.dofs]
    return g
