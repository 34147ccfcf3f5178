"""Benchmark: FP64 assembly of A + b for vector-P2 linear elasticity on the
64^3 x 6 unit-cube tetrahedral mesh (1,572,864 cells, 6,440,067 dofs) on one
MI355X, through the cuDOLFINx-style API.

One step = the timed regions of python/examples/poisson.py:91-114:
  A region: assemble_matrix (zero + cell integral + Dirichlet rows/cols +
            diagonal) + handoff (A.assemble(), a device synchronisation);
  b region: assemble_vector (cells + traction facets) + apply_lifting + set_bc.
Setup (mesh, dofmaps, sparsity, owner-computes schedule, hipRTC compile) is
outside the timed region.  Inputs live in HBM before timing starts.

Printed: one JSON line with DoF/s (value), the roofline of the dominant
kernel (owner-computes matrix kernel, timed live with HIP events on the
stream it runs on, beside the copy bandwidth measured on the same GPU), a
CPU baseline (the oracle on the full workload on every core the process may
use, and on one core on the 32^3 sample), and the metric's mixed-precision
half as the `c3` sub-record: Q3 Helmholtz on 48^3 jittered hexes, FP64 GPU
path against the FP32 / BF16 matrix-core paths, errors of A, b and a
fixed-tolerance CG solution against the FP64 oracle.  --workload c3 / c4 /
c5 run the Q3 precision comparison, the fused SIPG Newton step and the
512-instance batch as their own lines.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "cuda-dolfinx_amd"), str(ROOT)]

METRIC = "assembly DoF/s + % HBM roofline; mixed-prec vs FP64 speedup at rel err≤1e-6"
HBM_PEAK_GBS = 8000.0


def measured_peaks(bound: str) -> dict:
    """What this GPU reaches on plain library work, beside the published peaks:
    device copy and fill of 2 GiB (HBM), or 8192^3 FP32 / FP64 GEMMs (matrix
    cores).  Timed with HIP events after the benchmark, outside its timed region."""
    import torch

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    if bound == "hbm":
        n = (2 << 30) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        t_copy = timed(lambda: dst.copy_(src))
        t_fill = timed(lambda: dst.fill_(2.0))
        del src, dst
        return {"copy_gbs": 2 * 8 * n / t_copy / 1e9, "fill_gbs": 8 * n / t_fill / 1e9,
                "what": "torch copy_ (read + write) and fill_ (write) of 2 GiB float64"}
    n = 8192
    out = {"what": "torch.matmul of 8192x8192 matrices (rocBLAS / hipBLASLt)"}
    for name, dt in (("sgemm_tflops", torch.float32), ("dgemm_tflops", torch.float64)):
        a = torch.rand(n, n, dtype=dt, device="cuda")
        b = torch.rand(n, n, dtype=dt, device="cuda")
        out[name] = 2 * n ** 3 / timed(lambda: torch.matmul(a, b), reps=3) / 1e12
        del a, b
    return out


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(P, A):
    """Algorithmic HBM bytes of one matrix launch (owner-computes kernel).

    values written once (8 nnz) + blocked CSR structure read once
    (4 nnz_b + 4 (N_b + 1)) + geometry (24 Nv + 16 Nc) + packed Lame
    coefficients (16 Nc) + blocked dofmap (4 nd Nc).
    """
    nc, nv = P.mesh.num_cells, P.mesh.num_vertices
    nb = P.V.num_nodes
    nnz = A.nnz
    nnz_b = nnz // (P.V.bs * P.V.bs)
    nd = P.V.element.ndofs
    return 8 * nnz + 4 * nnz_b + 4 * (nb + 1) + 24 * nv + 16 * nc + 16 * nc + 4 * nd * nc


def measured_traffic(name: str = "c2_elasticity_p2_64_matrix_traffic.md"):
    """HBM bytes per launch from a committed rocprofv3 FETCH/WRITE summary."""
    p = ROOT / "profiles" / name
    try:
        for line in p.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def _cgroup_cpus():
    """CPUs the cgroup quota allows this process (None: no quota)."""
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            return max(1, -(-int(quota) // int(period)))
    except (OSError, ValueError):
        pass
    return None


def _cpu_threads() -> int:
    """Every core this process may run on: its CPU affinity, capped by the
    cgroup CPU quota when one is set (more threads than the quota would only
    time-slice)."""
    n = len(os.sched_getaffinity(0))
    q = _cgroup_cpus()
    return max(1, min(n, q) if q else n)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _cpu_fields() -> dict:
    """CPU model and core counts recorded with every CPU baseline."""
    return {"cpu": _cpu_model(), "affinity_cores": len(os.sched_getaffinity(0)),
            "cgroup_cpus": _cgroup_cpus(), "host_cpus": os.cpu_count()}


def cpu_baseline_c3(n_sample: int = 16):
    """Oracle (FP64) Q3 Helmholtz matrix on n_sample^3 jittered hexes, on all
    usable cores (coloured cells) and on one core."""
    from mifem import problems
    from oracle import cpu
    P = problems.helmholtz_q3(n_sample)
    Oa = cpu.OracleAssembler(P.a)
    rp, cols, _, _ = Oa.sparsity()
    t0 = time.perf_counter()
    Oa.assemble_matrix(rp, cols, P.bcs)
    dt1 = time.perf_counter() - t0
    threads = _cpu_threads()
    Oa.colouring()
    t0 = time.perf_counter()
    Oa.assemble_matrix_mt(rp, cols, P.bcs, threads=threads)
    dt = time.perf_counter() - t0
    nc = P.mesh.num_cells
    return {"value": nc / dt, "unit": "cells/s (A)", "cores": threads, "kind": "port",
            "est_full_ms": 110592 / (nc / dt) * 1e3, "serial_value": nc / dt1,
            "sample": f"helmholtz_q3 {n_sample}^3 hexes ({nc} cells): A in {dt:.2f} s on {threads} "
                      f"threads (oracle, C + OpenMP, coloured cells), {dt1:.2f} s serial",
            **_cpu_fields()}


def cpu_baseline_c4(n_sample: int = 24):
    """Oracle Jacobian + residual + energy of the SIPG DG1 step: J and F on all
    usable cores (coloured cells / facets), the energy serial; and all on one core."""
    from mifem import problems
    from mifem.dgforms import sipg_energy
    from oracle import cpu
    P = problems.dg_plaplace(n_sample)
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    OE = cpu.OracleAssembler(sipg_energy(P.V, P.extras["u"]))
    rp, cols, _, _ = Oa.sparsity()
    t0 = time.perf_counter()
    Oa.assemble_matrix(rp, cols)
    OL.assemble_vector()
    OE.assemble_scalar()
    dt1 = time.perf_counter() - t0
    threads = _cpu_threads()
    Oa.colouring()
    OL.colouring()
    t0 = time.perf_counter()
    Oa.assemble_matrix_mt(rp, cols, threads=threads)
    OL.assemble_vector_mt(threads=threads)
    OE.assemble_scalar()
    dt = time.perf_counter() - t0
    nc = P.mesh.num_cells
    return {"value": dt * 663552 / nc * 1e3, "unit": "ms per Newton step at 48^3 (scaled by cells)",
            "cores": threads, "kind": "port", "serial_value": dt1 * 663552 / nc * 1e3,
            "sample": f"sipg_plaplace DG1 {n_sample}^3x6 tets ({nc} cells): J + F + E in {dt:.2f} s "
                      f"on {threads} threads (oracle, C + OpenMP, coloured entities; E serial), "
                      f"{dt1:.2f} s serial",
            **_cpu_fields()}


def cpu_baseline_c5(B, per_degree: int = 2):
    """Oracle (single core) A + b for the first instances of each degree of the batch."""
    from mifem import batch as mb
    from mifem import forms as mf
    from mifem.fem import Constant, Function, dirichletbc, functionspace, locate_dofs_topological
    from oracle import cpu
    picked = []
    for g in B.groups:
        picked += [(i, g.degree) for i in g.instances[:per_degree]]
    total = 0.0
    for i, k in picked:
        mesh, V, kap = mb.poisson_instance(B.sizes[i], k)
        D = functionspace(mesh, ("DG", 0))
        kf = Function(D)
        kf.set_cellwise(kap)
        a, L = mf.stiffness(V, kf), mf.source(V, Constant([1.0]))
        bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
        Oa, OL = cpu.OracleAssembler(a), cpu.OracleAssembler(L)
        rp, cols, _, _ = Oa.sparsity()
        t0 = time.perf_counter()
        Oa.assemble_matrix(rp, cols, [bc])
        b = OL.assemble_vector()
        Oa.apply_lifting(b, [bc])
        cpu.set_bc(b, [bc], V)
        total += time.perf_counter() - t0
    return {"value": len(picked) / total, "unit": "assemblies/s", "cores": 1, "kind": "port",
            "sample": f"instances {[i for i, _ in picked]} (first {per_degree} of each degree): "
                      f"A + b in {total:.2f} s, oracle (C, serial)",
            **_cpu_fields()}


def cpu_baseline(n_full: int, n_serial: int):
    """Oracle (FP64) on the same workload: A + b (matrix, vector, lifting,
    set_bc) at n_full^3 x 6 cells on every core this process may use, the
    entities coloured so that no two of a colour share a row (colouring and
    sparsity are set-up, not timed); and serially on n_serial^3 x 6 cells."""
    from mifem import problems
    from oracle import cpu

    def run(n, threads):
        P = problems.elasticity_p2(n)
        Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
        rp, cols, _, _ = Oa.sparsity()
        if threads > 1:
            Oa.colouring()
            OL.colouring()
        t0 = time.perf_counter()
        if threads > 1:
            Oa.assemble_matrix_mt(rp, cols, P.bcs, threads=threads)
            b = OL.assemble_vector_mt(threads=threads)
            Oa.apply_lifting(b, P.bcs, threads=threads)
        else:
            Oa.assemble_matrix(rp, cols, P.bcs)
            b = OL.assemble_vector()
            Oa.apply_lifting(b, P.bcs)
        cpu.set_bc(b, P.bcs, P.V)
        return P, time.perf_counter() - t0

    threads = _cpu_threads()
    P, dtn = run(n_full, threads)
    Ps, dt1 = run(n_serial, 1)
    return {"value": P.V.size / dtn, "unit": "DoF/s", "cores": threads, "kind": "port",
            "sample": f"elasticity_p2 {n_full}^3x6 tets ({P.mesh.num_cells} cells, "
                      f"{P.V.size} dofs): A+b in {dtn:.2f} s on {threads} threads (oracle, C + "
                      f"OpenMP, coloured cells; set_bc serial); serial figure: {n_serial}^3x6 "
                      f"({Ps.V.size} dofs) in {dt1:.2f} s on one core",
            "serial_value": Ps.V.size / dt1,
            **_cpu_fields()}


def c3_record(args, oracle: bool = True) -> dict:
    """Q3 Helmholtz (k=5) on 48^3 jittered hexes: FP64 vs FP32 vs BF16 matrix-core paths.

    speedup = FP64 GPU path time / FP32 matrix-core time for A (each a whole
    matrix launch: per-cell record pass + owner kernel).  Errors are taken
    against the oracle's FP64 assembly of the same A and b (its coloured
    multi-core loops, timed here as the CPU baseline): relative Frobenius
    error of A, l2 error of b, and the relative error of the solution of
    Jacobi-CG at a fixed tolerance with each matrix against the solution with
    the oracle's matrix.  The <= 1e-6 gate applies to the assembled A (and b);
    the CG solution error is A's error amplified by the condition number of
    this near-resonant Helmholtz system (k^2 = 25 against the first Dirichlet
    eigenvalue 3 pi^2), and is reported beside it.
    """
    import numpy as np
    import torch
    import cudolfinx as cufem
    from cudolfinx import _lib
    from cudolfinx.solve import pcg
    from mifem import problems

    from cudolfinx.autotune import HEX_CANDIDATES

    t0 = time.perf_counter()
    P = problems.helmholtz_q3(args.n_hex)
    log(f"C3 problem: {P.mesh.num_cells} cells, {P.V.size} dofs ({time.perf_counter() - t0:.1f} s)")
    L = cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    b = asm.create_vector(L)
    bcs = asm.pack_bcs(P.bcs)
    stream = torch.cuda.current_stream()
    lib = _lib.lib()
    b0 = bcs._get_cpp_bcs(P.V)
    A = None

    def launch_ms(form, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            _lib.check(lib.mf_assemble_matrix(form.handle, A.handle, b0.handle, b0.handle, 1.0, 1,
                                              C.c_void_p(stream.cuda_stream)))
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    # autotuned shape per precision (one hipRTC build per candidate holds the
    # FP64, FP32 and BF16 kernels): chunk values, threads per workgroup, cells
    # per record batch
    cands = [{}] + [c for c in HEX_CANDIDATES if c != {"chunk_values": 4096}] if args.c3_tune else [{}]
    best, tuning = {}, []
    for cand in cands:
        a_c = cufem.form(P.a, options={"matrix_kernel": 2, "precision": 64, **cand})
        if A is None:
            A = asm.create_matrix(a_c)
        row = {"options": cand}
        for prec in (64, 32, 16):
            a_c.set_option("precision", prec)
            asm.assemble_matrix(a_c, A, bcs=bcs)
            asm.synchronize()
            row[str(prec)] = launch_ms(a_c, 3)
            if prec not in best or row[str(prec)] < best[prec][0]:
                best[prec] = (row[str(prec)], cand)
        tuning.append(row)
        log(f"C3 shape {cand}: FP64 {row['64']:.3f} ms, FP32 {row['32']:.3f} ms, BF16 {row['16']:.3f} ms")
    vals, times, shapes, forms_ = {}, {}, {}, {}
    for prec in (64, 32, 16):
        a = forms_[prec] = cufem.form(P.a, options={"matrix_kernel": 2, "precision": prec,
                                                    **best[prec][1]})
        buf = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
        _lib.check(lib.mf_matrix_borrow_values(A.handle, C.c_void_p(buf.data_ptr())))
        asm.assemble_matrix(a, A, bcs=bcs)
        asm.synchronize()
        times[prec] = launch_ms(a, max(3, args.steps // 2))
        vals[prec] = buf
        shapes[str(prec)] = best[prec][1]
        log(f"C3 precision {prec} (shape {best[prec][1]}): matrix {times[prec]:.3f} ms")
    # the reference's scheme in FP64: one thread per cell, full 64 x 64 tabulate_tensor,
    # FP64 atomics into the CSR (per-entity kernel)
    a = forms_[64]
    a.set_option("matrix_kernel", 1)
    gbuf = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    _lib.check(lib.mf_matrix_borrow_values(A.handle, C.c_void_p(gbuf.data_ptr())))
    asm.assemble_matrix(a, A, bcs=bcs)
    asm.synchronize()
    generic_ms = launch_ms(a, 2)
    generic_diff = float(torch.linalg.norm(gbuf - vals[64]) / torch.linalg.norm(vals[64]))
    del gbuf
    log(f"C3 per-entity FP64 kernel: matrix {generic_ms:.3f} ms")
    a.set_option("matrix_kernel", 2)
    asm.assemble_vector(L, b)
    asm.apply_lifting(b, [a], [bcs])
    asm.set_bc(b, bcs, P.V)
    asm.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        asm.assemble_vector(L, b)
        asm.apply_lifting(b, [a], [bcs])
        asm.set_bc(b, bcs, P.V)
    e1.record(stream)
    e1.synchronize()
    bms = e0.elapsed_time(e1) / 3
    bt = torch.from_numpy(b.to_host()).cuda()
    rec = {"workload": f"helmholtz_q3 {args.n_hex}^3 jittered hexes, k=5, 64 Gauss points",
           "cells": P.mesh.num_cells, "dofs": P.V.size, "nnz": A.nnz,
           "matrix_ms": {str(p): times[p] for p in times}, "b_region_ms": bms,
           "shapes": shapes, "autotune": tuning,
           "speedup_fp32_vs_fp64": times[64] / times[32],
           "speedup_bf16_vs_fp64": times[64] / times[16],
           "fp64_per_entity_matrix_ms": generic_ms,
           "speedup_fp32_vs_fp64_per_entity": generic_ms / times[32],
           "rel_diff_per_entity_vs_owner_fp64": generic_diff,
           "note": "FP64 = vector-ALU FP64 contraction, FP32 / BF16 = matrix cores (FP32 "
                   "accumulation), geometry and CSR accumulation FP64 in all; b on the FP64 "
                   "vector path in every variant"}
    flops = 2.175e6 * P.mesh.num_cells
    rec["fp32_tflops"] = flops / (times[32] * 1e-3) / 1e12
    rec["fp64_tflops"] = flops / (times[64] * 1e-3) / 1e12
    rec["bf16_tflops"] = flops / (times[16] * 1e-3) / 1e12
    if not oracle:
        return rec
    from oracle import cpu
    threads = _cpu_threads()
    Oa, OL = cpu.OracleAssembler(P.a), cpu.OracleAssembler(P.L)
    orp, ocols, _, _ = Oa.sparsity()
    rp, cols, _ = A.csr()
    rec["csr_bit_exact"] = bool(np.array_equal(rp, orp) and np.array_equal(cols, ocols))
    del rp, cols
    Oa.colouring()
    OL.colouring()
    t1 = time.perf_counter()
    ovals = Oa.assemble_matrix_mt(orp, ocols, P.bcs, threads=threads)
    t_oracle = time.perf_counter() - t1
    ob = OL.assemble_vector_mt(threads=threads)
    Oa.apply_lifting(ob, P.bcs, threads=threads)
    cpu.set_bc(ob, P.bcs, P.V)
    log(f"C3 oracle A on {threads} threads: {t_oracle:.1f} s")
    ref = torch.from_numpy(ovals).cuda()
    del ovals
    rnorm = torch.linalg.norm(ref)
    rec["rel_err_A"] = {str(p): float(torch.linalg.norm(vals[p] - ref) / rnorm) for p in vals}
    obt = torch.from_numpy(ob).cuda()
    rec["rel_err_b"] = float(torch.linalg.norm(bt - obt) / torch.linalg.norm(obt))
    # fixed-tolerance Jacobi-CG with each matrix (native CSR SpMV), against the
    # solution with the oracle's matrix and vector
    sol = {}
    for p, v, rhs in (("oracle", ref, obt), ("64", vals[64], bt), ("32", vals[32], bt),
                      ("16", vals[16], bt)):
        _lib.check(lib.mf_matrix_borrow_values(A.handle, C.c_void_p(v.data_ptr())))
        torch.cuda.synchronize()
        t_cg = time.perf_counter()
        x, it, res = pcg(A, rhs, rtol=1e-12, maxiter=args.cg_maxiter)
        torch.cuda.synchronize()
        sol[p] = (x, it, res, (time.perf_counter() - t_cg) / max(it, 1) * 1e3)
        log(f"C3 CG with the {p} matrix: {it} iterations, relative residual {res:.2e}")
    xref = sol["oracle"][0]
    xn = torch.linalg.norm(xref)
    rec["rel_err_cg_solution"] = {p: float(torch.linalg.norm(sol[p][0] - xref) / xn)
                                  for p in ("64", "32", "16")}
    rec["cg"] = {"rtol": 1e-12, "iterations": {p: sol[p][1] for p in sol},
                 "ms_per_iteration": {p: sol[p][3] for p in sol}}
    rec["gate"] = {"bound": 1e-6, "on": "rel_err_A and rel_err_b (assembled system)",
                   "fp32_meets": rec["rel_err_A"]["32"] <= 1e-6 and rec["rel_err_b"] <= 1e-6,
                   "fp32_cg_solution_meets": rec["rel_err_cg_solution"]["32"] <= 1e-6}
    rec["cpu_baseline"] = {"value": P.mesh.num_cells / t_oracle, "unit": "cells/s (A)",
                           "cores": threads, "kind": "port",
                           "sample": f"the full workload: oracle FP64 A in {t_oracle:.2f} s on "
                                     f"{threads} threads (C + OpenMP, coloured cells)",
                           **_cpu_fields()}
    del ref
    return rec


def bench_c3(args):
    """Q3 Helmholtz precision comparison as its own JSON line: value = FP64-path
    time / FP32-path time for A, with the errors of c3_record."""
    rec = c3_record(args, oracle=not args.no_cpu)
    flops = 2.175e6 * rec["cells"]
    out = {
        "metric": "mixed-prec vs FP64 speedup at rel err<=1e-6 (Q3 Helmholtz matrix assembly)",
        "value": rec["speedup_fp32_vs_fp64"],
        "unit": "x (FP64 time / FP32 time)",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": rec["matrix_ms"]["32"] + rec["b_region_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (MFMA contraction; FP64 geometry and accumulation)",
        "data": "synthetic (seeded jitter)",
        "config": {k: v for k, v in rec.items() if k != "cpu_baseline"},
        "roofline": {"bound": "mfma", "achieved": rec["fp32_tflops"], "peak": 157.3,
                     "unit": "TFLOP/s", "frac": rec["fp32_tflops"] / 157.3,
                     "traffic": (measured_traffic("c3_helmholtz_q3_48_traffic.md")
                                 if args.n_hex == 48 else None),
                     "kernel": "matrix launch: mf_hprep_cell_0 + mf_howner32_cell_0",
                     "fp64_achieved": rec["fp64_tflops"], "fp64_peak": 78.6,
                     "algorithmic_flops": flops},
        "cpu_baseline": rec.get("cpu_baseline"),
    }
    m = measured_peaks("mfma")
    m["frac_of_sgemm"] = out["roofline"]["achieved"] / m["sgemm_tflops"]
    m["fp64_frac_of_dgemm"] = out["roofline"]["fp64_achieved"] / m["dgemm_tflops"]
    out["roofline"]["measured"] = m
    print(json.dumps(out), flush=True)


def bench_c4(args):
    """SIPG DG1 p-Laplacian Newton on 48^3 x 6 jittered tets: the reassembly of the
    three forms of a Newton step -- Jacobian J(u), residual F(u) and the energy
    functional E(u) whose gradient is F -- captured once as a HIP graph.

    value = ms per graph replay (pack u for the three forms + J + F + E, all
    integral types); the eager time of the same sequence, each form's kernels
    alone (HIP events), and the 10-step Newton history (|F| and E; line search
    on E) are reported beside it.
    """
    import numpy as np
    import torch
    import cudolfinx as cufem
    from cudolfinx import _lib
    from cudolfinx.solve import pcg
    from mifem import problems
    from mifem.dgforms import sipg_energy

    t0 = time.perf_counter()
    P = problems.dg_plaplace(args.n_dg)
    n = P.V.size
    mesh = P.mesh
    nint, next_ = len(mesh.interior_facets), len(mesh.exterior_facets)
    log(f"C4 problem: {mesh.num_cells} cells, {n} dofs, {nint} interior + {next_} exterior facets "
        f"({time.perf_counter() - t0:.1f} s)")
    lib = _lib.lib()
    s = torch.cuda.Stream()
    J, F = cufem.form(P.a), cufem.form(P.L)
    Ef = cufem.form(sipg_energy(P.V, P.extras["u"]))
    asm = cufem.CUDAAssembler(stream=s)
    A = asm.create_matrix(J)
    ud = torch.from_numpy(P.extras["u"].x.copy()).cuda()
    for f in (J, F, Ef):
        f.set_coefficient_device(0, ud.data_ptr(), n)
    ev = torch.zeros(1, dtype=torch.float64, device="cuda")
    vals = torch.empty(A.nnz, dtype=torch.float64, device="cuda")
    _lib.check(lib.mf_matrix_borrow_values(A.handle, C.c_void_p(vals.data_ptr())))
    bt = torch.empty(n, dtype=torch.float64, device="cuda")
    dbc = asm.pack_bcs([])
    nob = dbc._get_cpp_bcs(P.V).handle
    sh = C.c_void_p(s.cuda_stream)

    def seq():
        _lib.check(lib.mf_form_pack_coefficients(J.handle, None, 0, sh))
        _lib.check(lib.mf_assemble_matrix(J.handle, A.handle, nob, nob, 1.0, 1, sh))
        _lib.check(lib.mf_form_pack_coefficients(F.handle, None, 0, sh))
        _lib.check(lib.mf_assemble_vector(F.handle, C.c_void_p(bt.data_ptr()), n, 1, sh))
        _lib.check(lib.mf_form_pack_coefficients(Ef.handle, None, 0, sh))
        _lib.check(lib.mf_assemble_scalar_device(Ef.handle, C.c_void_p(ev.data_ptr()), 1, sh))

    def energy_and_residual():
        _lib.check(lib.mf_form_pack_coefficients(F.handle, None, 0, sh))
        _lib.check(lib.mf_assemble_vector(F.handle, C.c_void_p(bt.data_ptr()), n, 1, sh))
        _lib.check(lib.mf_form_pack_coefficients(Ef.handle, None, 0, sh))
        _lib.check(lib.mf_assemble_scalar_device(Ef.handle, C.c_void_p(ev.data_ptr()), 1, sh))

    seq()
    s.synchronize()
    log(f"setup done: nnz={A.nnz} ({time.perf_counter() - t0:.1f} s)")

    def timed(fn, reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    for _ in range(args.warmup):
        seq()
    sep_eager_ms = timed(seq, args.steps)
    gs = cufem.CUDAGraph(s)
    with gs:
        seq()
    for _ in range(args.warmup):
        gs.launch()
    s.synchronize()
    sep_graph_ms = timed(gs.launch, args.steps)

    # the three forms in one persistent launch (cells per workgroup autotuned)
    from cudolfinx.autotune import autotune_system

    class _Vec:
        ptr, size = C.c_void_p(bt.data_ptr()), n
    t1 = time.perf_counter()
    system, sblock, stune = autotune_system(J, F, Ef, asm, A, _Vec, ev, stream=s, verbose=True)
    log(f"fused system: block {sblock} ({time.perf_counter() - t1:.1f} s)")
    bptr, eptr = C.c_void_p(bt.data_ptr()), C.c_void_p(ev.data_ptr())

    def fused():
        _lib.check(lib.mf_form_pack_coefficients(J.handle, None, 0, sh))
        _lib.check(lib.mf_system_assemble(system.handle, A.handle, nob, nob, 1.0, bptr, n, eptr, 1, sh))

    for _ in range(args.warmup):
        fused()
    eager_ms = timed(fused, args.steps)
    g = cufem.CUDAGraph(s)
    with g:
        fused()
    for _ in range(args.warmup):
        g.launch()
    s.synchronize()
    t = time.perf_counter()
    graph_ms = timed(g.launch, args.steps)
    wall_ms = (time.perf_counter() - t) / args.steps * 1e3
    sys_ms = timed(lambda: lib.mf_system_assemble(system.handle, A.handle, nob, nob, 1.0, bptr, n,
                                                  eptr, 1, sh), args.steps)
    jk_ms = timed(lambda: lib.mf_assemble_matrix(J.handle, A.handle, nob, nob, 1.0, 1, sh), args.steps)
    fk_ms = timed(lambda: lib.mf_assemble_vector(F.handle, C.c_void_p(bt.data_ptr()), n, 1, sh),
                  args.steps)
    ek_ms = timed(lambda: lib.mf_assemble_scalar_device(Ef.handle, C.c_void_p(ev.data_ptr()), 1, sh),
                  args.steps)
    log(f"fused step: eager {eager_ms:.3f} ms, graph {graph_ms:.3f} ms (system kernel {sys_ms:.3f} ms); "
        f"separate forms: eager {sep_eager_ms:.3f} ms, graph {sep_graph_ms:.3f} ms (J {jk_ms:.3f}, "
        f"F {fk_ms:.3f}, E {ek_ms:.3f} ms)")

    # Newton with an Armijo line search on the energy (F = grad E, J = Hessian,
    # J is SPD, so the Newton direction descends); one graph replay per step
    history, energies, cg_ms = [], [], []
    converged = False
    for it in range(args.newton):
        g.launch()
        s.synchronize()
        r0, e0 = float(torch.linalg.norm(bt)), float(ev.item())
        history.append(r0)
        energies.append(e0)
        if r0 <= 1e-10 * history[0]:
            converged = True
            break                          # converged: the energy no longer resolves a step
        s.synchronize()
        t_cg = time.perf_counter()
        du, cg_it, cg_res = pcg(A, -bt.clone(), rtol=1e-10, maxiter=args.cg_maxiter)
        torch.cuda.synchronize()
        cg_ms.append((time.perf_counter() - t_cg) / max(cg_it, 1) * 1e3)
        slope = float(torch.dot(bt, du))
        u_old = ud.clone()
        alpha = 1.0
        for _ in range(12):
            ud.copy_(u_old + alpha * du)
            energy_and_residual()
            s.synchronize()
            if float(ev.item()) <= e0 + 1e-4 * alpha * slope:
                break
            alpha *= 0.5
        log(f"Newton {it}: |F| = {r0:.3e}, E = {e0:.12e}, CG {cg_it} its (res {cg_res:.1e}, "
            f"{cg_ms[-1]:.3f} ms/it), step {alpha}")
    if not converged:
        g.launch()
        s.synchronize()
        history.append(float(torch.linalg.norm(bt)))
        energies.append(float(ev.item()))

    nc = mesh.num_cells
    # values written once, b written once (8 n), items (16 B per cell + 2 per interior facet +
    # 1 per exterior facet), cell pointers, geometry, geometry dofmap, packed state u (4 per cell)
    alg = (8 * A.nnz + 8 * n + 16 * (nc + 2 * nint + next_) + 4 * nc + 24 * mesh.num_vertices
           + 16 * nc + 8 * 4 * nc)
    out = {
        "metric": "hipGraph-captured Newton reassembly step time (SIPG DG1 J + F + energy, one launch)",
        "value": graph_ms, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": graph_ms, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded jitter, interpolated state)",
        "config": {"workload": f"sipg_plaplace DG1 {args.n_dg}^3x6 jittered tets, p=3, Robin alpha=2",
                   "cells": nc, "dofs": n, "nnz": A.nnz, "interior_facets": nint,
                   "exterior_facets": next_, "eager_ms": eager_ms, "graph_wall_ms": wall_ms,
                   "system_kernel_ms": sys_ms, "system_block": sblock,
                   "system_autotune": [{"block": b_, "ms": m_} for b_, m_ in stune],
                   "separate_forms_eager_ms": sep_eager_ms, "separate_forms_graph_ms": sep_graph_ms,
                   "J_kernel_ms": jk_ms, "F_kernel_ms": fk_ms, "E_kernels_ms": ek_ms,
                   "newton_residuals": history, "newton_energies": energies,
                   "cg_ms_per_iteration": cg_ms},
        "roofline": {"bound": "hbm", "achieved": alg / (sys_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / (sys_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": measured_traffic("c4_sipg_dg1_48_traffic.md") if args.n_dg == 48 else None,
                     "kernel": "mf_dgsys (J + F + E over cells and facets, one launch)",
                     "algorithmic_bytes": alg},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_c4(),
    }
    m = measured_peaks("hbm")
    m["frac_of_copy"] = out["roofline"]["achieved"] / m["copy_gbs"]
    out["roofline"]["measured"] = m
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """512 independent Poisson problems (n ~ U{4..24}, degree ~ U{1,2,3}, seed 3;
    kappa ~ lognormal(0, 1) per cell), merged per degree into block-diagonal
    systems whose CSR blocks are the per-instance matrices.

    value = instances assembled per second (A + b with lifting and set_bc for
    all 512); per-instance errors against the oracle on a seeded sample.
    """
    import numpy as np
    import torch
    import cudolfinx as cufem
    from cudolfinx import _lib
    from mifem import batch, forms
    from mifem.fem import Constant, Function, dirichletbc, functionspace, locate_dofs_topological

    t0 = time.perf_counter()
    B = batch.many_poisson(args.instances, seed=3)
    ncells = sum(g.mesh.num_cells for g in B.groups)
    ndofs = sum(g.V.size for g in B.groups)
    log(f"C5 batch: {B.num_instances} instances, {ncells} cells, {ndofs} dofs "
        f"({time.perf_counter() - t0:.1f} s)")
    asm = cufem.CUDAAssembler()
    sys_ = []
    for g in B.groups:
        # degree >= 2: element tensors in FP32 (FP64 geometry, LDS accumulation and CSR
        # values); --c5-fp64 keeps every group in FP64
        a = cufem.form(g.a, options={"precision": 32} if g.degree >= 2 and not args.c5_fp64 else None)
        L = cufem.form(g.L)
        A, b = asm.create_matrix(a), asm.create_vector(L)
        bcs = asm.pack_bcs(g.bcs)
        sys_.append((g, a, L, A, b, bcs))
        log(f"  group P{g.degree}: {len(g.instances)} instances, {g.mesh.num_cells} cells, "
            f"{g.V.size} dofs, nnz {A.nnz} ({time.perf_counter() - t0:.1f} s)")

    # every group's matrix from one persistent launch (after one record pass per group)
    mats = [A for g, a, L, A, b, bcs in sys_]
    gbcs = [bcs for g, a, L, A, b, bcs in sys_]
    batch_tune = None
    if args.no_tune:
        fused = cufem.CUDABatch([a for g, a, L, A, b, bcs in sys_])
    else:
        # member shapes (rows per workgroup, pairs per chunk) tuned together
        from cudolfinx.autotune import autotune_batch
        base = [{"precision": 32} if g.degree >= 2 and not args.c5_fp64 else {} for g in B.groups]
        _, fused, _, batch_tune = autotune_batch([g.a for g in B.groups], asm, mats, gbcs,
                                                 base_options=base, verbose=True)
        log(f"batch autotuned ({time.perf_counter() - t0:.1f} s)")
    # every group's right-hand side (cell integrals) in one launch, the
    # Dirichlet values of all groups in one launch
    rhs = cufem.CUDABatch([L for g, a, L, A, b, bcs in sys_])
    vecs = [b for g, a, L, A, b, bcs in sys_]

    def step():
        fused.assemble(asm, mats, gbcs)
        rhs.assemble_vectors(asm, vecs)
        for g, a, L, A, b, bcs in sys_:
            asm.apply_lifting(b, [a], [bcs])
        rhs.set_bc(asm, vecs, gbcs)

    step()
    asm.synchronize()
    log(f"setup done ({time.perf_counter() - t0:.1f} s)")
    for _ in range(args.warmup):
        step()
    asm.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    asm.synchronize()
    ms = (time.perf_counter() - t) / args.steps * 1e3
    # matrix launches alone (HIP events on the null stream they run on)
    lib = _lib.lib()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    n = len(sys_)
    hA = (C.c_void_p * n)(*[A.handle.value for A in mats])
    hB = (C.c_void_p * n)(*[bcs._get_cpp_bcs(g.V).handle.value for g, a, L, A, b, bcs in sys_])
    e0.record(stream)
    for _ in range(args.steps):
        _lib.check(lib.mf_batch_assemble(fused.handle, hA, hB, 1.0, 1, None))
    e1.record(stream)
    e1.synchronize()
    kms = e0.elapsed_time(e1) / args.steps
    # the same matrices one owner-computes launch per group, for comparison
    e0.record(stream)
    for _ in range(args.steps):
        for g, a, L, A, b, bcs in sys_:
            b0 = bcs._get_cpp_bcs(g.V)
            _lib.check(lib.mf_assemble_matrix(a.handle, A.handle, b0.handle, b0.handle, 1.0, 1, None))
    e1.record(stream)
    e1.synchronize()
    kms_sep = e0.elapsed_time(e1) / args.steps
    alg = 0
    for g, a, L, A, b, bcs in sys_:
        nd = g.V.element.ndofs
        nc = g.mesh.num_cells
        alg += (8 * A.nnz + 4 * A.nnz + 4 * (A.nrows + 1) + 24 * g.mesh.num_vertices + 16 * nc
                + 8 * nc + 4 * nd * nc)
    log(f"step {ms:.3f} ms ({B.num_instances / (ms * 1e-3):.0f} assemblies/s); matrices in one "
        f"launch {kms:.3f} ms (one launch per group: {kms_sep:.3f} ms)")

    # per-instance errors on a seeded sample of instances small enough for the oracle
    sys.path.insert(0, str(ROOT))
    from oracle import cpu
    rng = np.random.default_rng(11)
    errs = []
    for g, a, L, A, b, bcs in sys_:
        rp = np.empty(A.nrows + 1, dtype=np.int32)
        d_rp, d_cols, d_vals = A.device_arrays()
        ctx = cufem.get_cuda_context()
        _lib.check(lib.mf_memcpy_dtoh(ctx.handle, rp.ctypes.data_as(C.c_void_p), C.c_void_p(d_rp),
                                      rp.nbytes, None))
        bh = b.to_host()
        small = [j for j, i in enumerate(g.instances) if B.sizes[i] <= args.c5_oracle_n]
        for j in rng.permutation(small)[:args.c5_sample]:
            i = g.instances[j]
            r0, r1 = g.dof_offsets[j], g.dof_offsets[j + 1]
            p0, p1 = int(rp[r0]), int(rp[r1])
            cols = np.empty(p1 - p0, dtype=np.int32)
            vals = np.empty(p1 - p0, dtype=np.float64)
            _lib.check(lib.mf_memcpy_dtoh(ctx.handle, cols.ctypes.data_as(C.c_void_p),
                                          C.c_void_p(d_cols + 4 * p0), cols.nbytes, None))
            _lib.check(lib.mf_memcpy_dtoh(ctx.handle, vals.ctypes.data_as(C.c_void_p),
                                          C.c_void_p(d_vals + 8 * p0), vals.nbytes, None))
            mesh, V, kap = batch.poisson_instance(B.sizes[i], g.degree)
            D = functionspace(mesh, ("DG", 0))
            kf = Function(D)
            kf.set_cellwise(kap)
            bc = dirichletbc(0.0, locate_dofs_topological(V, 2, mesh.exterior_facets), V)
            o = cpu.OracleAssembler(forms.stiffness(V, kf))
            orp, ocols, _, _ = o.sparsity()
            ov = o.assemble_matrix(orp, ocols, [bc])
            ob = cpu.OracleAssembler(forms.source(V, Constant([1.0]))).assemble_vector()
            cpu.set_bc(ob, [bc], V)
            exact = np.array_equal(rp[r0:r1 + 1] - p0, orp) and np.array_equal(cols - r0, ocols)
            ea = float(np.linalg.norm(vals - ov) / np.linalg.norm(ov))
            eb = float(np.linalg.norm(bh[r0:r1] - ob) / np.linalg.norm(ob))
            errs.append({"instance": int(i), "n": int(B.sizes[i]), "degree": int(g.degree),
                         "csr_exact": bool(exact), "rel_err_A": ea, "rel_err_b": eb})
            log(f"  instance {i} (n={B.sizes[i]}, P{g.degree}): csr exact {exact}, A {ea:.2e}, b {eb:.2e}")
    out = {
        "metric": "independent-instance assemblies/s (A+b, 512 Poisson meshes, P1-P3)",
        "value": B.num_instances / (ms * 1e-3), "unit": "assemblies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.c5_fp64 else "f32 element tensors (P2, P3) / f64 (P1)",
        "data": "synthetic (seeded sizes, degrees and lognormal conductivity)",
        "config": {"workload": f"many_poisson {B.num_instances} instances n~U{{4..24}} P1-P3",
                   "cells": ncells, "dofs": ndofs,
                   "groups": [{"degree": g.degree, "instances": len(g.instances), "nnz": A.nnz,
                               "precision": 32 if g.degree >= 2 and not args.c5_fp64 else 64}
                              for g, a, L, A, b, bcs in sys_],
                   "dof_per_s": ndofs / (ms * 1e-3), "matrix_ms": kms, "matrix_ms_per_group_launches": kms_sep,
                   "batch_autotune": batch_tune,
                   "sampled_errors": errs,
                   "max_rel_err_A": max((e["rel_err_A"] for e in errs), default=None),
                   "max_rel_err_b": max((e["rel_err_b"] for e in errs), default=None)},
        "roofline": {"bound": "hbm", "achieved": alg / (kms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "traffic": (measured_traffic("c5_many_poisson_512_traffic.md")
                                 if args.instances == 512 else None), "kernel": "mf_owner_batch (all groups in one launch) + one record pass per group",
                     "algorithmic_bytes": alg},
        "cpu_baseline": None if args.no_cpu else cpu_baseline_c5(B),
    }
    m = measured_peaks("hbm")
    m["frac_of_copy"] = out["roofline"]["achieved"] / m["copy_gbs"]
    out["roofline"]["measured"] = m
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64, help="cubes per direction (64: 1.57M cells)")
    ap.add_argument("--kernel", type=int, default=0, help="matrix kernel: 0 auto, 1 atomic, 2 owner")
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="oracle size for the multi-core CPU baseline (0: skip)")
    ap.add_argument("--cpu-serial", type=int, default=32, help="oracle size for the one-core figure")
    ap.add_argument("--no-c3", action="store_true",
                    help="c2: skip the Q3 mixed-precision sub-record")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the owner-kernel autotuner")
    ap.add_argument("--workload", default="c2", choices=["c2", "c3", "c4", "c5"],
                    help="c2: P2 elasticity A+b (headline); c3: Q3 Helmholtz precision "
                         "comparison; c4: SIPG DG1 Newton reassembly (HIP graph); "
                         "c5: 512 independent Poisson instances")
    ap.add_argument("--instances", type=int, default=512, help="instances for c5")
    ap.add_argument("--c5-sample", type=int, default=4, help="oracle-checked instances per group")
    ap.add_argument("--c5-oracle-n", type=int, default=10, help="largest n checked by the oracle")
    ap.add_argument("--c5-fp64", action="store_true",
                    help="c5: every group in FP64 (default: degree >= 2 element tensors in FP32)")
    ap.add_argument("--n-dg", type=int, default=48, help="cubes per direction for c4")
    ap.add_argument("--newton", type=int, default=10, help="Newton steps for c4 (after the timed region)")
    ap.add_argument("--n-hex", type=int, default=48, help="hexahedra per direction for c3")
    ap.add_argument("--no-c3-tune", dest="c3_tune", action="store_false",
                    help="c3: the default shape only (no autotuning)")
    ap.add_argument("--cg-maxiter", type=int, default=4000)
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (PMC traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py runs the assembly on one GPU; --gpus above 1 is "
                                   "not supported", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "c3":
        bench_c3(args)
        return
    if args.workload == "c4":
        bench_c4(args)
        return
    if args.workload == "c5":
        bench_c5(args)
        return

    import numpy as np
    import torch
    import cudolfinx as cufem
    from cudolfinx import _lib
    from mifem import problems

    t0 = time.perf_counter()
    P = problems.elasticity_p2(args.n)
    log(f"problem: {P.mesh.num_cells} cells, {P.V.size} dofs ({time.perf_counter() - t0:.1f} s)")
    a = cufem.form(P.a, options={"matrix_kernel": args.kernel})
    L = cufem.form(P.L)
    asm = cufem.CUDAAssembler()
    A = asm.create_matrix(a)
    b = asm.create_vector(L)
    bcs = asm.pack_bcs(P.bcs)
    log(f"setup done: nnz={A.nnz} ({time.perf_counter() - t0:.1f} s)")
    tuned = None
    if args.kernel != 1 and not args.no_tune:
        from cudolfinx.autotune import autotune_matrix
        opts, a, timings = autotune_matrix(P.a, asm, A, bcs, verbose=True)
        tuned = {"options": opts, "candidates": timings}
        log(f"autotuned owner kernel: {opts} ({time.perf_counter() - t0:.1f} s)")

    def step():
        asm.assemble_matrix(a, A, bcs=bcs)
        A.assemble()
        asm.assemble_vector(L, b)
        asm.apply_lifting(b, [a], [bcs])
        asm.set_bc(b, bcs, P.V)

    for _ in range(args.warmup):
        step()
    asm.synchronize()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        step()
    asm.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    ms = wall / args.steps * 1e3
    log(f"{args.steps} steps: {ms:.3f} ms/step")

    # dominant kernel: the matrix assembly launch alone, timed with HIP events
    # on the stream it runs on (the null stream, torch's default here)
    lib = _lib.lib()
    b0 = bcs._get_cpp_bcs(P.V)
    stream = torch.cuda.current_stream()
    reps = max(5, args.steps)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(reps):
        _lib.check(lib.mf_assemble_matrix(a.handle, A.handle, b0.handle, b0.handle, 1.0, 1,
                                          C.c_void_p(stream.cuda_stream)))
    ev1.record(stream)
    ev1.synchronize()
    kms = ev0.elapsed_time(ev1) / reps
    vb = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    vb[0].record(stream)
    for _ in range(reps):
        asm.assemble_vector(L, b)
        asm.apply_lifting(b, [a], [bcs])
        asm.set_bc(b, bcs, P.V)
    vb[1].record(stream)
    vb[1].synchronize()
    vms = vb[0].elapsed_time(vb[1]) / reps
    alg = algorithmic_bytes(P, A)
    achieved = alg / (kms * 1e-3) / 1e9
    log(f"matrix kernel {kms:.3f} ms -> {achieved:.0f} GB/s algorithmic; b region {vms:.3f} ms")

    # property check at full size: symmetry of sampled rows, relative to the matrix scale
    rp, cols, vals = A.csr()
    rng = np.random.default_rng(0)
    rows = rng.choice(A.nrows, size=200, replace=False)
    vmax = float(np.abs(vals).max())
    asym = 0.0
    for r in rows:
        for p in range(rp[r], rp[r + 1]):
            c = cols[p]
            q = rp[c] + np.searchsorted(cols[rp[c]:rp[c + 1]], r)
            asym = max(asym, abs(vals[p] - vals[q]) / vmax)
    log(f"sampled symmetry defect {asym:.2e} (relative to max |A|)")

    out = {
        "metric": METRIC,
        "value": P.V.size / (ms * 1e-3),
        "unit": "DoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Lame fields, generated mesh)",
        "config": {"workload": f"elasticity_p2 {args.n}^3x6 tets A+b (matrix+vector+lifting+set_bc)",
                   "cells": P.mesh.num_cells, "dofs": P.V.size, "nnz": A.nnz,
                   "matrix_kernel": {0: "auto(owner-computes)", 1: "per-entity atomic",
                                     2: "owner-computes"}[args.kernel],
                   "matrix_ms": kms, "b_region_ms": vms, "autotune": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic() if args.n == 64 and args.kernel != 1 else None,
                     "kernel": ("matrix launch: mf_prep_cell_0 + mf_owner_cell_0"
                                if args.kernel != 1 else "mf_mat_cell_0"),
                     "algorithmic_bytes": alg},
        "cpu_baseline": None,
    }
    m = measured_peaks("hbm")
    m["frac_of_copy"] = achieved / m["copy_gbs"]
    out["roofline"]["measured"] = m
    if args.copy_calib:
        nbytes = args.copy_calib * (1 << 20)
        src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
        src.fill_(1.0)                    # writes exactly nbytes (WRITE_SIZE calibration)
        tot = torch.sum(src)              # reads exactly nbytes (FETCH_SIZE calibration)
        torch.cuda.synchronize()
        log(f"traffic calibration: fill + sum of {args.copy_calib} MiB (sum={float(tot):.0f})")
    if not args.no_cpu and args.cpu_sample > 0:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample, args.cpu_serial)
    if not args.no_c3:
        # the metric's mixed-precision half: Q3 FP32 / BF16 matrix cores against
        # the FP64 GPU path, errors against the FP64 oracle
        log("c3: Q3 Helmholtz mixed precision ...")
        c3 = c3_record(args, oracle=not args.no_cpu)
        m = measured_peaks("mfma")
        c3["mfma_roofline"] = {"bound": "mfma", "achieved": c3["fp32_tflops"], "peak": 157.3,
                               "unit": "TFLOP/s", "frac": c3["fp32_tflops"] / 157.3,
                               "measured": m, "frac_of_sgemm": c3["fp32_tflops"] / m["sgemm_tflops"],
                               "fp64_frac_of_dgemm": c3["fp64_tflops"] / m["dgemm_tflops"]}
        out["c3"] = c3
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
