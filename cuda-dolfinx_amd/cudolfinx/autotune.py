"""Autotuners: the cooperative owner-computes matrix kernel, the batched
(several element groups in one launch) matrix kernel and the fused
Newton-system kernel.

Picks the kernel shape -- LDS value budget per workgroup ("chunk_values",
i.e. how many CSR rows a workgroup owns), pairs per chunk ("owner_pairs"),
threads per workgroup ("owner_block", "hex_block" for hexahedra), cells per
record batch ("owner_nb"), an optional register cap ("owner_wpe": waves per
SIMD), the warp-specialised variant ("owner_pipe") -- by timing each
candidate on the actual form, matrix and boundary conditions.  Every
candidate is a separate hipRTC build of the same form; the matrix (sparsity)
is shared.  Results are cached in a JSON table keyed by the form description
and problem size, so a later run reuses the winner without timing.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from pathlib import Path

from .form import CUDAForm

DEFAULT_CANDIDATES = [
    {"chunk_values": 4096, "owner_pairs": 256, "owner_block": 256},
    {"chunk_values": 4096, "owner_pairs": 256, "owner_block": 256, "owner_wpe": 4},
    {"chunk_values": 4096, "owner_pairs": 384, "owner_block": 256},
    {"chunk_values": 4608, "owner_pairs": 384, "owner_block": 256, "owner_lds": 73728},
    {"chunk_values": 5120, "owner_pairs": 384, "owner_block": 256, "owner_lds": 81920},
    {"chunk_values": 3072, "owner_pairs": 256, "owner_block": 256},
    {"chunk_values": 6144, "owner_pairs": 512, "owner_block": 256, "owner_lds": 98304},
    # warp-specialised variant: three compute waves and one store wave,
    # double-buffered accumulators
    {"chunk_values": 2304, "owner_pairs": 192, "owner_pipe": 1, "owner_cwaves": 3,
     "owner_lds": 98304},
]

# scalar reference-tensor forms (kappa-weighted stiffness, any degree): the
# (nd x nd x 6) table shares LDS with the accumulator
TRACE_CANDIDATES = [
    {"chunk_values": 2048, "owner_pairs": 384, "owner_block": 256},
    {"chunk_values": 1536, "owner_pairs": 384, "owner_block": 256},
    {"chunk_values": 3072, "owner_pairs": 384, "owner_block": 256, "owner_lds": 98304},
    {"chunk_values": 2048, "owner_pairs": 256, "owner_block": 128},
    {"chunk_values": 2048, "owner_pairs": 384, "owner_pipe": 1, "owner_cwaves": 3,
     "owner_lds": 98304},
]

# Q_k hexahedra (matrix-core / FP64 owner kernel): value chunk, threads per
# workgroup and cells per record batch
HEX_CANDIDATES = [
    {"chunk_values": 4096},
    {"chunk_values": 4096, "hex_block": 256},
    {"chunk_values": 4096, "owner_nb": 8},
    {"chunk_values": 6144, "owner_lds": 110000},
    {"chunk_values": 3072, "owner_lds": 70000},
]


def candidates_for(form):
    """The candidate shapes for `form`'s owner-computes kernel."""
    spec = form.integrals[0].staged.source() if form.integrals and form.integrals[0].staged else ""
    if "HX_NQ1" in spec:
        return HEX_CANDIDATES
    if "SK_TRACE" in spec:
        return TRACE_CANDIDATES
    return DEFAULT_CANDIDATES


def _key(form, A) -> str:
    desc = "|".join([form.description, str(A.nrows), str(A.nnz)] +
                    [I.name for I in form.integrals])
    return hashlib.sha1(desc.encode()).hexdigest()[:16]


def _table_path() -> Path:
    base = os.environ.get("MIFEM_CACHE_DIR") or str(Path.home() / ".cache" / "mifem")
    return Path(base) / "autotune.json"


def autotune_matrix(form, asm, A, bcs, candidates=None, reps: int = 5, verbose: bool = False,
                    base_options=None):
    """Return (best options, best CUDAForm, timings[list]) for assembling `form` into A.

    candidates: shapes to time (default: by the form's kind, candidates_for);
    base_options: options every candidate carries (e.g. {"precision": 32}).
    """
    table_path = _table_path()
    key = _key(form, A) + ("" if not base_options else
                          "|" + ",".join(f"{k}={v}" for k, v in sorted(base_options.items())))
    table = {}
    if table_path.exists():
        try:
            table = json.loads(table_path.read_text())
        except (OSError, ValueError):
            table = {}
    if candidates is None and key in table:
        opts = dict(table[key]["options"])
        return opts, CUDAForm(form, options=opts), table[key].get("timings", [])
    timings = []
    best = None
    for cand in candidates or candidates_for(form):
        opts = dict(base_options or {})
        opts.update(cand)
        opts["matrix_kernel"] = 2
        try:
            a = CUDAForm(form, options=opts)
            asm.assemble_matrix(a, A, bcs=bcs)         # plan build + warm-up
            asm.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                asm.assemble_matrix(a, A, bcs=bcs)
            asm.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
        except (RuntimeError, NotImplementedError, ValueError) as e:
            if verbose:
                print(f"[autotune] {cand}: unavailable ({e})", file=sys.stderr, flush=True)
            continue
        timings.append({"options": cand, "ms": ms})
        if verbose:
            print(f"[autotune] {cand}: {ms:.3f} ms", file=sys.stderr, flush=True)
        if best is None or ms < best[0]:
            best = (ms, opts, a)
    if best is None:
        raise RuntimeError("no owner-computes candidate could run")
    table[key] = {"options": best[1], "ms": best[0], "timings": timings}
    try:
        table_path.parent.mkdir(parents=True, exist_ok=True)
        table_path.write_text(json.dumps(table, indent=1))
    except OSError:
        pass
    return best[1], best[2], timings


SYSTEM_BLOCKS = (64, 128, 256)


def autotune_system(a, L, M, asm, A, b, value, blocks=SYSTEM_BLOCKS, reps: int = 10,
                    stream=None, verbose: bool = False):
    """Cells per workgroup of the fused Newton-system kernel: times one
    `CUDASystem` per candidate block size on the actual forms and matrix (HIP
    events on `stream`, the torch stream the assembler runs on; default the
    current stream) and returns (best system, best block, [(block, ms)])."""
    import torch
    from .system import CUDASystem

    timings, best = [], None
    for blk in blocks:
        system = CUDASystem(a, L, M, block=blk)
        system.assemble(asm, A, b, value)
        asm.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s = stream if stream is not None else torch.cuda.current_stream()
        e0.record(s)
        for _ in range(reps):
            system.assemble(asm, A, b, value, pack=False)
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        timings.append((blk, ms))
        if verbose:
            print(f"[autotune] system block {blk}: {ms:.3f} ms", file=sys.stderr, flush=True)
        if best is None or ms < best[1]:
            best = (system, ms, blk)
    return best[0], best[2], timings


def autotune_matrix_native(a, asm, A, bcs=(), reps: int = 5):
    """The C-ABI tuner (mf_autotune_matrix) on an existing CUDAForm: the form
    is recompiled in place with the fastest owner-computes shape.  Returns
    (description "key=value,...", ms per assembly)."""
    import ctypes as C

    from ._lib import check, lib
    coll = asm._bc_collection(list(bcs) if isinstance(bcs, tuple) else bcs)
    bc = coll._get_cpp_bcs(a.function_spaces[0])
    a.to_device(force=False)
    buf = C.create_string_buffer(256)
    ms = C.c_double(0.0)
    check(lib().mf_autotune_matrix(a.handle, A.handle, bc.handle, int(reps), asm._stream, buf,
                                   256, C.byref(ms)))
    return buf.value.decode(), ms.value


# member shapes of a CUDABatch: value chunk (rows per workgroup) and pairs
# per chunk; the batch kernel's threads per workgroup are shared by all members
BATCH_CANDIDATES = [
    {"chunk_values": 2048, "owner_pairs": 384},
    {"chunk_values": 1536, "owner_pairs": 256},
    {"chunk_values": 3072, "owner_pairs": 384, "owner_lds": 98304},
    {"chunk_values": 4096, "owner_pairs": 384, "owner_lds": 98304},
]


def autotune_batch(forms, asm, mats, bcs, base_options=None, candidates=None, reps: int = 5,
                   sweeps: int = 1, verbose: bool = False):
    """Member shapes of a CUDABatch of bilinear forms, tuned together.

    The batch kernel's occupancy follows its largest member's LDS, so one
    member's shape changes the speed of all: the tuner walks the members one
    at a time (coordinate descent, `sweeps` passes), timing each candidate as
    the whole batch launch on the actual matrices and boundary conditions
    while the other members keep their best shape so far.  forms: the
    members' (mifem) bilinear forms; base_options[g]: options member g always
    carries (e.g. {"precision": 32}).

    Returns (options per member, the CUDABatch, its CUDAForms, timings).
    """
    from .batch import CUDABatch

    cands = list(candidates or BATCH_CANDIDATES)
    n = len(forms)
    base = [dict(o or {}) for o in (base_options or [None] * n)]
    compiled = {}

    def member(g, c):
        key = (g, c)
        if key not in compiled:
            opts = dict(base[g])
            opts.update(cands[c])
            opts["matrix_kernel"] = 2
            compiled[key] = CUDAForm(forms[g], options=opts)
        return compiled[key]

    def time_batch(choice):
        fs = [member(g, c) for g, c in enumerate(choice)]
        batch = CUDABatch(fs)
        batch.assemble(asm, mats, bcs)                # warm-up (and coefficient upload)
        asm.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            batch.assemble(asm, mats, bcs, pack=False)
        asm.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3, batch, fs

    # start from the first candidate every member can take
    for c0 in range(len(cands)):
        choice = [c0] * n
        try:
            best_ms, best_batch, best_fs = time_batch(choice)
            break
        except (RuntimeError, ValueError) as e:
            if verbose:
                print(f"[autotune] batch {choice}: unavailable ({e})", file=sys.stderr, flush=True)
    else:
        raise RuntimeError("no batch shape could run")
    timings = [{"members": list(choice), "ms": best_ms}]
    if verbose:
        print(f"[autotune] batch {choice}: {best_ms:.3f} ms", file=sys.stderr, flush=True)
    for _ in range(sweeps):
        for g in range(n):
            for c in range(len(cands)):
                if c == choice[g]:
                    continue
                trial = list(choice)
                trial[g] = c
                try:
                    ms, batch, fs = time_batch(trial)
                except (RuntimeError, ValueError) as e:
                    if verbose:
                        print(f"[autotune] batch {trial}: unavailable ({e})", file=sys.stderr, flush=True)
                    continue
                timings.append({"members": trial, "ms": ms})
                if verbose:
                    print(f"[autotune] batch {trial}: {ms:.3f} ms", file=sys.stderr, flush=True)
                if ms < best_ms:
                    best_ms, best_batch, best_fs, choice = ms, batch, fs, trial
    opts = [dict(base[g], **cands[c]) for g, c in enumerate(choice)]
    for t in timings:
        t["members"] = [cands[c] for c in t["members"]]
    return opts, best_batch, best_fs, timings
