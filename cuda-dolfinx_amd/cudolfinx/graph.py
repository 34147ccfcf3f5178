"""HIP-graph capture of repeated assembly sequences.

The reference launches every kernel and synchronises after each one
(cpp/cudolfinx/fem/CUDAFormIntegral.cpp:2354-2361); a Newton or
time-stepping loop here instead records its assembly calls once on a
non-default stream and replays them with a single graph launch:

    asm = CUDAAssembler(stream=s)
    asm.assemble_matrix(J, A); asm.assemble_vector(F, b)   # first call builds schedules
    with CUDAGraph(s) as g:
        asm.pack_coefficients(J); asm.assemble_matrix(J, A)
        asm.pack_coefficients(F); asm.assemble_vector(F, b)
    g.launch()                                             # every Newton step

Coefficients should be bound to device memory
(`CUDAForm.set_coefficient_device`) so each replay reads the current state.
"""
from __future__ import annotations

import ctypes as C

from ._lib import check, lib
from .assemble import _stream
from .context import get_cuda_context


class CUDAGraph:
    def __init__(self, stream):
        if stream is None:
            raise ValueError("graph capture needs a non-default stream")
        self._ctx = get_cuda_context()
        self._stream = _stream(stream)
        self._handle = None

    def __enter__(self):
        check(lib().mf_graph_begin(self._ctx.handle, self._stream))
        return self

    def __exit__(self, exc_type, exc, tb):
        h = C.c_void_p()
        rc = lib().mf_graph_end(self._ctx.handle, self._stream, C.byref(h))
        if exc_type is None:
            check(rc)
            self._handle = h
        return False

    def launch(self, stream=None) -> None:
        if self._handle is None:
            raise RuntimeError("graph was not captured")
        s = self._stream if stream is None else _stream(stream)
        check(lib().mf_graph_launch(self._handle, s))

    def __del__(self):
        if getattr(self, "_handle", None):
            try:
                lib().mf_graph_destroy(self._handle)
            except Exception:
                pass
            self._handle = None
