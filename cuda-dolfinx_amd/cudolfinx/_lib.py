"""ctypes binding of libmifem.so (the engine's C ABI, include/mifem.h).

The library is loaded from the package tree (``cuda-dolfinx_amd/lib``); if it
is missing the import fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = Path(os.environ.get("MIFEM_LIBRARY", _ROOT / "lib" / "libmifem.so"))

MF_OK, MF_EINVAL, MF_ESHAPE, MF_EHIP, MF_ECOMPILE, MF_ENOMEM, MF_ESTATE, MF_ENOTFOUND, \
    MF_EUNSUPPORTED = range(9)
MF_CELL, MF_EXTERIOR_FACET, MF_INTERIOR_FACET = 0, 1, 2

# every symbol include/mifem.h declares, with (restype, argtypes)
_vp, _i32p, _dp, _i64, _i32 = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int64, C.c_int
SIGNATURES = {
    "mf_last_error": (C.c_char_p, []),
    "mf_version": (C.c_int, []),
    "mf_context_create": (C.c_int, [C.c_int, C.c_char_p, C.POINTER(_vp)]),
    "mf_context_destroy": (C.c_int, [_vp]),
    "mf_context_arch": (C.c_int, [_vp, C.c_char_p, C.c_int]),
    "mf_alloc": (C.c_int, [_vp, _i64, C.POINTER(_vp)]),
    "mf_free": (C.c_int, [_vp, _vp]),
    "mf_memcpy_htod": (C.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "mf_memcpy_dtoh": (C.c_int, [_vp, _vp, _vp, _i64, _vp]),
    "mf_synchronize": (C.c_int, [_vp, _vp]),
    "mf_mesh_create": (C.c_int, [_vp, _dp, _i64, _i32p, _i64, C.c_int, C.POINTER(C.c_uint8),
                                 C.c_int, C.POINTER(_vp)]),
    "mf_mesh_destroy": (C.c_int, [_vp]),
    "mf_dofmap_create": (C.c_int, [_vp, _i32p, _i64, C.c_int, C.c_int, _i64, C.POINTER(_vp)]),
    "mf_dofmap_destroy": (C.c_int, [_vp]),
    "mf_form_create": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.POINTER(_vp)]),
    "mf_form_add_integral": (C.c_int, [_vp, C.c_int, C.c_int, C.c_char_p, C.c_char_p, _i32p, _i64,
                                       C.c_char_p]),
    "mf_form_add_coefficient": (C.c_int, [_vp, _vp, _i64, C.POINTER(C.c_int)]),
    "mf_form_set_coefficient": (C.c_int, [_vp, C.c_int, _vp, _i64, C.c_int]),
    "mf_form_set_constants": (C.c_int, [_vp, _dp, C.c_int]),
    "mf_form_compile": (C.c_int, [_vp, C.c_int]),
    "mf_form_pack_coefficients": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int, _vp]),
    "mf_form_set_option": (C.c_int, [_vp, C.c_char_p, _i64]),
    "mf_form_source": (C.c_int, [_vp, C.c_char_p, C.POINTER(_i64)]),
    "mf_form_destroy": (C.c_int, [_vp]),
    "mf_matrix_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mf_matrix_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "mf_matrix_device_arrays": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "mf_matrix_borrow_values": (C.c_int, [_vp, _vp]),
    "mf_spmv": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_matrix_diagonal": (C.c_int, [_vp, _vp, _vp]),
    "mf_matrix_copy_csr": (C.c_int, [_vp, _i32p, _i32p, _dp]),
    "mf_matrix_destroy": (C.c_int, [_vp]),
    "mf_bc_create": (C.c_int, [_vp, _i64, _i32p, _i64, _i64, _dp, C.POINTER(_vp)]),
    "mf_bc_update": (C.c_int, [_vp, _dp]),
    "mf_bc_destroy": (C.c_int, [_vp]),
    "mf_zero": (C.c_int, [_vp, _vp, _i64, _vp]),
    "mf_assemble_matrix": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.c_int, _vp]),
    "mf_assemble_vector": (C.c_int, [_vp, _vp, _i64, C.c_int, _vp]),
    "mf_assemble_scalar": (C.c_int, [_vp, _vp, _vp]),
    "mf_form_set_restriction": (C.c_int, [_vp, C.c_int, _vp, _i64, _i64]),
    "mf_matrix_create_block": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
    "mf_matrix_create_from_csr": (C.c_int, [_vp, _i32p, _i32p, _i64, _i64, _vp, C.POINTER(_vp)]),
    "mf_autotune_matrix": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, _i64, _vp]),
    "mf_batch_create": (C.c_int, [_vp, C.c_int, _vp]),
    "mf_batch_assemble": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int, _vp]),
    "mf_batch_assemble_vector": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_i64), C.c_int, _vp]),
    "mf_set_bc_batch": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_vp),
                                  C.c_double, _vp]),
    "mf_batch_source": (C.c_int, [_vp, _vp, _vp]),
    "mf_batch_destroy": (C.c_int, [_vp]),
    "mf_system_create": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp]),
    "mf_system_assemble": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, _vp, _i64, _vp, C.c_int, _vp]),
    "mf_system_source": (C.c_int, [_vp, _vp, _vp]),
    "mf_system_destroy": (C.c_int, [_vp]),
    "mf_interpolator_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mf_interpolate": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mf_interpolator_destroy": (C.c_int, [_vp]),
    "mf_assemble_scalar_device": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "mf_apply_lifting": (C.c_int, [C.POINTER(_vp), C.c_int, _vp, _i64, C.POINTER(_vp),
                                   C.POINTER(_vp), C.c_double, _vp]),
    "mf_apply_lifting_noalloc": (C.c_int, [C.POINTER(_vp), C.c_int, _vp, _i64, C.POINTER(_vp),
                                           C.POINTER(_vp), C.c_double, _vp]),
    "mf_set_bc": (C.c_int, [_vp, _vp, _i64, _vp, _vp, C.c_double, _vp]),
    "mf_graph_begin": (C.c_int, [_vp, _vp]),
    "mf_graph_end": (C.c_int, [_vp, _vp, _vp]),
    "mf_graph_launch": (C.c_int, [_vp, _vp]),
    "mf_graph_destroy": (C.c_int, [_vp]),
}

_lib = None


class MifemError(RuntimeError):
    pass


def lib():
    """Load libmifem.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"libmifem.so not found at {LIB_PATH}; run __graft_entry__.build() "
                              f"or `make -C cuda-dolfinx_amd`")
        try:
            # torch's ROCm wheel ships its own HIP runtime; loading torch first
            # makes libmifem bind to that same runtime instead of a second copy
            # (two runtimes in one process cannot share the device)
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc == MF_OK:
        return
    msg = lib().mf_last_error().decode(errors="replace")
    if rc in (MF_EINVAL, MF_ESHAPE):
        raise ValueError(msg)
    if rc == MF_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise MifemError(f"[mifem error {rc}] {msg}")


def ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))
