"""ctypes binding of libcac_hip.so (the C-ABI in include/cac_hip.h and
include/cac_io.h).

Loading fails loudly when the library is missing: there is no CPU fallback on
the product path.  Build it with ``python -m cac_code_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcac_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HEADER = os.path.join(INCLUDE, "cac_hip.h")
HEADERS = (HEADER, os.path.join(INCLUDE, "cac_io.h"))

CAC_OK, CAC_ERR_ARG, CAC_ERR_HIP, CAC_ERR_OVERFLOW, CAC_ERR_BOX, CAC_ERR_DATA, CAC_ERR_IO = range(7)
PREC_FP64, PREC_MIXED = 0, 1
ATOM_SITE = 0xFFFFFFFF  # CAC_ATOM_SITE: site field of a list entry naming an atom

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)
_u32 = C.POINTER(C.c_uint32)
_u64 = C.POINTER(C.c_uint64)
_f = C.POINTER(C.c_float)
_ctx = C.c_void_p
_i64 = C.POINTER(C.c_int64)


class DataView(C.Structure):
    """cac_data_view (include/cac_io.h)."""
    _fields_ = [("comment", C.c_char_p), ("lo", C.c_double * 3), ("hi", C.c_double * 3),
                ("ntypes", C.c_int), ("entypes", C.c_int), ("masses", _d), ("nae", _i), ("numint", _i),
                ("nintp", C.c_int64), ("intp_ijk", _i), ("intp_w", _d),
                ("natoms", C.c_int64), ("nelements", C.c_int64),
                ("atom_tag", _i64), ("atom_type", _i), ("x", _d), ("atom_image", _i),
                ("elem_tag", _i64), ("elem_type", _i), ("xc", _d), ("elem_image", _i),
                ("nodex", _d), ("nodev", _d), ("nodetag", _i64), ("node_velocities", C.c_int)]


# name -> (restype, argtypes)
SIGNATURES = {
    "cac_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cac_ctx_destroy": (C.c_int, [_ctx]),
    "cac_last_error": (C.c_char_p, [_ctx]),
    "cac_synchronize": (C.c_int, [_ctx]),
    "cac_set_box": (C.c_int, [_ctx, _d, _d, _i]),
    "cac_upload_atoms": (C.c_int, [_ctx, C.c_int, _d, _d, C.c_double]),
    "cac_upload_elements": (C.c_int, [_ctx, C.c_int, _i, _d, _d, C.c_double]),
    "cac_set_quadrature": (C.c_int, [_ctx, C.c_int, _i, _d]),
    "cac_eam_tables_from_funcfl": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_double, _d, _d, _d,
                                             _i, _d, _i, _d, _d, _d, _d]),
    "cac_set_eam": (C.c_int, [_ctx, C.c_int, C.c_double, _d, C.c_int, C.c_double, _d, _d, C.c_double]),
    "cac_set_lj": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    "cac_set_neighbor": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cac_set_coupling": (C.c_int, [_ctx, C.c_int]),
    "cac_set_list_order": (C.c_int, [_ctx, C.c_double]),
    "cac_set_ip_clusters": (C.c_int, [_ctx, C.c_int, _i, _i]),
    "cac_cluster_info": (C.c_int, [_ctx, _i, _i, _i]),
    "cac_deform": (C.c_int, [_ctx, _d]),
    "cac_set_deform_rate": (C.c_int, [_ctx, _d]),
    "cac_build_neighbors": (C.c_int, [_ctx]),
    "cac_decide": (C.c_int, [_ctx, _i]),
    "cac_force": (C.c_int, [_ctx, C.c_int, C.c_int, C.c_int]),
    "cac_nve_initial": (C.c_int, [_ctx, C.c_double, C.c_double]),
    "cac_nve_final": (C.c_int, [_ctx, C.c_double]),
    "cac_run": (C.c_int, [_ctx, C.c_int, C.c_double, C.c_double, C.c_int, _i]),
    "cac_download_atoms": (C.c_int, [_ctx, _d, _d, _d]),
    "cac_download_elements": (C.c_int, [_ctx, _d, _d, _d]),
    "cac_get_energy_virial": (C.c_int, [_ctx, _d, _d]),
    "cac_centro_atoms": (C.c_int, [_ctx, C.c_int, _d]),
    "cac_interpolate_sites": (C.c_int, [_ctx, _d]),
    "cac_get_capacities": (C.c_int, [_ctx, _i, _i]),
    "cac_thermo": (C.c_int, [_ctx, C.c_double, C.c_double, C.c_double, _d]),
    "cac_eam_atoms_noalloc": (C.c_int, [_ctx, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "cac_get_atom_neighbors": (C.c_int, [_ctx, C.c_int, _i, _u64]),
    "cac_get_ip_neighbors": (C.c_int, [_ctx, C.c_int, _i, _u64]),
    "cac_get_rho": (C.c_int, [_ctx, _d, _d]),
    "cac_get_erho": (C.c_int, [_ctx, _d]),
    "cac_get_ip_forces": (C.c_int, [_ctx, _d]),
    "cac_profile_enable": (C.c_int, [_ctx, C.c_int]),
    "cac_profile_avg_ms": (C.c_double, [_ctx, C.c_char_p, _i]),
    # learned element surrogate
    "cac_mlp_create": (C.c_int, [_ctx, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64]),
    "cac_mlp_set_weights": (C.c_int, [_ctx, C.c_int, _f, _f]),
    "cac_mlp_get_weights": (C.c_int, [_ctx, C.c_int, _f, _f]),
    "cac_mlp_forward": (C.c_int, [_ctx, C.c_int, _f, _f]),
    "cac_mlp_train_step": (C.c_int, [_ctx, C.c_int, _f, _f, C.c_float, _f]),
    "cac_mlp_grads": (C.c_int, [_ctx, C.c_int, _f, _f, _f, C.POINTER(_f)]),
    "cac_mlp_set_dataset": (C.c_int, [_ctx, C.c_int, _f, _f]),
    "cac_mlp_train": (C.c_int, [_ctx, C.c_int, C.c_int, C.c_float, C.c_uint64, _f]),
    "cac_surrogate_setup": (C.c_int, [_ctx, _d, _i, _f, _f, C.c_double]),
    "cac_surrogate_set_fit": (C.c_int, [_ctx, _f, _f, _f]),
    "cac_surrogate_enable": (C.c_int, [_ctx, C.c_int]),
    "cac_surrogate_features": (C.c_int, [_ctx, _f, _i]),
    "cac_surrogate_info": (C.c_int, [_ctx, _i, _i, _i]),
    # on-disk formats (include/cac_io.h)
    "cac_io_last_error": (C.c_char_p, []),
    "cac_data_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "cac_data_get": (C.POINTER(DataView), [C.c_void_p]),
    "cac_data_free": (None, [C.c_void_p]),
    "cac_data_write": (C.c_int, [C.c_char_p, C.POINTER(DataView), C.c_int]),
    "cac_dump_atom_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, _d, _d, _i, C.c_int64, _i64, _i, _d, _d, _d,
                                      C.c_int64, _i64, _d, _d, _d]),
    "cac_dump_element_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int64, C.c_int64, _d, _d, _d]),
    "cac_eam_tables_from_setfl": (C.c_int, [C.c_int, C.c_double, C.c_int, C.c_double, _d, _d, _d, _d, _d, _d]),
}


class CacError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"CAC status {code}: {msg}")
        self.code = code


def header_symbols():
    """Function names declared in include/*.h."""
    names = set()
    for h in HEADERS:
        text = open(h).read()
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(cac_\w+)\s*\(", text, re.M))
    return sorted(names)


_lib = None


def lib():
    """Load libcac_hip.so; raise if it is not built (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m cac_code_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib
