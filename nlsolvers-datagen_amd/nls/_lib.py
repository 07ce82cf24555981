"""ctypes binding of libnls.so (the C ABI declared in include/nls.h).

The library is loaded from this directory.  There is no fallback: if libnls.so
is missing or cannot be loaded, every call raises NLSError, so nothing silently
runs on the CPU.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnls.so")


class NLSError(RuntimeError):
    pass


class KGEParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("num_snapshots", C.c_int), ("krylov_dim", C.c_int),
                ("method", C.c_int), ("mu2", C.c_double), ("q6", C.c_int),
                ("variant", C.c_int)]


class NLSEParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("num_snapshots", C.c_int), ("krylov_dim", C.c_int),
                ("method", C.c_int), ("q6", C.c_int), ("variant", C.c_int)]


class SP4Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int), ("L", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("num_snapshots", C.c_int), ("mode", C.c_int), ("q6", C.c_int),
                ("gamma", C.c_double), ("sigma", C.c_double), ("seed", C.c_uint64),
                ("stream_id", C.c_uint32)]


class BatchParams(C.Structure):
    _fields_ = [("equation", C.c_int), ("method", C.c_int), ("dim", C.c_int), ("dt", C.c_double),
                ("krylov_dim", C.c_int), ("mu2", C.c_double), ("q6", C.c_int),
                ("gamma", C.c_double), ("use_graph", C.c_int), ("snapshot_every", C.c_int),
                ("max_snapshots", C.c_int), ("cells_per_thread", C.c_int),
                ("pass_kernel", C.c_int), ("lanes", C.c_int), ("reorth", C.c_int),
                ("reorth_tol", C.c_double), ("always_load_fields", C.c_int),
                ("variant", C.c_int), ("persistent", C.c_int)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int), ("L", C.c_double),
                ("mode", C.c_int), ("gamma", C.c_double), ("max_steps_per_group", C.c_int),
                ("h5_dir", C.c_char_p), ("run_id", C.c_char_p)]


class H5Run(C.Structure):
    _fields_ = [("path", C.c_char_p), ("complex_field", C.c_int), ("dim", C.c_int),
                ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("T", C.c_double),
                ("nt", C.c_int), ("num_snapshots", C.c_int),
                ("problem_type", C.c_char_p), ("run_id", C.c_char_p), ("phenomenon", C.c_char_p),
                ("m_type", C.c_char_p), ("timestamp", C.c_char_p), ("run_index", C.c_int),
                ("elapsed_time", C.c_double), ("n_params", C.c_int),
                ("param_keys", C.POINTER(C.c_char_p)), ("param_values", C.POINTER(C.c_char_p)),
                ("u0", C.POINTER(C.c_double)), ("v0", C.POINTER(C.c_double)),
                ("m", C.POINTER(C.c_double)), ("c", C.POINTER(C.c_double)),
                ("ns_out", C.c_int), ("ox", C.c_int), ("oy", C.c_int), ("oz", C.c_int),
                ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
                ("c_type", C.c_char_p)]


class Request(C.Structure):
    _fields_ = [("n", C.c_int), ("nt", C.c_int), ("dt", C.c_double), ("sigma", C.c_double),
                ("seed", C.c_uint64), ("num_snapshots", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("failed", C.c_int64), ("active", C.c_int64),
                ("queued", C.c_int64), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("cell_steps", C.c_double), ("busy_s", C.c_double),
                ("peak_bytes", C.c_size_t), ("peak_active", C.c_int)]


class TrajDesc(C.Structure):
    _fields_ = [("n", C.c_int), ("L", C.c_double), ("sigma", C.c_double), ("seed", C.c_uint64)]


# constants mirrored from nls.h
F_COS_SQRT, F_SINC2_SQRT, F_ID_SQRT, F_SINC2_HALF, F_COSM1_SQRT = 0, 1, 2, 3, 4
F_EXP, F_SINC = 8, 9
KGE_GAUTSCHI, KGE_SV, KGE_SEWI = 0, 1, 2
NLSE_STRANG, NLSE_SEWI = 0, 1
SP4_REFERENCE, SP4_DAMPED = 0, 1
EQ_KGE, EQ_NLSE, EQ_SP4 = 0, 1, 2

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)

# name -> argtypes (all functions return int status unless listed in _RESTYPE)
_SIGS = {
    "nls_last_error": [],
    "nls_version": [],
    "nls_device_info": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_size_t)],
    "nls_dev_malloc": [C.POINTER(C.c_void_p), C.c_size_t],
    "nls_dev_free": [_P],
    "nls_memcpy_h2d": [_P, _P, C.c_size_t],
    "nls_memcpy_d2h": [_P, _P, C.c_size_t],
    "nls_dev_synchronize": [],
    "nls_laplacian_nnz": [C.c_int, C.c_int, C.c_int],
    "nls_build_laplacian_2d": [C.c_int, C.c_double, C.c_double, _D, C.c_int, C.c_double, _I32,
                               _I32, _D, _I64],
    "nls_build_laplacian_3d": [C.c_int, C.c_double, _D, C.c_int, C.c_double, C.c_int, _I32, _I32,
                               _D, _I64],
    "nls_csr_spmv_f64": [C.c_int64, _P, _P, _P, _P, _P, _P],
    "nls_csr_spmv_f32": [C.c_int64, _P, _P, _P, _P, _P, _P],
    "nls_csr_spmv_c128": [C.c_int64, _P, _P, _P, _P, _P, _P],
    "nls_csr_spmv_c64": [C.c_int64, _P, _P, _P, _P, _P, _P],
    "nls_matfunc_create": [C.c_int64, C.c_int64, _I32, _I32, _D, C.c_int, C.c_int,
                           C.POINTER(C.c_void_p)],
    "nls_matfunc_destroy": [_P],
    "nls_matfunc_apply_real": [_P, _P, _P, C.c_double, C.c_int],
    "nls_matfunc_apply_complex": [_P, _P, _P, C.c_double, C.c_double, C.c_int],
    "nls_nc_class_create": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                            C.c_double, C.c_double, C.c_double, C.POINTER(C.c_void_p)],
    "nls_nc_class_put": [_P, C.c_int, _F, _F, _F, _F, _F, _F],
    "nls_nc_class_close": [_P],
    "nls_nc_std_create": [C.c_char_p, C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_int,
                          C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, C.c_char_p,
                          C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)],
    "nls_nc_std_put": [_P, C.c_int, _F, _F, _F, _F],
    "nls_nc_std_close": [_P],
    "nls_h5_read_range_f32": [C.c_char_p, C.c_char_p, C.c_int64, C.c_int64, _F],
    "nls_matfunc_apply_real_noalloc": [C.c_int64, _P, _P, _P, _P, _P, C.c_double, C.c_int,
                                       C.c_int, _P],
    "nls_matfunc_last_krylov": [_P, _D, _D],
    "nls_small_matfunc": [C.c_int, _D, _D, C.c_int, C.c_double, C.c_double, C.c_double, _D, _D],
    "nls_kge_create": [C.POINTER(KGEParams), _D, _D, _D, _D, C.POINTER(C.c_void_p)],
    "nls_kge_create_csr": [C.POINTER(KGEParams), C.c_int64, _I32, _I32, _D, _D, _D, _D,
                           C.POINTER(C.c_void_p)],
    "nls_kge_destroy": [_P],
    "nls_kge_step": [_P],
    "nls_kge_step_sv": [_P],
    "nls_kge_apply_bc": [_P],
    "nls_kge_store_snapshot": [_P, C.c_int],
    "nls_kge_transfer_snapshots": [_P, _D, _D],
    "nls_kge_run": [_P],
    "nls_kge_get_state": [_P, _D, _D],
    "nls_nlse_create": [C.POINTER(NLSEParams), _D, _D, _D, C.POINTER(C.c_void_p)],
    "nls_nlse_destroy": [_P],
    "nls_nlse_step": [_P, C.c_int],
    "nls_nlse_step_sewi": [_P, C.c_int],
    "nls_nlse_apply_bc": [_P],
    "nls_nlse_run": [_P],
    "nls_nlse_transfer_snapshots": [_P, _D],
    "nls_nlse_get_state": [_P, _D],
    "nls_nlse_store_snapshot_online": [_P, _D],
    "nls_sp4_create": [C.POINTER(SP4Params), _D, _D, _D, _D, C.POINTER(C.c_void_p)],
    "nls_sp4_destroy": [_P],
    "nls_sp4_step": [_P],
    "nls_sp4_apply_bc": [_P],
    "nls_sp4_store_snapshot": [_P, C.c_int],
    "nls_sp4_transfer_snapshots": [_P, _D, _D],
    "nls_sp4_run": [_P],
    "nls_philox_normals": [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int64, _D],
    "nls_stream_copy_bench": [C.c_size_t, C.c_int, _D, _D],
    "nls_batch_create": [C.POINTER(BatchParams), C.c_int, C.POINTER(TrajDesc),
                         C.POINTER(C.c_void_p)],
    "nls_batch_destroy": [_P],
    "nls_batch_set_initial": [_P, C.c_int, _D, _D, _D, _D],
    "nls_batch_run": [_P, C.c_int, C.POINTER(C.c_float)],
    "nls_batch_get_snapshots": [_P, C.c_int, _D, _D, C.POINTER(C.c_int)],
    "nls_batch_get_state": [_P, C.c_int, _D, _D],
    "nls_batch_info": [_P, C.POINTER(C.c_int), _I64, C.POINTER(C.c_size_t)],
    "nls_batch_set_profiling": [_P, C.c_int],
    "nls_batch_persistent_steps": [_P, _I64],
    "nls_batch_abort": [_P],
    "nls_batch_monitor": [_P, C.c_int, C.POINTER(C.c_int), _D, _D, C.POINTER(C.c_int)],
    "nls_ic_draw": [C.c_int, C.c_uint64, C.c_int, C.c_double, _D, C.c_int, C.POINTER(C.c_int)],
    "nls_field_sample": [C.c_char_p, C.c_char_p, C.c_int, C.c_double, C.c_uint64, C.c_int,
                         C.c_double, _P, _P, _P],
    "nls_phenomenon_sample": [C.c_char_p, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                              _P, _P, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), _P],
    "nls_batch_set_initial_sampled": [_P, C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_double,
                                      C.c_double],
    "nls_batch_tile": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "nls_batch_pass_timing": [_P, _D, C.POINTER(C.c_int), _D],
    "nls_batch_estimate_bytes": [C.POINTER(BatchParams), C.c_int, C.POINTER(TrajDesc),
                                 C.POINTER(C.c_size_t)],
    "nls_server_create": [C.POINTER(ServerParams), C.POINTER(C.c_void_p)],
    "nls_server_destroy": [_P],
    "nls_server_submit": [_P, C.POINTER(Request), _D, _D, _D, _D, _I64],
    "nls_server_status": [_P, C.c_int64, C.POINTER(C.c_int)],
    "nls_server_wait": [_P, C.c_int64, C.c_double],
    "nls_server_drain": [_P, C.c_double],
    "nls_server_fetch": [_P, C.c_int64, _D, C.POINTER(C.c_int), _D, _D],
    "nls_server_release": [_P, C.c_int64],
    "nls_server_get_stats": [_P, C.POINTER(ServerStats)],
    "nls_pool_simulate": [C.c_size_t, C.c_int, _I64, _I64],
    "nls_h5_available": [C.c_char_p, C.c_int],
    "nls_h5_write_run": [C.POINTER(H5Run)],
    "nls_h5_writer_create": [C.c_size_t, C.POINTER(C.c_void_p)],
    "nls_h5_writer_submit": [_P, C.POINTER(H5Run)],
    "nls_h5_writer_drain": [_P, _I64],
    "nls_h5_writer_destroy": [_P],
    "nls_h5_read_dataset": [C.c_char_p, C.c_char_p, _D, C.c_int64, _I64, C.POINTER(C.c_int),
                            C.POINTER(C.c_int)],
    "nls_h5_read_attr": [C.c_char_p, C.c_char_p, C.c_char_p, _D, C.c_char_p, C.c_int,
                         C.POINTER(C.c_int)],
}
_RESTYPE = {"nls_last_error": C.c_char_p, "nls_laplacian_nnz": C.c_int64}

_lib = None


def exported_symbols():
    """Names the binding expects libnls.so to export."""
    return sorted(_SIGS)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NLSError(f"libnls.so not found at {LIB_PATH}; run __graft_entry__.build() "
                       "(make -C nlsolvers-datagen_amd)")
    lib = C.CDLL(LIB_PATH)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int)
    _lib = lib
    return lib


def call(name, *args):
    """Calls a status-returning entry point and raises NLSError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise NLSError(f"{name}: {lib.nls_last_error().decode(errors='replace')}")
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def fptr(a):
    return a.ctypes.data_as(_F)


def i32ptr(a):
    return a.ctypes.data_as(_I32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_c128_view(a):
    """complex128 array -> interleaved float64 view (no copy when contiguous)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)
