"""ctypes bindings of liba2afft.so (the C ABI declared in include/a2afft.h).

The shared library is built in-tree by ``make -C 3d-ffts-mpi_amd/csrc`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "liba2afft.so"))

OK = 0
W_DIVERGES_FROM_REFERENCE = 1
E_INVALID_ARG = -1
E_DIVISIBILITY = -2
E_RANKS = -3
E_CHUNK = -4
E_HIP = -5
E_ALLOC = -6
E_UNSUPPORTED = -7
E_NO_SCRATCH = -8
E_SIZE = -9

C128, C64 = 0, 1
FP64, FP32, SPLIT_BF16, SPLIT_FP16 = 0, 1, 2, 3
EXEC_FUSED, EXEC_STAGED = 0, 1


class DistInfo(ctypes.Structure):
    _fields_ = [("Ng", ctypes.c_int), ("nlocal", ctypes.c_int64), ("world_size", ctypes.c_int),
                ("local_grid_size", ctypes.c_int * 3), ("dims", ctypes.c_int * 3),
                ("blockSize", ctypes.c_int), ("dtype", ctypes.c_int),
                ("reference_domain", ctypes.c_int)]


class A2AFFTError(RuntimeError):
    def __init__(self, status, where, detail):
        super().__init__(f"{where}: status {status} ({detail})")
        self.status = status


_lib = None
_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_IA = ctypes.POINTER(ctypes.c_int)

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "a2afft_status_string": [_I],
    "a2afft_last_error": [],
    "a2afft_set_sync": [_I],
    "a2afft_set_tuning_cache": [ctypes.c_char_p],
    "a2afft_version": [],
    "a2afft_dims_create": [_I, _IA],
    "a2afft_dist_create": [ctypes.POINTER(_P), _I, _I, _IA, _I, _I],
    "a2afft_dist_get_info": [_P, ctypes.POINTER(DistInfo)],
    "a2afft_dist_rank_coords": [_P, _I, _IA, _IA],
    "a2afft_dist_write_dims": [_P, ctypes.c_char_p],
    "a2afft_dist_destroy": [_P],
    "a2afft_get_z_pencils": [_P, _P, _P, _P],
    "a2afft_return_z_pencils": [_P, _P, _P, _P],
    "a2afft_get_x_pencils": [_P, _P, _P, _P],
    "a2afft_return_x_pencils": [_P, _P, _P, _P],
    "a2afft_get_y_pencils": [_P, _P, _P, _P],
    "a2afft_return_y_pencils": [_P, _P, _P, _P],
    "a2afft_launch_d_z_a2a_to_z_pencils": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_z_pencils_to_z_a2a": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_x_a2a_to_x_pencils": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_x_pencils_to_x_a2a": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_y_a2a_to_y_pencils": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_y_pencils_to_y_a2a": [_P, _P, _I, _I, _I64, _IA, _IA, _I, _I, _P],
    "a2afft_launch_d_fast_z_to_x": [_P, _P, _IA, _I, _I64, _I, _I, _P],
    "a2afft_launch_d_fast_x_to_y": [_P, _P, _IA, _I, _I64, _I, _I, _P],
    "a2afft_launch_d_fast_y_to_z": [_P, _P, _IA, _I, _I64, _I, _I, _P],
    "a2afft_alltoall": [_P, _P, _I, _I64, _I, _P],
    "a2afft_forward_1d_fft": [_P, _I, _I64, _I, _P],
    "a2afft_inverse_1d_fft": [_P, _I, _I64, _I, _P],
    "a2afft_launch_scale_fft": [_P, _P, _I, _I64, _I, _I, _P],
    "a2afft_launch_fast_copy_fft": [_P, _P, _I64, _I, _I, _P],
    "a2afft_plan_create": [ctypes.POINTER(_P), _P, _I, _I],
    "a2afft_grid_plan_create": [ctypes.POINTER(_P), _I, _I, _I, _I],
    "a2afft_make_plans": [_P, _P],
    "a2afft_forward": [_P, _P, _P],
    "a2afft_inverse": [_P, _P, _P],
    "a2afft_graph_launch": [_P, _P, _I, _P],
    "a2afft_plan_destroy": [_P],
    "a2afft_execute_pass": [_P, _P, _I, _I, _P],
    "a2afft_plan_autotune": [_P, _P, _I, _P],
    "a2afft_plan_set_kernel": [_P, _I, _I, _I, _I],
    "a2afft_plan_describe": [_P, ctypes.c_char_p, _I],
    "a2afft_plan_pattern_copy": [_P, _P, _I, _I, ctypes.POINTER(ctypes.c_float), _P],
    "a2afft_plan_set_chunking": [_P, _I],
    "a2afft_plan_num_launches": [_P],
    "a2afft_execute_launch": [_P, _P, _I, _I, _P],
    "a2afft_fill_random": [_P, _I, _I64, ctypes.c_uint64, _I, _I64, _P],
    "a2afft_batch_create": [ctypes.POINTER(_P), _I, _IA, _I],
    "a2afft_batch_create_ranks": [ctypes.POINTER(_P), _I, _IA, _IA, _I],
    "a2afft_batch_offsets": [_P, ctypes.POINTER(_I64)],
    "a2afft_batch_execute": [_P, _P, _I, _P],
    "a2afft_batch_graph_launch": [_P, _P, _I, _P],
    "a2afft_batch_status": [_P, _P],
    "a2afft_batch_set_schedule": [_P, _I, _I, _I],
    "a2afft_batch_set_grouping": [_P, _I64, _I],
    "a2afft_batch_autotune": [_P, _P, _I, _P],
    "a2afft_batch_set_shape": [_P, _I, _I, _I, _I, _I],
    "a2afft_batch_describe": [_P, ctypes.c_char_p, _I],
    "a2afft_batch_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I), ctypes.POINTER(_I)],
    "a2afft_batch_destroy": [_P],
    "a2afft_error_metrics": [_P, _I, _P, _I64, ctypes.POINTER(ctypes.c_double),
                             ctypes.POINTER(ctypes.c_double), _P],
    "a2afft_error_metrics_noalloc": [_P, _I, _P, _I64, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), _P],
}
_RESTYPES = {
    "a2afft_status_string": ctypes.c_char_p,
    "a2afft_last_error": ctypes.c_char_p,
    "a2afft_set_sync": None,
}


def lib():
    """The loaded liba2afft.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"liba2afft.so not built at {LIB_PATH}: run `make -C 3d-ffts-mpi_amd/csrc` "
                               "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(status, where):
    """Raises A2AFFTError for negative status codes; returns warnings (>0)."""
    if status < 0:
        L = lib()
        detail = (L.a2afft_last_error() or b"").decode() or L.a2afft_status_string(status).decode()
        raise A2AFFTError(status, where, detail)
    return status


def ints(vals):
    return (ctypes.c_int * len(vals))(*vals)
