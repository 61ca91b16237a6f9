"""a2afft: MI355X-native pencil 3D FFT (Python mirror of the reference interface).

The compute path is liba2afft.so (HIP kernels for gfx950 behind the C ABI of
include/a2afft.h); this package only binds it.  Import with
``sys.path.insert(0, "<repo>/3d-ffts-mpi_amd")``.
"""
from ._lib import (A2AFFTError, C64, C128, EXEC_FUSED, EXEC_STAGED, FP32, FP64, SPLIT_BF16, SPLIT_FP16,
                   LIB_PATH, lib)
from .core import (Batch, Dfft, Distribution, GridDfft, alltoall, dims_create, error_metrics, execute_pass, fast_copy_fft,
                   fill_random, forward_1d_fft, inverse_1d_fft, launch_remap, scale_fft, set_sync, set_tuning_cache)
from . import inputs

__all__ = [
    "A2AFFTError", "C64", "C128", "EXEC_FUSED", "EXEC_STAGED", "FP32", "FP64", "SPLIT_BF16", "SPLIT_FP16",
    "LIB_PATH", "lib", "Batch", "Dfft", "Distribution", "GridDfft", "alltoall", "dims_create", "error_metrics", "execute_pass",
    "fast_copy_fft", "fill_random",
    "forward_1d_fft", "inverse_1d_fft", "launch_remap", "scale_fft", "set_sync", "set_tuning_cache", "inputs",
]
