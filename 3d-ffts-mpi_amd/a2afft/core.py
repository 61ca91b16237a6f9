"""Python mirror of the reference's host interface over the C ABI.

``Distribution`` mirrors ``a2aDistribution`` (src/a2afft/distribution/
a2adistribution.h:5-42) and ``Dfft`` mirrors ``a2adfft`` (src/a2afft/dfft/
a2adfft.h:9-30): same method names, same buffer roles, same stage order.
The reference's P MPI ranks are P virtual ranks on one GPU, so every buffer is
a device tensor of shape (P, nlocal) complex; torch supplies the device memory
and the stream, the arithmetic runs in liba2afft.so.
"""
import ctypes

import torch

from . import _lib as L

DTYPES = {"c128": L.C128, "c64": L.C64}
TORCH_DTYPES = {L.C128: torch.complex128, L.C64: torch.complex64}
PRECISIONS = {"fp64": L.FP64, "fp32": L.FP32, "split_bf16": L.SPLIT_BF16, "split_fp16": L.SPLIT_FP16}
EXECS = {"fused": L.EXEC_FUSED, "staged": L.EXEC_STAGED}


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def dims_create(nnodes):
    """MPI_Dims_create(nnodes, 3) (a2adistribution.cpp:59)."""
    d = (ctypes.c_int * 3)()
    L.check(L.lib().a2afft_dims_create(nnodes, d), "dims_create")
    return tuple(d)


def set_tuning_cache(path):
    """Tuning cache file for every plan of the process (None: no cache)."""
    return L.check(L.lib().a2afft_set_tuning_cache(path.encode() if path else None), "set_tuning_cache")


def set_sync(on):
    """Block after every call, as the reference's launchers do."""
    L.lib().a2afft_set_sync(1 if on else 0)


class Distribution:
    """a2aDistribution over ``world_size`` virtual ranks (a2adistribution.cpp:44-113)."""

    def __init__(self, Ng, world_size, dims=None, blockSize=64, dtype="c128"):
        self._h = ctypes.c_void_p()
        self.dtype_code = DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        dims_arg = L.ints(dims) if dims is not None else None
        L.check(L.lib().a2afft_dist_create(ctypes.byref(self._h), Ng, world_size, dims_arg, blockSize,
                                           self.dtype_code), "a2aDistribution")
        info = L.DistInfo()
        L.check(L.lib().a2afft_dist_get_info(self._h, ctypes.byref(info)), "dist_get_info")
        self.Ng = info.Ng
        self.nlocal = info.nlocal
        self.world_size = info.world_size
        self.local_grid_size = tuple(info.local_grid_size)
        self.dims = tuple(info.dims)
        self.blockSize = info.blockSize
        self.reference_domain = bool(info.reference_domain)
        self.torch_dtype = TORCH_DTYPES[self.dtype_code]

    @property
    def handle(self):
        return self._h

    def coords(self, rank):
        """rank2coords + get_local_grid_start (a2adistribution.cpp:14-42)."""
        c = (ctypes.c_int * 3)()
        s = (ctypes.c_int * 3)()
        L.check(L.lib().a2afft_dist_rank_coords(self._h, rank, c, s), "rank_coords")
        return tuple(c), tuple(s)

    def write_dims(self, path):
        """The saveDims header (a2adistribution.cpp:101-111)."""
        L.check(L.lib().a2afft_dist_write_dims(self._h, str(path).encode()), "write_dims")

    def empty(self, device="cuda"):
        return torch.empty((self.world_size, self.nlocal), dtype=self.torch_dtype, device=device)

    def _stage(self, name, b1, b2):
        self._check_buf(b1)
        self._check_buf(b2)
        return L.check(getattr(L.lib(), name)(self._h, _ptr(b1), _ptr(b2), _stream()), name)

    def _check_buf(self, b):
        if b.dtype != self.torch_dtype or b.numel() != self.world_size * self.nlocal or not b.is_contiguous():
            raise ValueError(f"buffer must be a contiguous {self.torch_dtype} tensor of "
                             f"{self.world_size}x{self.nlocal} elements")
        if not b.is_cuda:
            raise ValueError("buffer must be a device tensor")

    def getZPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_get_z_pencils", d_Buff1, d_Buff2)

    def returnZPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_return_z_pencils", d_Buff1, d_Buff2)

    def getXPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_get_x_pencils", d_Buff1, d_Buff2)

    def returnXPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_return_x_pencils", d_Buff1, d_Buff2)

    def getYPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_get_y_pencils", d_Buff1, d_Buff2)

    def returnYPencils(self, d_Buff1, d_Buff2):
        return self._stage("a2afft_return_y_pencils", d_Buff1, d_Buff2)

    def finalize(self):
        """a2aDistribution::finalize (a2adistribution.cpp:217-223)."""
        if self._h:
            L.lib().a2afft_dist_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.finalize()
        except Exception:
            pass


class _Plan:
    """Execution methods shared by distribution plans and single-grid plans."""

    _h = ctypes.c_void_p()

    def _check_buf(self, b):
        raise NotImplementedError

    def forward(self, d_data):
        self._check_buf(d_data)
        return L.check(L.lib().a2afft_forward(self._h, _ptr(d_data), _stream()), "forward")

    def inverse(self, d_data):
        self._check_buf(d_data)
        return L.check(L.lib().a2afft_inverse(self._h, _ptr(d_data), _stream()), "inverse")

    def graph_launch(self, d_data, direction=-1):
        """forward (-1) / inverse (+1) replayed from a hipGraph captured on first use for this buffer."""
        self._check_buf(d_data)
        return L.check(L.lib().a2afft_graph_launch(self._h, _ptr(d_data), direction, _stream()), "graph_launch")

    def autotune(self, work, iters=5):
        """Pick the fastest kernel variant per pass; `work` contents are destroyed."""
        self._check_buf(work)
        return L.check(L.lib().a2afft_plan_autotune(self._h, _ptr(work), iters, _stream()), "autotune")

    def set_kernel(self, pass_index, T, E, flags=0):
        """Force a kernel variant: flags 1 XCD order, 2 LDS pad, 4 non-temporal, 8 z-tiles outermost,
        32 half-width LDS exchanges; 128 the runtime-schedule Stockham kernel (T = tile lines, E its
        elements per thread), (1, N, 2048) the chirp-z pass, (1, N, 64) the direct DFT.  Split plans:
        T=16, E=0 selects the interleaved-K split-BF16 kernel (flags 1, 4, 8 as above), E=N the
        split-K-halves kernels.  Raises A2AFFTError for a variant not built or not fitting the pass."""
        return L.check(L.lib().a2afft_plan_set_kernel(self._h, pass_index, T, E, flags), "set_kernel")

    def set_chunking(self, z_columns):
        """x/y passes alternate over windows of z_columns z columns (0 = whole passes)."""
        return L.check(L.lib().a2afft_plan_set_chunking(self._h, z_columns), "set_chunking")

    def num_launches(self):
        return L.check(L.lib().a2afft_plan_num_launches(self._h), "num_launches")

    def execute_launch(self, data, index, direction=-1):
        """Launch `index` of a fused transform; returns its pass (0 z, 1 x, 2 y)."""
        self._check_buf(data)
        return L.check(L.lib().a2afft_execute_launch(self._h, _ptr(data), index, direction, _stream()),
                       "execute_launch")

    def describe(self):
        buf = ctypes.create_string_buffer(32768)
        L.check(L.lib().a2afft_plan_describe(self._h, buf, len(buf)), "describe")
        return buf.value.decode()

    def pattern_copy(self, d_data, pass_index, iters=10):
        """Mean ms of a copy kernel with fused pass `pass_index`'s (0 z, 1 x,
        2 y) access pattern, tile order, cache policy and LDS footprint; the
        data is left unchanged."""
        self._check_buf(d_data)
        ms = ctypes.c_float()
        L.check(L.lib().a2afft_plan_pattern_copy(self._h, _ptr(d_data), pass_index, iters, ctypes.byref(ms),
                                                 _stream()), "pattern_copy")
        return ms.value

    def destroy(self):
        if self._h:
            L.lib().a2afft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Dfft(_Plan):
    """a2adfft (a2adfft.cpp:7-158) with a precision and an execution strategy."""

    def __init__(self, dist, precision=None, exec="fused"):
        if precision is None:
            precision = "fp64" if dist.dtype_code == L.C128 else "fp32"
        self.dist = dist
        self.precision = precision
        self.exec = exec
        self._h = ctypes.c_void_p()
        L.check(L.lib().a2afft_plan_create(ctypes.byref(self._h), dist.handle, PRECISIONS[precision],
                                           EXECS[exec]), "a2adfft")
        self._scratch = None

    def _check_buf(self, b):
        self.dist._check_buf(b)

    def make_plans(self, d_scratch=None):
        """a2adfft::make_plans (a2adfft.cpp:18-24)."""
        if d_scratch is not None:
            self.dist._check_buf(d_scratch)
        self._scratch = d_scratch
        L.check(L.lib().a2afft_make_plans(self._h, _ptr(d_scratch) if d_scratch is not None else None),
                "make_plans")


class GridDfft(_Plan):
    """Fused plan over one nx x ny x nz C-order grid (no rank decomposition)."""

    def __init__(self, shape, precision="fp32"):
        self.shape = tuple(int(s) for s in shape)
        self.precision = precision
        self.torch_dtype = torch.complex128 if precision == "fp64" else torch.complex64
        self._h = ctypes.c_void_p()
        L.check(L.lib().a2afft_grid_plan_create(ctypes.byref(self._h), *self.shape, PRECISIONS[precision]),
                "grid_plan_create")

    def _check_buf(self, b):
        n = self.shape[0] * self.shape[1] * self.shape[2]
        if b.dtype != self.torch_dtype or b.numel() != n or not b.is_contiguous() or not b.is_cuda:
            raise ValueError(f"buffer must be a contiguous device {self.torch_dtype} tensor of {n} elements")

    def empty(self, device="cuda"):
        return torch.empty(self.shape, dtype=self.torch_dtype, device=device)


class Batch:
    """Many independent complex64 C-order grids, one persistent launch per transform."""

    def __init__(self, shapes, precision="fp32", dims=None):
        """dims: per grid (d0, d1, d2) virtual-rank decomposition (None: undecomposed C-order grids);
        a decomposed grid is stored as its rank blocks, rank-major, like a Distribution buffer."""
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.dims = None if dims is None else [tuple(int(v) for v in d) for d in dims]
        flat = [v for s in self.shapes for v in s]
        self._h = ctypes.c_void_p()
        if self.dims is None:
            L.check(L.lib().a2afft_batch_create(ctypes.byref(self._h), len(self.shapes), L.ints(flat),
                                                PRECISIONS[precision]), "batch_create")
        else:
            if len(self.dims) != len(self.shapes):
                raise ValueError("one dims triple per grid")
            L.check(L.lib().a2afft_batch_create_ranks(ctypes.byref(self._h), len(self.shapes), L.ints(flat),
                                                      L.ints([v for d in self.dims for v in d]),
                                                      PRECISIONS[precision]), "batch_create_ranks")
        offs = (ctypes.c_int64 * (len(self.shapes) + 1))()
        L.check(L.lib().a2afft_batch_offsets(self._h, offs), "batch_offsets")
        self.offsets = list(offs)
        self.total = self.offsets[-1]

    def empty(self, device="cuda"):
        return torch.empty(self.total, dtype=torch.complex64, device=device)

    def views(self, buf):
        """Per-grid views of a batch buffer: (nx, ny, nz), or (P, nlocal) rank blocks for a decomposed grid."""
        out = []
        for i, s in enumerate(self.shapes):
            v = buf[self.offsets[i]:self.offsets[i + 1]]
            d = self.dims[i] if self.dims is not None else (1, 1, 1)
            P = d[0] * d[1] * d[2]
            out.append(v.view(*s) if P == 1 else v.view(P, -1))
        return out

    def _check(self, b):
        if b.dtype != torch.complex64 or b.numel() != self.total or not b.is_contiguous() or not b.is_cuda:
            raise ValueError(f"batch buffer must be a contiguous device complex64 tensor of {self.total} elements")

    def execute(self, data, direction=-1):
        self._check(data)
        return L.check(L.lib().a2afft_batch_execute(self._h, _ptr(data), direction, _stream()), "batch_execute")

    def forward(self, data):
        return self.execute(data, -1)

    def inverse(self, data):
        return self.execute(data, 1)

    def graph_launch(self, data, direction=-1):
        """Replay (capturing on first use) the hipGraph of this transform for `data`."""
        self._check(data)
        return L.check(L.lib().a2afft_batch_graph_launch(self._h, _ptr(data), direction, _stream()),
                       "batch_graph_launch")

    def status(self):
        return L.check(L.lib().a2afft_batch_status(self._h, _stream()), "batch_status")

    def set_schedule(self, workgroups=0, xcd_queues=False, chunk=1):
        """Persistent workgroups (0 = all resident slots), per-XCD queue ranges
        and units per queue take; every schedule computes the same bits."""
        return L.check(L.lib().a2afft_batch_set_schedule(self._h, workgroups, 1 if xcd_queues else 0, chunk),
                       "batch_set_schedule")

    def set_grouping(self, group_points=0, skewed=False):
        """Groups of consecutive grids of at most group_points points whose
        passes run group by group (0 = all grids in one group), optionally
        skewed; every order computes the same bits."""
        return L.check(L.lib().a2afft_batch_set_grouping(self._h, int(group_points), 1 if skewed else 0),
                       "batch_set_grouping")

    def autotune(self, work, iters=3):
        """Time schedules and phase orders on `work` (overwritten) and keep the fastest."""
        self._check(work)
        return L.check(L.lib().a2afft_batch_autotune(self._h, _ptr(work), iters, _stream()), "batch_autotune")

    def set_shape(self, N, strided, T=0, E=0, pad=True):
        """Tile shape of the (N, strided) class of passes (T=0: the primary)."""
        return L.check(L.lib().a2afft_batch_set_shape(self._h, N, 1 if strided else 0, T, E, 1 if pad else 0),
                       "batch_set_shape")

    def describe(self):
        buf = ctypes.create_string_buffer(16384)
        L.check(L.lib().a2afft_batch_describe(self._h, buf, len(buf)), "batch_describe")
        return buf.value.decode()

    def schedule(self):
        """The current schedule: workgroups, xcd_queues, chunk, group_points, skewed (and sizes)."""
        first = self.describe().splitlines()[0]
        return {k: int(v) for k, v in (kv.split("=") for kv in first.split())}

    def info(self):
        u, w, m = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        L.check(L.lib().a2afft_batch_info(self._h, ctypes.byref(u), ctypes.byref(w), ctypes.byref(m)), "batch_info")
        return {"units": u.value, "workgroups": w.value, "max_workgroups": m.value}

    def destroy(self):
        if self._h:
            L.lib().a2afft_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- launcher level --------------------------------------------------------

_REMAP_FNS = [
    "a2afft_launch_d_z_a2a_to_z_pencils", "a2afft_launch_d_z_pencils_to_z_a2a",
    "a2afft_launch_d_x_a2a_to_x_pencils", "a2afft_launch_d_x_pencils_to_x_a2a",
    "a2afft_launch_d_y_a2a_to_y_pencils", "a2afft_launch_d_y_pencils_to_y_a2a",
]
_TRANSPOSE_FNS = ["a2afft_launch_d_fast_z_to_x", "a2afft_launch_d_fast_x_to_y", "a2afft_launch_d_fast_y_to_z"]


def _dtype_code(t):
    if t.dtype == torch.complex128:
        return L.C128
    if t.dtype == torch.complex64:
        return L.C64
    raise ValueError("complex64 or complex128 tensor expected")


def launch_remap(kind, source, dest, world_size, nlocal, local_grid_size, dims, nbuffers, blockSize=64):
    """One of the six pack/unpack launchers (kind 0..5) or a transpose (6..8)."""
    lib = L.lib()
    if kind < 6:
        fn = getattr(lib, _REMAP_FNS[kind])
        st = fn(_ptr(source), _ptr(dest), blockSize, world_size, nlocal, L.ints(local_grid_size), L.ints(dims),
                _dtype_code(source), nbuffers, _stream())
        return L.check(st, _REMAP_FNS[kind])
    fn = getattr(lib, _TRANSPOSE_FNS[kind - 6])
    st = fn(_ptr(source), _ptr(dest), L.ints(local_grid_size), blockSize, nlocal, _dtype_code(source), nbuffers,
            _stream())
    return L.check(st, _TRANSPOSE_FNS[kind - 6])


def alltoall(send, recv, world_size, nlocal):
    return L.check(L.lib().a2afft_alltoall(_ptr(send), _ptr(recv), world_size, nlocal, _dtype_code(send),
                                           _stream()), "alltoall")


def forward_1d_fft(data, Ng):
    return L.check(L.lib().a2afft_forward_1d_fft(_ptr(data), Ng, data.numel(), _dtype_code(data), _stream()),
                   "forward_1d_fft")


def inverse_1d_fft(data, Ng):
    return L.check(L.lib().a2afft_inverse_1d_fft(_ptr(data), Ng, data.numel(), _dtype_code(data), _stream()),
                   "inverse_1d_fft")


def scale_fft(out, data, Ng, blockSize=64):
    return L.check(L.lib().a2afft_launch_scale_fft(_ptr(out), _ptr(data), Ng, data.numel(), blockSize,
                                                   _dtype_code(data), _stream()), "scale_fft")


def fast_copy_fft(out, data, blockSize=64):
    return L.check(L.lib().a2afft_launch_fast_copy_fft(_ptr(out), _ptr(data), data.numel(), blockSize,
                                                       _dtype_code(data), _stream()), "fast_copy_fft")


def error_metrics(y, ref128):
    """(rel_l2, max_abs) of y against a complex128 device reference, reduced in FP64 on the GPU."""
    if ref128.dtype != torch.complex128:
        raise ValueError("reference must be complex128")
    if y.numel() != ref128.numel():
        raise ValueError("size mismatch")
    rel = ctypes.c_double()
    mx = ctypes.c_double()
    L.check(L.lib().a2afft_error_metrics(_ptr(y), _dtype_code(y), _ptr(ref128), y.numel(), ctypes.byref(rel),
                                         ctypes.byref(mx), _stream()), "error_metrics")
    return rel.value, mx.value


def fill_random(out, seed, kind="uniform", g0=0):
    """Philox4x32-10 synthetic input on the device (kind 'uniform' or 'normal')."""
    k = {"uniform": 0, "normal": 1}[kind]
    return L.check(L.lib().a2afft_fill_random(_ptr(out), _dtype_code(out), out.numel(), seed, k, g0, _stream()),
                   "fill_random")


def execute_pass(plan, data, pass_index, direction=-1):
    """One axis pass of a fused plan (0 z, 1 x, 2 y), as forward/inverse run it."""
    return L.check(L.lib().a2afft_execute_pass(plan._h, _ptr(data), pass_index, direction, _stream()),
                   "execute_pass")
