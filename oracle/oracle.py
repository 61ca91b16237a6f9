"""ctypes wrapper over the a2afft oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liba2afft_oracle.so`` (the C restatement of the reference
dataflow) and, when built, ``oracle/_ref/libref_twins.so`` (the reference's own
CPU C twins of the nine remaps, compiled from
src/pencils_demo/{a2a_to_pencils,change_fast_axis,helpers}/*.c).

Only tests/, ``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py
may import this module; it is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

Z_A2A_TO_Z_PENCILS, Z_PENCILS_TO_Z_A2A = 0, 1
X_A2A_TO_X_PENCILS, X_PENCILS_TO_X_A2A = 2, 3
Y_A2A_TO_Y_PENCILS, Y_PENCILS_TO_Y_A2A = 4, 5
FAST_Z_TO_X, FAST_X_TO_Y, FAST_Y_TO_Z = 6, 7, 8
REMAP_NAMES = [
    "z_a2a_to_z_pencils", "z_pencils_to_z_a2a", "x_a2a_to_x_pencils",
    "x_pencils_to_x_a2a", "y_a2a_to_y_pencils", "y_pencils_to_y_a2a",
    "fast_z_to_x", "fast_x_to_y", "fast_y_to_z",
]
GET_Z, RETURN_Z, GET_X, RETURN_X, GET_Y, RETURN_Y = range(6)


class Decomp(ctypes.Structure):
    _fields_ = [("Ng", ctypes.c_int), ("P", ctypes.c_int),
                ("dims", ctypes.c_int * 3), ("lgs", ctypes.c_int * 3),
                ("nlocal", ctypes.c_int64)]

    @property
    def dims_t(self):
        return tuple(self.dims)

    @property
    def lgs_t(self):
        return tuple(self.lgs)


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liba2afft_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_dims_create.argtypes = [_I, ctypes.POINTER(_I)]
        L.orc_decompose.argtypes = [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(Decomp)]
        L.orc_decompose.restype = _I
        L.orc_y_return_in_reference_domain.argtypes = [ctypes.POINTER(Decomp)]
        L.orc_y_return_in_reference_domain.restype = _I
        L.orc_remap_index.argtypes = [_I, _I64, ctypes.POINTER(Decomp), _I]
        L.orc_remap_index.restype = _I64
        L.orc_remap_apply.argtypes = [_I, ctypes.POINTER(Decomp), _P, _P, _I]
        L.orc_alltoall.argtypes = [ctypes.POINTER(Decomp), _P, _P, _I]
        L.orc_stage.argtypes = [_I, ctypes.POINTER(Decomp), _P, _P, _I]
        L.orc_fft_batched.argtypes = [_P, _I, _I64, _I]
        L.orc_fft_batched.restype = _I
        L.orc_dft_naive.argtypes = [_P, _P, _I, _I]
        L.orc_twiddle.argtypes = [_I64, _I64, _I, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double)]
        L.orc_forward.argtypes = [ctypes.POINTER(Decomp), _P, _P]
        L.orc_forward.restype = _I
        L.orc_inverse.argtypes = [ctypes.POINTER(Decomp), _P, _P]
        L.orc_inverse.restype = _I
        L.orc_blocks_to_global.argtypes = [ctypes.POINTER(Decomp), _P, _P, _I]
        L.orc_global_to_blocks.argtypes = [ctypes.POINTER(Decomp), _P, _P, _I]
        L.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_uint32)]
        L.orc_fill_uniform.argtypes = [ctypes.c_uint64, _I64, _P]
        L.orc_fill_normal.argtypes = [ctypes.c_uint64, _I64, _P]
        L.orc_set_threads.argtypes = [_I]
        L.orc_get_threads.restype = _I
        _LIB = L
    return _LIB


def ref_twins():
    """The reference's compiled C twins, or None when _ref/ was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_twins.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        ip = ctypes.POINTER(ctypes.c_int)
        for name in REMAP_NAMES[:6]:
            getattr(R, name).argtypes = [ip, ip, _I, _I, _I, ip, ip]
        for name in REMAP_NAMES[6:]:
            getattr(R, name).argtypes = [ip, ip, ip]
        R.init_grid_cells.argtypes = [_I, ip, ip, ip, _I]
        _REF = R
    return _REF


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def dims_create(P):
    d = (ctypes.c_int * 3)()
    lib().orc_dims_create(P, d)
    return tuple(d)


def decompose(Ng, P, dims=None):
    d = Decomp()
    arr = (ctypes.c_int * 3)(*dims) if dims is not None else None
    rc = lib().orc_decompose(Ng, P, arr, ctypes.byref(d))
    return rc, d


def y_return_in_reference_domain(d):
    return bool(lib().orc_y_return_in_reference_domain(ctypes.byref(d)))


def remap_index(kind, d, ref_formula=False):
    """Vector of orc_remap_index over one rank's [0, nlocal)."""
    L = lib()
    return np.array([L.orc_remap_index(kind, i, ctypes.byref(d), int(ref_formula))
                     for i in range(d.nlocal)], dtype=np.int64)


def remap_perm(kind, d):
    """Scatter permutation of one rank: perm[src_index] = dst_index."""
    idx = remap_index(kind, d)
    if kind >= FAST_Z_TO_X:  # gather: dest[i] = src[idx[i]]  ->  perm[idx[i]] = i
        perm = np.empty_like(idx)
        perm[idx] = np.arange(d.nlocal)
        return perm
    return idx


def remap_apply(kind, d, src):
    src = np.ascontiguousarray(src)
    dst = np.empty_like(src)
    lib().orc_remap_apply(kind, ctypes.byref(d), _ptr(src), _ptr(dst), src.itemsize)
    return dst


def alltoall(d, send):
    send = np.ascontiguousarray(send)
    recv = np.empty_like(send)
    lib().orc_alltoall(ctypes.byref(d), _ptr(send), _ptr(recv), send.itemsize)
    return recv


def stage(kind, d, b1, b2):
    """Runs one a2aDistribution stage in place on (b1, b2)."""
    assert b1.flags.c_contiguous and b2.flags.c_contiguous
    lib().orc_stage(kind, ctypes.byref(d), _ptr(b1), _ptr(b2), b1.itemsize)


def fft_batched(x, sign=-1):
    """In-place-style batched 1D FFT over the last axis (complex128)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    N = a.shape[-1]
    rc = lib().orc_fft_batched(_ptr(a), N, a.size // N, sign)
    assert rc == 0, rc
    return a


def dft_naive(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft_naive(_ptr(x), _ptr(out), x.shape[-1], sign)
    return out


def twiddle(m, M, sign=-1):
    re, im = ctypes.c_double(), ctypes.c_double()
    lib().orc_twiddle(m, M, sign, ctypes.byref(re), ctypes.byref(im))
    return complex(re.value, im.value)


def forward(d, blocks):
    """3D forward in the reference stage order; blocks: [P][nlocal] complex128."""
    D = np.ascontiguousarray(blocks, dtype=np.complex128).copy()
    S = np.zeros_like(D)
    assert lib().orc_forward(ctypes.byref(d), _ptr(D), _ptr(S)) == 0
    return D


def inverse(d, blocks):
    D = np.ascontiguousarray(blocks, dtype=np.complex128).copy()
    S = np.zeros_like(D)
    assert lib().orc_inverse(ctypes.byref(d), _ptr(D), _ptr(S)) == 0
    return D


def blocks_to_global(d, blocks):
    blocks = np.ascontiguousarray(blocks)
    out = np.empty((d.Ng,) * 3, dtype=blocks.dtype)
    lib().orc_blocks_to_global(ctypes.byref(d), _ptr(blocks), _ptr(out), blocks.itemsize)
    return out


def global_to_blocks(d, g):
    g = np.ascontiguousarray(g)
    out = np.empty((d.P, d.nlocal), dtype=g.dtype)
    lib().orc_global_to_blocks(ctypes.byref(d), _ptr(g), _ptr(out), g.itemsize)
    return out


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return tuple(o)


def fill_uniform(seed, n):
    out = np.empty(n, dtype=np.complex128)
    lib().orc_fill_uniform(seed, n, _ptr(out))
    return out


def fill_normal(seed, n):
    out = np.empty(n, dtype=np.complex128)
    lib().orc_fill_normal(seed, n, _ptr(out))
    return out


def set_threads(n):
    lib().orc_set_threads(n)


def get_threads():
    return lib().orc_get_threads()


# --- reference twins (compiled reference C) -------------------------------

def twin_perm(kind, d, P=None):
    """Scatter permutation from the reference's own C twin, fed with iota.

    Returns perm[src] = dst (for the transposes the twin's gather is inverted).
    For y_pencils_to_y_a2a outside the reference's valid domain the twin writes
    out of bounds; callers must not ask for that case.
    """
    R = ref_twins()
    if R is None:
        return None
    n = int(d.nlocal)
    src = np.arange(n, dtype=np.int32)
    lgs = np.array(d.lgs_t, dtype=np.int32)
    dims = np.array(d.dims_t, dtype=np.int32)
    if kind == Y_PENCILS_TO_Y_A2A and not y_return_in_reference_domain(d):
        raise ValueError("twin y_pencils_to_y_a2a writes out of bounds for these dims")
    dst = np.full(n, -1, dtype=np.int32)
    name = REMAP_NAMES[kind]
    if kind < FAST_Z_TO_X:
        getattr(R, name)(_ip(src), _ip(dst), d.P, 0, n, _ip(lgs), _ip(dims))
        # dst[new] = src[idx] = idx  ->  perm[idx] = new
        perm = np.empty(n, dtype=np.int64)
        perm[dst] = np.arange(n)
        assert (dst >= 0).all()
        return perm
    getattr(R, name)(_ip(src), _ip(dst), _ip(lgs))
    # dst[i] = src[g(i)] = g(i)  ->  perm[g(i)] = i
    perm = np.empty(n, dtype=np.int64)
    perm[dst.astype(np.int64)] = np.arange(n)
    return perm
