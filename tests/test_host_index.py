"""CPU checks of the HIP library's index arithmetic (no GPU needed).

tests/hostcheck/host_check.cpp is compiled with g++ against the library's own
host+device headers (common.h, remap_index.h, fft_shape.h), so these tests
evaluate the exact expressions the kernels run:
  - FastDiv against integer division;
  - the six pack/unpack index maps against the oracle (itself pinned to the
    reference's compiled C twins), every element;
  - every address each FFT launch loads / stores: in bounds, each element of
    the [P][nlocal] buffer exactly once, and each line holding one global line
    in natural order.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def hc(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("hc") / "libhostcheck.so")
    subprocess.run([os.path.join(HERE, "hostcheck", "build.sh"), out], check=True)
    L = ctypes.CDLL(out)
    P = ctypes.c_void_p
    L.hc_fastdiv.argtypes = [ctypes.c_uint32, P, P, ctypes.c_int64]
    L.hc_remap_dest.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, P, P, P]
    L.hc_fft_enumerate.argtypes = [ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int64, P, ctypes.c_uint32,
                                   ctypes.c_uint32]
    L.hc_fft_enumerate.restype = ctypes.c_int64
    L.hc_shape_fits.argtypes = [ctypes.c_int, ctypes.c_int, P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.hc_xcd_blocks.argtypes = [ctypes.c_uint32, P]
    L.hc_grid_enumerate.argtypes = [ctypes.c_int] * 6 + [P, P, P, ctypes.c_int64]
    L.hc_grid_enumerate.restype = ctypes.c_int64
    L.hc_long_split.argtypes = [ctypes.c_int64, P, P]
    L.hc_long_transpose.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int64]
    L.hc_long_transpose.restype = ctypes.c_int64
    L.hc_long_square.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_int64]
    L.hc_long_square.restype = ctypes.c_int64
    L.hc_long_pass_enumerate.argtypes = [ctypes.c_int] * 6 + [P, P, P, ctypes.c_int64, P]
    L.hc_long_pass_enumerate.restype = ctypes.c_int64
    L.hc_cz_fused_enumerate.argtypes = [P, P, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, P, P, P,
                                        ctypes.c_int64]
    L.hc_cz_fused_enumerate.restype = ctypes.c_int64
    L.hc_grid_check.argtypes = [ctypes.c_int] * 3 + [P] + [ctypes.c_int] * 4 + [ctypes.c_uint32, ctypes.c_uint32, P, P]
    L.hc_grid_check.restype = ctypes.c_int64
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def test_fastdiv_exact(hc):
    rng = np.random.default_rng(0)
    n = np.concatenate([np.arange(0, 70000, dtype=np.uint64),
                        rng.integers(0, 2 ** 32, 200000, dtype=np.uint64),
                        np.array([2 ** 32 - 1, 2 ** 32 - 2, 2 ** 31, 2 ** 31 - 1], dtype=np.uint64)]).astype(np.uint32)
    divisors = list(range(1, 300)) + [2 ** k for k in range(32)] + [2 ** k - 1 for k in range(2, 32)] + \
        [2 ** k + 1 for k in range(2, 31)] + list(rng.integers(1, 2 ** 32, 60))
    q = np.empty_like(n)
    for d in divisors:
        hc.hc_fastdiv(int(d), _p(n), _p(q), n.size)
        assert np.array_equal(q, n // np.uint32(d)), d


REMAP_CONFIGS = [
    (8, 8, None), (32, 8, None), (12, 6, None), (4, 4, None), (64, 4, (2, 2, 1)), (32, 8, (1, 2, 4)),
    (32, 16, (1, 4, 4)), (64, 64, (1, 8, 8)), (16, 16, (4, 2, 2)), (32, 64, (4, 4, 4)), (16, 8, (2, 4, 1)),
    (24, 12, None), (48, 24, None), (64, 32, None), (128, 16, (1, 4, 4)),
]


@pytest.mark.parametrize("cfg", REMAP_CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_remap_index_maps_match_oracle(hc, cfg):
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    lgs = np.array(d.lgs_t, dtype=np.int32)
    dm = np.array(d.dims_t, dtype=np.int32)
    out = np.empty(d.nlocal, dtype=np.uint32)
    for kind in range(6):
        hc.hc_remap_dest(kind, d.P, d.nlocal, _p(lgs), _p(dm), _p(out))
        ref = O.remap_index(kind, d)
        assert np.array_equal(out.astype(np.int64), ref), O.REMAP_NAMES[kind]
        assert np.array_equal(np.sort(out), np.arange(d.nlocal)), "not a permutation"


FFT_CONFIGS = [
    (64, 4, (2, 2, 1)), (8, 8, (2, 2, 2)), (32, 8, (2, 2, 2)), (4, 4, (2, 2, 1)), (32, 8, (1, 2, 4)),
    (64, 16, (1, 4, 4)), (64, 16, (4, 2, 2)), (128, 64, (1, 8, 8)), (32, 64, (4, 4, 4)), (256, 8, (1, 2, 4)),
    (16, 1, (1, 1, 1)), (128, 16, (1, 4, 4)), (2, 1, (1, 1, 1)), (512, 16, (1, 4, 4)),
    # mixed radix: the reference's (P=6, Ng=12) and two more
    (12, 6, (3, 2, 1)), (24, 12, (3, 2, 2)), (60, 4, (1, 2, 2)),
]


def _enumerate(hc, Ng, P, dims, dbl, pass_, T=0, E=0, xcd=0, start=0, count=0):
    cap = (P * (Ng // dims[0]) * (Ng // dims[1]) * (Ng // dims[2])) if pass_ >= 0 else P * Ng
    addr = np.empty(cap, dtype=np.int64)
    n = np.empty(cap, dtype=np.int32)
    line = np.empty(cap, dtype=np.int64)
    shape = np.zeros(4, dtype=np.int64)
    dm = np.array(dims, dtype=np.int32)
    c = hc.hc_fft_enumerate(Ng, P, _p(dm), int(dbl), pass_, T, E, xcd, _p(addr), _p(n), _p(line), cap, _p(shape),
                            start, count)
    assert c >= 0, c
    return addr[:c], n[:c], line[:c], shape


@pytest.mark.parametrize("dbl", [0, 1])
@pytest.mark.parametrize("cfg", FFT_CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_fused_pass_addresses(hc, cfg, dbl):
    Ng, P, dims = cfg
    if Ng ** 3 > 2 ** 24 and dbl:
        pytest.skip("covered by the complex64 case")
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    total = P * d.nlocal
    # global index of every block-layout element
    gid = O.global_to_blocks(d, np.arange(Ng ** 3, dtype=np.int64).reshape((Ng,) * 3)).reshape(-1)
    for pass_, axis in ((0, 2), (1, 0), (2, 1)):
        addr, n, line, shape = _enumerate(hc, Ng, P, dims, dbl, pass_)
        assert addr.size == total
        assert addr.min() >= 0 and addr.max() < total
        assert np.array_equal(np.sort(addr), np.arange(total)), f"pass {pass_}: not every element exactly once"
        coords = np.stack(np.unravel_index(gid[addr], (Ng,) * 3))
        assert np.array_equal(coords[axis], n), f"pass {pass_}: element n is not coordinate n along the axis"
        other = [a for a in range(3) if a != axis]
        key = coords[other[0]] * Ng + coords[other[1]]
        order = np.argsort(line, kind="stable")
        k_sorted = key[order].reshape(-1, Ng)
        assert (k_sorted == k_sorted[:, :1]).all(), f"pass {pass_}: a line mixes global lines"


@pytest.mark.parametrize("N", [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 3, 5, 6, 12, 15, 30, 45, 60])
@pytest.mark.parametrize("lines", [1, 7, 33])
def test_batched_pencil_addresses(hc, N, lines):
    for dbl in (0, 1):
        addr, n, line, shape = _enumerate(hc, N, lines, (1, 1, 1), dbl, -1)
        assert np.array_equal(np.sort(addr), np.arange(N * lines))
        assert np.array_equal(addr % N, n)
        assert np.array_equal(addr // N, line)


def test_xcd_renumbering_is_a_bijection(hc):
    for G in list(range(1, 300)) + [4095, 4096, 4097, 16384, 32768, 100003]:
        out = np.empty(G, dtype=np.uint32)
        hc.hc_xcd_blocks(G, _p(out))
        assert np.array_equal(np.sort(out), np.arange(G)), G
        if G >= 16:
            # each XCD (b % 8) receives a contiguous logical range
            for x in range(8):
                r = out[x::8].astype(np.int64)
                assert np.all(np.diff(r) == 1), (G, x)


SHAPE_CONFIGS = [(64, 16, (1, 4, 4)), (32, 8, (1, 2, 4)), (16, 4, (2, 2, 1)), (128, 16, (1, 4, 4)),
                 (12, 6, (3, 2, 1)), (60, 4, (1, 2, 2))]


@pytest.mark.parametrize("cfg", SHAPE_CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_every_fitting_shape_addresses(hc, cfg):
    """Every (T, E) the dispatcher accepts, with and without XCD renumbering:
    each element exactly once, each line one global line in natural order."""
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    total = P * d.nlocal
    gid = O.global_to_blocks(d, np.arange(Ng ** 3, dtype=np.int64).reshape((Ng,) * 3)).reshape(-1)
    dm = np.array(dims, dtype=np.int32)
    checked = 0
    for pass_, axis in ((0, 2), (1, 0), (2, 1)):
        for T in (1, 2, 4, 8, 16, 32, 64, 128):
            for E in sorted({2, 4, 8, 16, 32, 64, Ng}):
                if not hc.hc_shape_fits(Ng, P, _p(dm), pass_, T, E):
                    continue
                for xcd in (0, 1, 2, 3):
                    addr, n, line, shape = _enumerate(hc, Ng, P, dims, 0, pass_, T, E, xcd)
                    assert np.array_equal(np.sort(addr), np.arange(total)), (pass_, T, E, xcd)
                    coords = np.stack(np.unravel_index(gid[addr], (Ng,) * 3))
                    assert np.array_equal(coords[axis], n)
                    other = [a for a in range(3) if a != axis]
                    key = coords[other[0]] * Ng + coords[other[1]]
                    k_sorted = key[np.argsort(line, kind="stable")].reshape(-1, Ng)
                    assert (k_sorted == k_sorted[:, :1]).all()
                    checked += 1
    assert checked > 10


@pytest.mark.parametrize("cfg", [(64, 16, (1, 4, 4), 16, 16), (128, 16, (1, 4, 4), 32, 32), (32, 8, (1, 2, 4), 8, 8)],
                         ids=lambda c: f"Ng{c[0]}_{c[2]}_T{c[3]}")
def test_chunk_windows_cover_each_element_once(hc, cfg):
    """x and y launches over z windows (chunked execution): the windows of one
    pass together touch every element exactly once, each window only its own z
    columns."""
    Ng, P, dims, T, E = cfg
    rc, d = O.decompose(Ng, P, dims)
    total = P * d.nlocal
    gid = O.global_to_blocks(d, np.arange(Ng ** 3, dtype=np.int64).reshape((Ng,) * 3)).reshape(-1)
    for chunk in (T, 2 * T, Ng // 2):
        for pass_ in (1, 2):
            seen = []
            for z0 in range(0, Ng, chunk):
                addr, n, line, shape = _enumerate(hc, Ng, P, dims, 0, pass_, T, E, 1, z0, chunk)
                z = np.unravel_index(gid[addr], (Ng,) * 3)[2]
                assert z.min() >= z0 and z.max() < z0 + chunk
                seen.append(addr)
            allv = np.concatenate(seen)
            assert np.array_equal(np.sort(allv), np.arange(total)), (chunk, pass_)


# (N, strided) -> (T, E) shapes of the batched kernel (csrc/batch.hip kShapes)
BATCH_SHAPES = {(16, 0): [(256, 16)], (32, 0): [(64, 8)], (64, 0): [(32, 8)], (128, 0): [(32, 16)],
                (256, 0): [(16, 16)], (512, 0): [(8, 16)],
                (16, 1): [(32, 16), (1, 16)], (32, 1): [(32, 8), (1, 8)], (64, 1): [(32, 8), (1, 8)],
                (128, 1): [(32, 16), (1, 16)], (256, 1): [(16, 16), (1, 16)], (512, 1): [(8, 16), (1, 8)]}


@pytest.mark.parametrize("shape", [(32, 32, 32), (128, 256, 64), (16, 32, 64), (64, 16, 128), (32, 64, 16),
                                   (16, 16, 16), (256, 64, 512), (512, 16, 32)], ids=str)
def test_non_cubic_grid_addresses(hc, shape):
    """Plain C-order grids (single-grid and batched plans): every (T, E) the
    batched kernel may use, each element once, lines in natural order."""
    nx, ny, nz = shape
    total = nx * ny * nz
    coords_all = np.stack(np.unravel_index(np.arange(total), shape))
    for pass_, axis in ((0, 2), (1, 0), (2, 1)):
        N = shape[axis]
        for T, E in BATCH_SHAPES[(N, 0 if pass_ == 0 else 1)]:
            addr = np.empty(total, dtype=np.int64)
            n = np.empty(total, dtype=np.int32)
            line = np.empty(total, dtype=np.int64)
            c = hc.hc_grid_enumerate(nx, ny, nz, pass_, T, E, _p(addr), _p(n), _p(line), total)
            if c == -7:          # shape does not fit this grid (e.g. T wider than nz): dispatcher falls back
                continue
            assert c == total, (pass_, T, E, c)
            assert np.array_equal(np.sort(addr), np.arange(total))
            coords = coords_all[:, addr]
            assert np.array_equal(coords[axis], n)
            other = [a for a in range(3) if a != axis]
            key = coords[other[0]] * 4096 + coords[other[1]]
            k_sorted = key[np.argsort(line, kind="stable")].reshape(-1, N)
            assert (k_sorted == k_sorted[:, :1]).all()


def test_long_line_split(hc):
    """Lines above 4096 split into N1 * N2, both in 2..4096, N1 >= N2; primes
    above 4096 and lengths above 4096^2 have no split."""
    n1, n2 = ctypes.c_int(), ctypes.c_int()
    for N in [4097 * 2, 8192, 10000, 12288, 8194, 1 << 20, 4096 * 4096, 4095 * 4093, 6000, 4100]:
        assert hc.hc_long_split(N, ctypes.byref(n1), ctypes.byref(n2)) == 1, N
        assert n1.value * n2.value == N and 2 <= n2.value <= n1.value <= 4096, (N, n1.value, n2.value)
        # no divisor between sqrt(N) and N1: the most balanced split
        assert not any(N % d == 0 and N // d <= 4096 for d in range(int(np.ceil(np.sqrt(N))), n1.value)), N
    for N in [4096, 4099, 8191, 4099 * 4097, 4096 * 4096 + 2, 1 << 25]:
        assert hc.hc_long_split(N, ctypes.byref(n1), ctypes.byref(n2)) == 0, N


@pytest.mark.parametrize("case", [(128, 64, 3), (241, 34, 2), (100, 100, 1), (128, 96, 2), (1024, 1024, 1),
                                  (33, 31, 5)], ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}")
def test_long_line_transpose_addresses(hc, case):
    """The twiddle-transpose's workgroups move every element of the chunk
    exactly once, in bounds, as the per-line transposition [N1][N2] ->
    [N2][N1] (numpy's transpose as the reference)."""
    N1, N2, lines = case
    total = lines * N1 * N2
    src = np.empty(total + 1, np.int64)
    dst = np.empty(total + 1, np.int64)
    c = hc.hc_long_transpose(N1, N2, lines, _p(src), _p(dst), total + 1)
    assert c == total
    src, dst = src[:c], dst[:c]
    assert np.array_equal(np.sort(src), np.arange(total)) and np.array_equal(np.sort(dst), np.arange(total))
    moved = np.empty(total, np.int64)
    moved[dst] = src                           # out[dst] = in[src]
    ref = np.arange(total).reshape(lines, N1, N2).transpose(0, 2, 1).reshape(-1)
    assert np.array_equal(moved, ref)


@pytest.mark.parametrize("case", [(128, 3), (100, 2), (33, 4), (256, 1), (1024, 1), (31, 2)],
                         ids=lambda c: f"{c[0]}sq_x{c[1]}")
def test_long_line_square_transpose_in_place(hc, case):
    """The in-place transpose of square lines: every element moved once, as
    numpy's transpose, and each workgroup writes exactly the elements it read
    (no two workgroups touch the same element)."""
    M, lines = case
    total = lines * M * M
    src = np.empty(total + 1, np.int64)
    dst = np.empty(total + 1, np.int64)
    wg = np.empty(total + 1, np.int64)
    c = hc.hc_long_square(M, lines, _p(src), _p(dst), _p(wg), total + 1)
    assert c == total
    src, dst, wg = src[:c], dst[:c], wg[:c]
    assert np.array_equal(np.sort(src), np.arange(total)) and np.array_equal(np.sort(dst), np.arange(total))
    moved = np.empty(total, np.int64)
    moved[dst] = src
    ref = np.arange(total).reshape(lines, M, M).transpose(0, 2, 1).reshape(-1)
    assert np.array_equal(moved, ref)
    order = np.argsort(wg, kind="stable")
    w, s_, d_ = wg[order], src[order], dst[order]
    cuts = np.flatnonzero(np.diff(w)) + 1
    for ss, dd in zip(np.split(s_, cuts), np.split(d_, cuts)):
        assert np.array_equal(np.sort(ss), np.sort(dd))


@pytest.mark.parametrize("case", [(128, 64, 3), (128, 128, 2), (256, 256, 1), (100, 100, 2), (241, 34, 2),
                                  (128, 96, 2), (1024, 1024, 1), (64, 128, 2)], ids=lambda c: f"{c[0]}x{c[1]}x{c[2]}")
def test_long_line_column_pass_addresses(hc, case):
    """The four-step's column FFT pass (N1-point FFTs down the N2 columns of
    each line), for the default shape and every power-of-two tile that fits:
    each element of the lines exactly once, and each column its N1 elements
    in natural order."""
    N1, N2, lines = case
    total = lines * N1 * N2
    pow2 = N1 & (N1 - 1) == 0
    # power-of-two lengths: the default and the Stockham shapes; others: the
    # (T, E) shapes of the runtime-schedule kernel (E dividing N1) and the
    # direct DFT's one line per workgroup (T = 1, E = N1)
    Es = (2, 4, 8, 16, 32) if pow2 else [e for e in range(2, 65) if N1 % e == 0] + [N1]
    shapes = ([(0, 0)] if pow2 else []) + [(T, E) for T in (1, 2, 4, 8, 16, 32, 64) for E in Es]
    checked = 0
    for dbl in (0, 1):
        for T, E in shapes:
            addr = np.empty(total + 1, np.int64)
            n = np.empty(total + 1, np.int32)
            col = np.empty(total + 1, np.int64)
            shape = np.zeros(2, np.int64)
            c = hc.hc_long_pass_enumerate(N1, N2, lines, dbl, T, E, _p(addr), _p(n), _p(col), total + 1, _p(shape))
            if c < 0:
                assert (T, E) != (0, 0), (N1, N2, dbl)   # the default shape always fits
                continue
            assert c == total, (T, E)
            addr, n, col = addr[:c], n[:c], col[:c]
            assert np.array_equal(np.sort(addr), np.arange(total))
            line, r = addr // (N1 * N2), addr % (N1 * N2)
            assert np.array_equal(r // N2, n)                    # element n1 of its column
            assert np.array_equal(line * N2 + r % N2, col)      # the column it belongs to
            checked += 1
    assert checked >= 2


# (grid, dims): extents with prime factors above 5, over virtual ranks
CZ_GRIDS = [((22, 26, 34), (2, 2, 2)), ((13, 17, 19), (1, 1, 1)), ((33, 14, 26), (3, 2, 2)),
            ((11, 12, 97), (1, 2, 1))]


@pytest.mark.parametrize("cfg", CZ_GRIDS + [((251, 6, 20), (1, 2, 2)), ((9, 37, 1000), (1, 1, 2)),
                                            ((4093, 2, 4), (1, 1, 2)), ((3, 2050, 8), (1, 2, 1))],
                         ids=lambda c: f"{c[0]}_{c[1]}")
def test_chirp_pass_addresses(hc, cfg):
    """The chirp-z pass (one kernel, M up to 8192): every element of the pass
    (or of a z window) loaded and stored exactly once by the thread the
    kernel's lane order gives it, element n of a line at coordinate n along the
    axis, each line one global line."""
    shape, dims = cfg
    n3 = np.array(shape, dtype=np.int32)
    dm = np.array(dims, dtype=np.int32)
    P = int(np.prod(dims))
    lgs = [s // d for s, d in zip(shape, dims)]
    total = P * lgs[0] * lgs[1] * lgs[2]
    g = np.arange(total, dtype=np.int64).reshape(shape)
    blocks = g.reshape(dims[0], lgs[0], dims[1], lgs[1], dims[2], lgs[2]).transpose(0, 2, 4, 1, 3, 5).reshape(-1)
    for pass_, axis in ((0, 2), (1, 0), (2, 1)):
        N = shape[axis]
        if N < 9:
            continue
        other = [a for a in range(3) if a != axis]
        for start, count in ([(0, 0)] + ([(2, shape[2] - 3)] if pass_ > 0 else [])):
            cnt = count if count else shape[2]
            cap = (total // shape[2]) * cnt if count else total
            addr = np.empty(cap, np.int64)
            n = np.empty(cap, np.int32)
            line = np.empty(cap, np.int64)
            c = hc.hc_cz_fused_enumerate(_p(n3), _p(dm), pass_, start, count, _p(addr), _p(n), _p(line), cap)
            assert c == cap, (pass_, c)
            if not count:
                assert np.array_equal(np.sort(addr), np.arange(total)), pass_
            else:
                assert np.unique(addr).size == cap
            coords = np.stack(np.unravel_index(blocks[addr], shape))
            assert np.array_equal(coords[axis], n), pass_
            key = coords[other[0]] * 4096 + coords[other[1]]
            ks = key[np.argsort(line, kind="stable")].reshape(-1, N)
            assert (ks == ks[:, :1]).all(), pass_


# (grid, dims, pass, T, E, flags): the default launches of whole grids of
# 2^32 points and more (2048^3 complex64 as one grid and over 64 virtual
# ranks, 2048^2 x 1024 complex128), and small grids where every workgroup is
# checked.
BIG_GRID_CASES = [((2048, 2048, 2048), (1, 1, 1), 0, 2, 16, 7), ((2048, 2048, 2048), (1, 1, 1), 1, 16, 64, 39),
                  ((2048, 2048, 2048), (1, 1, 1), 1, 16, 64, 41),
                  ((2048, 2048, 2048), (1, 1, 1), 2, 16, 64, 33), ((2048, 2048, 1024), (1, 1, 1), 0, 4, 16, 1),
                  ((2048, 2048, 1024), (1, 1, 1), 1, 8, 32, 39), ((2048, 2048, 1024), (1, 1, 1), 2, 8, 32, 33),
                  ((1024, 1024, 1024), (1, 1, 1), 1, 16, 32, 37), ((64, 64, 64), (1, 1, 1), 1, 16, 16, 9),
                  ((2048, 2048, 2048), (1, 8, 8), 0, 2, 16, 7), ((2048, 2048, 2048), (1, 8, 8), 1, 16, 64, 41),
                  ((2048, 2048, 2048), (1, 8, 8), 2, 16, 64, 33), ((2048, 2048, 2048), (4, 4, 4), 0, 2, 16, 7),
                  ((2048, 2048, 2048), (4, 4, 4), 1, 16, 64, 41), ((2048, 2048, 2048), (4, 4, 4), 2, 16, 64, 33),
                  ((64, 64, 64), (2, 2, 2), 0, 4, 16, 1), ((64, 64, 64), (1, 4, 4), 2, 16, 16, 8)]


@pytest.mark.parametrize("case", BIG_GRID_CASES,
                         ids=lambda c: f"{'x'.join(map(str, c[0]))}_dims{''.join(map(str, c[1]))}_pass{c[2]}")
def test_grid_addresses_64bit(hc, case):
    """Every address a pass computes (line_base + elem_off, XCD renumbering and
    unit order included) equals the element's index computed directly in 64
    bits from its global (x, y, z) in the block layout over dims: for the
    first and last 512 workgroups and 36 runs of 64 spread over the launch
    (every workgroup on the 64^3 grids).  A 32-bit overflow anywhere in the
    geometry shows up as a mismatch or an address beyond the grid."""
    shape, dims, pass_, T, E, flags = case
    d = (ctypes.c_int * 3)(*dims)
    grid, cnt = ctypes.c_int64(), ctypes.c_int64()
    assert hc.hc_grid_check(*shape, d, pass_, T, E, flags, 0, 1, ctypes.byref(grid), ctypes.byref(cnt)) == 0
    G = grid.value
    runs = [(0, G)] if G <= 4096 else [(0, 512), (G - 512, G)] + [(G * k // 37, G * k // 37 + 64) for k in range(1, 37)]
    total = 0
    for a, b in runs:
        bad = hc.hc_grid_check(*shape, d, pass_, T, E, flags, a, b, ctypes.byref(grid), ctypes.byref(cnt))
        assert bad == 0, (a, b, bad)
        total += cnt.value
    assert total > 0
