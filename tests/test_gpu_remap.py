"""GPU parity of the remap kernels, the device-local all-to-all and the six
a2aDistribution stages: bit-exact against the oracle (itself pinned against
the reference's compiled C twins) on seeded, position-dependent data."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (Ng, P, dims): the reference's own test list (src/a2afft/tests/run_tests.py:18-21)
# under MPI_Dims_create, the decompositions the benchmark configs name, and two
# decompositions outside the reference's y-return domain.
CONFIGS = [
    (8, 8, None), (32, 8, None), (12, 6, None), (4, 4, None),
    (64, 4, (2, 2, 1)), (32, 8, (1, 2, 4)), (32, 16, (1, 4, 4)), (64, 64, (1, 8, 8)),
    (16, 16, (4, 2, 2)), (32, 64, (4, 4, 4)), (16, 8, (2, 4, 1)),
]
DTYPES = ["c128", "c64"]


def _rand_blocks(d, dtype, seed):
    rng = np.random.default_rng(seed)
    n = d.P * d.nlocal
    a = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return a.astype(np.complex128 if dtype == "c128" else np.complex64).reshape(d.P, d.nlocal)


def _to_dev(a):
    import torch
    return torch.from_numpy(a).cuda()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_remap_launchers_bit_exact(a2afft_mod, cuda, cfg, dtype):
    import torch
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    src = _rand_blocks(d, dtype, seed=Ng * 131 + P)
    s_dev = _to_dev(src)
    for kind in range(9):
        out = torch.zeros_like(s_dev)
        st = a2afft_mod.launch_remap(kind, s_dev, out, d.P, d.nlocal, d.lgs_t, d.dims_t, nbuffers=d.P)
        torch.cuda.synchronize()
        ref = O.remap_apply(kind, d, src)
        got = out.cpu().numpy()
        assert np.array_equal(_bits(got), _bits(ref)), f"remap {O.REMAP_NAMES[kind]} differs"
        if kind == O.Y_PENCILS_TO_Y_A2A:
            expect = 0 if O.y_return_in_reference_domain(d) else 1
            assert st == expect


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("cfg", CONFIGS[:8], ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_alltoall_bit_exact(a2afft_mod, cuda, cfg, dtype):
    import torch
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    src = _rand_blocks(d, dtype, seed=7)
    s_dev = _to_dev(src)
    out = torch.zeros_like(s_dev)
    a2afft_mod.alltoall(s_dev, out, d.P, d.nlocal)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(out.cpu().numpy()), _bits(O.alltoall(d, src)))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_stages_bit_exact(a2afft_mod, cuda, cfg, dtype):
    """The six get/return stages in the order a2adfft::forward uses them, with
    the two buffers compared after every stage (a2adistribution.cpp:115-215)."""
    import torch
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    dist = a2afft_mod.Distribution(Ng, P, dims=d.dims_t, dtype=dtype)
    b1 = _rand_blocks(d, dtype, seed=11)
    b2 = _rand_blocks(d, dtype, seed=12)
    g1, g2 = _to_dev(b1), _to_dev(b2)
    h1, h2 = b1.copy(), b2.copy()
    methods = ["getZPencils", "returnZPencils", "getXPencils", "returnXPencils", "getYPencils", "returnYPencils"]
    for k, name in enumerate(methods):
        st = getattr(dist, name)(g1, g2)
        O.stage(k, d, h1, h2)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(g1.cpu().numpy()), _bits(h1)), f"{name}: d_Buff1 differs"
        assert np.array_equal(_bits(g2.cpu().numpy()), _bits(h2)), f"{name}: d_Buff2 differs"
        if name == "returnYPencils":
            assert st == (0 if dist.reference_domain else 1)


@pytest.mark.parametrize("n,off", [(1000, 0), (1001, 0), (1, 0), (4099, 1), (1000, 1), (3, 1)])
def test_copy_and_scale(a2afft_mod, cuda, n, off):
    """Counts and offsets that take the 16-byte lanes (complex128, even
    complex64 counts at aligned addresses) and the 8-byte forms (odd complex64
    counts, views starting 8 bytes into an allocation)."""
    import torch
    rng = np.random.default_rng(3 + n + off)
    for dt in (np.complex128, np.complex64):
        a = (rng.standard_normal(n + off) + 1j * rng.standard_normal(n + off)).astype(dt)
        xb = _to_dev(a)
        yb = torch.zeros_like(xb)
        x, y = xb[off:], yb[off:]
        a2afft_mod.fast_copy_fft(y, x)
        torch.cuda.synchronize()
        out = yb.cpu().numpy()
        assert np.array_equal(_bits(out[off:]), _bits(a[off:]))
        assert not out[:off].any()                   # nothing written before the view
        yb.zero_()
        a2afft_mod.scale_fft(y, x, 12)
        torch.cuda.synchronize()
        # scale_fft divides each component by (double)(Ng^3) (a2affts.cu:83-88)
        ref = (a.real / dt(0).real.dtype.type(12 ** 3)) + 1j * (a.imag / dt(0).real.dtype.type(12 ** 3))
        out = yb.cpu().numpy()
        assert np.array_equal(_bits(out[off:]), _bits(ref[off:].astype(dt)))
        assert not out[:off].any()
