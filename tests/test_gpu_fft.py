"""GPU parity of the Stockham kernels and of the 3D transform.

1D: forward_1d_fft / inverse_1d_fft against the FP64 oracle FFT and against
known answers.  3D: a2adfft forward / inverse, both execution strategies,
against the oracle's reference-stage-order transform; the fused strategy must
equal the staged one bit for bit (the remaps are permutations and both use the
same per-line arithmetic)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LENGTHS = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
# mixed radix 2^a 3^b 5^c: (N, dtypes built for it)
MIXED = [(3, ("c128", "c64")), (5, ("c128", "c64")), (6, ("c128", "c64")), (12, ("c128", "c64")),
         (15, ("c128", "c64")), (27, ("c128", "c64")), (30, ("c128", "c64")), (45, ("c64",)), (48, ("c64",)),
         (60, ("c64",))]
# rel-L2 bounds against the FP64 oracle
TOL_1D = {"c128": 1e-15, "c64": 1.5e-7}


def _relerr(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("N", LENGTHS + [n for n, _ in MIXED])
def test_batched_1d_fft_vs_oracle(a2afft_mod, cuda, N, dtype):
    """Stockham lengths; mixed lengths with no kernel in a precision run as the direct DFT."""
    import torch
    rng = np.random.default_rng(N)
    lines = max(3, 4096 // N) + 1          # ragged batch: not a multiple of the lines per block
    x = rng.standard_normal((lines, N)) + 1j * rng.standard_normal((lines, N))
    x = x.astype(np.complex128 if dtype == "c128" else np.complex64)
    ref_f = O.fft_batched(x.astype(np.complex128), -1)
    ref_i = O.fft_batched(x.astype(np.complex128), +1)
    d = torch.from_numpy(x).cuda()
    a2afft_mod.forward_1d_fft(d, N)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), ref_f) < TOL_1D[dtype] * max(1, np.log2(N))
    d = torch.from_numpy(x).cuda()
    a2afft_mod.inverse_1d_fft(d, N)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), ref_i) < TOL_1D[dtype] * max(1, np.log2(N))


# 2^a 3^b 5^c lengths above the one-thread-per-line limit: the Stockham kernel
# with the radix schedule in its arguments (radix 4/2, then 3s, then 5s).
MIXED_LONG = [72, 96, 100, 120, 192, 225, 243, 384, 625, 720, 768, 900, 1000, 1536, 1800, 2187, 3072, 3125, 3375, 3600,
              3840, 4000]


KF_DFT, KF_MIX, KF_CZ, KF_MIXL = 64, 128, 2048, 4096


def _chosen_flags(plan, name):
    for ln in plan.describe().splitlines():
        if ln.startswith(f"chosen {name} "):
            return int(ln.split("flags=")[1].split()[0])
    raise AssertionError(plan.describe())


def _smooth(n):
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
    return n == 1


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("N", MIXED_LONG)
def test_long_mixed_lengths_vs_oracle(a2afft_mod, cuda, N, dtype):
    import torch
    rng = np.random.default_rng(N + 1)
    lines = 7
    x = rng.standard_normal((lines, N)) + 1j * rng.standard_normal((lines, N))
    x = x.astype(np.complex128 if dtype == "c128" else np.complex64)
    for sign, fn in ((-1, a2afft_mod.forward_1d_fft), (1, a2afft_mod.inverse_1d_fft)):
        ref = O.fft_batched(x.astype(np.complex128), sign)
        d = torch.from_numpy(x).cuda()
        fn(d, N)
        torch.cuda.synchronize()
        err = _relerr(d.cpu().numpy(), ref)
        assert err < TOL_1D[dtype] * np.log2(N), (N, dtype, sign, err)


def test_long_mixed_impulses(a2afft_mod, cuda):
    """N = 384 = 2^7 * 3 and 3125 = 5^5: an impulse at n0 transforms to exp(-2 pi i k n0 / N)."""
    import torch
    for N, n0 in ((384, 7), (3125, 1234)):
        x = np.zeros((1, N), dtype=np.complex128)
        x[0, n0] = 1.0
        d = torch.from_numpy(x).cuda()
        a2afft_mod.forward_1d_fft(d, N)
        k = np.arange(N)
        want = np.exp(-2j * np.pi * ((k * n0) % N) / N)   # reduced first: exact phases for numpy's exp
        assert np.abs(d.cpu().numpy()[0] - want).max() < 1e-13, N


# Lengths with other prime factors (7, 11, 13, 97): the direct O(N^2) DFT
# below 9 points, whose rounding grows like sqrt(N), and chirp-z convolutions
# above (two power-of-two FFTs: log2 N times a few).
DIRECT = [7, 11, 14, 49, 63, 97, 4095]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("N", DIRECT)
def test_direct_dft_lengths_vs_oracle(a2afft_mod, cuda, N, dtype):
    import torch
    rng = np.random.default_rng(N)
    lines = 5
    x = rng.standard_normal((lines, N)) + 1j * rng.standard_normal((lines, N))
    x = x.astype(np.complex128 if dtype == "c128" else np.complex64)
    tol = TOL_1D[dtype] * max(np.sqrt(N), 3 * np.log2(N))
    for sign, fn in ((-1, a2afft_mod.forward_1d_fft), (1, a2afft_mod.inverse_1d_fft)):
        ref = O.fft_batched(x.astype(np.complex128), sign)
        d = torch.from_numpy(x).cuda()
        fn(d, N)
        torch.cuda.synchronize()
        err = _relerr(d.cpu().numpy(), ref)
        assert err < tol, (N, dtype, sign, err)


def test_direct_dft_known_answers(a2afft_mod, cuda):
    """N = 97 (prime): an impulse at n0 gives exp(-2 pi i k n0 / N); a pure
    tone exp(2 pi i k0 n / N) gives N at k0 and ~0 elsewhere."""
    import torch
    N, n0, k0 = 97, 5, 11
    x = np.zeros((2, N), dtype=np.complex128)
    x[0, n0] = 1.0
    x[1] = np.exp(2j * np.pi * k0 * np.arange(N) / N)
    d = torch.from_numpy(x).cuda()
    a2afft_mod.forward_1d_fft(d, N)
    y = d.cpu().numpy()
    k = np.arange(N)
    assert np.abs(y[0] - np.exp(-2j * np.pi * ((k * n0) % N) / N)).max() < 1e-14
    e = np.zeros(N, dtype=np.complex128)
    e[k0] = N
    assert np.abs(y[1] - e).max() < 1e-12


@pytest.mark.parametrize("N", [8, 64, 512, 2048])
def test_1d_known_answers(a2afft_mod, cuda, N):
    """delta -> ones, ones -> N*delta, single tone -> N*delta_k0 (FP64)."""
    import torch
    k0 = 3
    x = np.zeros((3, N), dtype=np.complex128)
    x[0, 0] = 1.0
    x[1, :] = 1.0
    x[2, :] = np.exp(2j * np.pi * k0 * np.arange(N) / N)
    d = torch.from_numpy(x).cuda()
    a2afft_mod.forward_1d_fft(d, N)
    torch.cuda.synchronize()
    y = d.cpu().numpy()
    assert np.allclose(y[0], 1.0, atol=1e-15)
    e = np.zeros(N, dtype=np.complex128)
    e[0] = N
    assert np.allclose(y[1], e, atol=1e-12 * N)
    e = np.zeros(N, dtype=np.complex128)
    e[k0] = N
    assert np.allclose(y[2], e, atol=1e-12 * N)


# (Ng, P, dims, dtype) -- C1 shape, the reference's own small configs and the
# decompositions of the benchmark configs at reduced Ng.
CONFIGS_3D = [
    (64, 4, (2, 2, 1), "c128"),
    (14, 2, None, "c128"),        # the direct DFT (a factor 7) on every axis, blocks of 7 x 14 x 14
    (14, 2, None, "c64"),
    (21, 3, None, "c128"),        # 3 * 7 over dims [3,1,1]
    (100, 4, (2, 2, 1), "c64"),   # 2^2 * 5^2, above the one-thread-per-line limit: radix-10 passes (4 * 5 * 5)
    (45, 9, None, "c128"),        # no fixed complex128 kernel for 45: runtime schedule
    (96, 8, (1, 2, 4), "c128"),   # 2^5 * 3 over blocks of 96 x 48 x 24
    (96, 8, (1, 2, 4), "c64"),
    (120, 8, (2, 2, 2), "c128"),  # 2^3 * 3 * 5: the per-pass-mapped schedule (radices 8, 3, 5)
    (120, 8, (1, 2, 4), "c64"),
    (120, 8, (1, 1, 8), "c64"),    # z blocks of 15: the first pass's inputs step across rank blocks
    (120, 15, (1, 1, 15), "c128"),  # z blocks of 8: the last pass's outputs step across rank blocks
    (12, 6, None, "c128"),
    (12, 6, None, "c64"),
    (24, 12, None, "c128"),
    (60, 4, (1, 2, 2), "c64"),
    (8, 8, None, "c128"),
    (32, 8, None, "c128"),
    (4, 4, None, "c128"),
    (32, 8, (1, 2, 4), "c128"),
    (64, 16, (1, 4, 4), "c64"),
    (64, 16, (4, 2, 2), "c64"),
    (128, 64, (1, 8, 8), "c64"),
    (32, 64, (4, 4, 4), "c128"),
]
TOL_3D = {"c128": 2e-15, "c64": 6e-7}


def _tol3d(Ng, dtype):
    """Stockham lengths (2^a 3^b 5^c): TOL_3D; direct-DFT lengths round like sqrt(Ng)."""
    return TOL_3D[dtype] * (1.0 if _smooth(Ng) else max(1.0, np.sqrt(Ng) / 2))


def _blocks(d, dtype, seed):
    g = O.fill_uniform(seed, d.Ng ** 3).reshape((d.Ng,) * 3)
    b = O.global_to_blocks(d, g)
    return b.astype(np.complex128 if dtype == "c128" else np.complex64)


@pytest.mark.parametrize("cfg", CONFIGS_3D, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}_{c[3]}")
def test_3d_forward_inverse_vs_oracle(a2afft_mod, cuda, cfg):
    import torch
    Ng, P, dims, dtype = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    x = _blocks(d, dtype, seed=Ng + P)
    ref_f = O.forward(d, x.astype(np.complex128))
    dist = a2afft_mod.Distribution(Ng, P, dims=d.dims_t, dtype=dtype)
    outs = {}
    for ex in ("staged", "fused"):
        plan = a2afft_mod.Dfft(dist, exec=ex)
        scratch = dist.empty()
        plan.make_plans(scratch)
        data = torch.from_numpy(x).cuda()
        plan.forward(data)
        torch.cuda.synchronize()
        y = data.cpu().numpy()
        assert _relerr(y, ref_f) < _tol3d(Ng, dtype), ex
        outs[ex] = y
        plan.inverse(data)
        torch.cuda.synchronize()
        z = data.cpu().numpy()
        assert _relerr(z, x.astype(np.complex128)) < _tol3d(Ng, dtype) * 2, ex
        outs[ex + "_inv"] = z
    assert np.array_equal(_bits(outs["staged"]), _bits(outs["fused"])), "fused != staged (forward)"
    assert np.array_equal(_bits(outs["staged_inv"]), _bits(outs["fused_inv"])), "fused != staged (inverse)"


def test_3d_matches_numpy_fftn(a2afft_mod, cuda):
    """The reference's own pin: np.fft.fftn of the stitched global grid (tests/stitch.py:89)."""
    import torch
    from a2afft import inputs
    Ng, dims = 32, (2, 2, 2)
    g = inputs.uniform(5, Ng ** 3).reshape((Ng,) * 3)
    dist = a2afft_mod.Distribution(Ng, 8, dims=dims, dtype="c128")
    plan = a2afft_mod.Dfft(dist)
    plan.make_plans(dist.empty())
    data = torch.from_numpy(inputs.global_to_blocks(g, dims)).cuda()
    plan.forward(data)
    out = inputs.blocks_to_global(data.cpu().numpy(), dims, g.shape)
    assert _relerr(out, np.fft.fftn(g)) < 1e-15
    plan.inverse(data)
    inv = inputs.blocks_to_global(data.cpu().numpy(), dims, g.shape)
    assert _relerr(inv, g) < 1e-15


@pytest.mark.parametrize("cfg", [(64, 16, (1, 4, 4), "c64"), (128, 16, (1, 4, 4), "c64"), (64, 8, (1, 2, 4), "c128")],
                         ids=lambda c: f"Ng{c[0]}_{c[2]}_{c[3]}")
def test_chunked_and_tuned_plans_bit_identical(a2afft_mod, cuda, cfg):
    """Chunk widths and kernel variants change launch order and shape only:
    forward and inverse must equal the default plan bit for bit, and the staged
    reference-order plan."""
    import torch
    Ng, P, dims, dtype = cfg
    rc, d = O.decompose(Ng, P, dims)
    x = _blocks(d, dtype, seed=3)
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype=dtype)
    base = a2afft_mod.Dfft(dist)
    base.make_plans(None)
    ref = torch.from_numpy(x).cuda()
    base.forward(ref)
    staged = a2afft_mod.Dfft(dist, exec="staged")
    staged.make_plans(dist.empty())
    st = torch.from_numpy(x).cuda()
    staged.forward(st)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(ref.cpu().numpy()), _bits(st.cpu().numpy()))
    tuned = a2afft_mod.Dfft(dist)
    tuned.make_plans(None)
    tuned.autotune(dist.empty().copy_(torch.from_numpy(x)), iters=1)
    for chunk in (0, 16, 32, Ng // 2):
        try:
            tuned.set_chunking(chunk)
        except a2afft_mod.A2AFFTError:
            continue
        y = torch.from_numpy(x).cuda()
        tuned.forward(y)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref.cpu().numpy())), (chunk, tuned.describe())
        tuned.inverse(y)
        torch.cuda.synchronize()
        assert _relerr(y.cpu().numpy(), x.astype(np.complex128)) < TOL_3D[dtype] * 2
    # launch-by-launch execution equals forward
    y = torch.from_numpy(x).cuda()
    for i in range(tuned.num_launches()):
        tuned.execute_launch(y, i, -1)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref.cpu().numpy()))


@pytest.mark.parametrize("shape,prec", [((32, 64, 16), "fp64"), ((128, 256, 64), "fp32"), ((64, 32, 128), "fp64"),
                                        ((16, 512, 8), "fp32"), ((2, 4, 2048), "fp64"),
                                        ((1024, 8, 1024), "fp32"), ((1024, 8, 512), "fp64"),
                                        ((8, 1024, 64), "fp32"), ((8, 1024, 64), "fp64"),
                                        ((2048, 64, 64), "fp64"), ((8, 2048, 64), "fp64")],
                         ids=lambda v: str(v))
def test_grid_plan_non_cubic_vs_numpy(a2afft_mod, cuda, shape, prec):
    """Single-grid plans (no decomposition) of non-cubic shapes, the shapes of
    the batched configuration, against numpy's fftn in FP64."""
    import torch
    from a2afft import inputs
    n = int(np.prod(shape))
    g = inputs.uniform(11, n).reshape(shape)
    x = g.astype(np.complex128 if prec == "fp64" else np.complex64)
    plan = a2afft_mod.GridDfft(shape, prec)
    d = torch.from_numpy(x).cuda()
    plan.forward(d)
    torch.cuda.synchronize()
    ref = np.fft.fftn(x.astype(np.complex128))
    tol = 2e-15 if prec == "fp64" else 6e-7
    assert _relerr(d.cpu().numpy(), ref) < tol
    plan.inverse(d)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), x.astype(np.complex128)) < 2 * tol


def test_chirp_on_smooth_lengths(a2afft_mod, cuda):
    """The chirp-z pass also takes 2^a 3^b 5^c lengths when forced (a
    set_kernel candidate beside the runtime-schedule tiles): 120-point x
    passes, complex128 and complex64, against numpy's fftn."""
    import torch
    from a2afft import inputs
    shape = (120, 4, 64)
    x = inputs.uniform(19, int(np.prod(shape))).reshape(shape)
    ref = np.fft.fftn(x)
    for prec, dt in (("fp64", "c128"), ("fp32", "c64")):
        tol = TOL_1D[dt] * 3 * np.log2(120)
        for forced in (False, True):
            plan = a2afft_mod.GridDfft(shape, prec)
            if forced:
                plan.set_kernel(1, 1, 120, KF_CZ)
            assert _chosen_flags(plan, "x") == (KF_CZ if forced else KF_MIXL | 1)
            d = torch.from_numpy(x.astype(np.complex128 if prec == "fp64" else np.complex64)).cuda()
            plan.forward(d)
            torch.cuda.synchronize()
            assert _relerr(d.cpu().numpy(), ref) < tol, (prec, forced)


# (grid shape, precision, a tile the plan must choose): strided passes of long
# 2^a 3^b 5^c lines whose tiles need more than 256 threads (wide runtime-schedule
# kernels, up to 1024 threads and 128 KiB of LDS)
WIDE_MIX_CASES = [((384, 384, 16), "fp32", "x T=16 E=12 NT=512"), ((768, 32, 16), "fp32", "x T=16 E=12 NT=1024"),
                  ((768, 4, 16), "fp64", "x T=8 E=12 NT=512"), ((2560, 2, 16), "fp32", "x T=4 E=20 NT=512"),
                  ((1536, 2, 16), "fp32", "x T=8 E=12 NT=1024"), ((640, 8, 16), "fp32", "x T=16 E=20 NT=512"),
                  ((16, 768, 32), "fp32", "y T=16 E=12 NT=1024")]


@pytest.mark.parametrize("case", WIDE_MIX_CASES, ids=lambda c: f"{c[0]}_{c[1]}")
def test_wide_mixed_tiles_vs_numpy(a2afft_mod, cuda, case):
    """Wide strided tiles of the runtime-schedule Stockham kernel: the plan
    picks them (set by hand for lengths that default to the per-pass-mapped
    kernel: multiples of 20, and 1536 in complex64), and forward and
    inverse match numpy's fftn in FP64 to the precision's bound."""
    import torch
    from a2afft import inputs
    shape, prec, chosen = case
    n = int(np.prod(shape))
    x = inputs.uniform(13, n).reshape(shape).astype(np.complex128 if prec == "fp64" else np.complex64)
    plan = a2afft_mod.GridDfft(shape, prec)
    if not _chosen_flags(plan, chosen[0]) & KF_MIX:   # per-pass-mapped by default: force the tile
        T, E = (int(f.split("=")[1]) for f in chosen.split()[1:3])
        plan.set_kernel("zxy".index(chosen[0]), T, E, KF_MIX)
    assert chosen in plan.describe(), plan.describe()
    d = torch.from_numpy(x).cuda()
    plan.forward(d)
    torch.cuda.synchronize()
    ref = np.fft.fftn(x.astype(np.complex128))
    lg = np.log2(max(shape))
    tol = 1e-15 * lg if prec == "fp64" else 1.5e-7 * lg
    assert _relerr(d.cpu().numpy(), ref) < tol
    plan.inverse(d)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), x.astype(np.complex128)) < 2 * tol


@pytest.mark.parametrize("case", [((768, 32, 16), "fp32", 1, 12, (1, 2, 4, 8, 16)),
                                  ((768, 4, 16), "fp64", 1, 12, (1, 2, 4, 8)),
                                  ((16, 16, 768), "fp32", 0, 12, (1, 2, 4)),
                                  ((16, 960, 16), "fp32", 2, 30, (1, 2, 4, 8)),
                                  ((120, 4, 64), "fp32", 1, 30, (1, 2, 4, 8, 16, 32, 64)),
                                  ((4, 72, 64), "fp32", 2, 12, (1, 2, 4, 8, 16, 32))],
                         ids=lambda c: f"{c[0]}_{c[1]}_pass{c[2]}")
def test_mixed_tiles_bit_identical(a2afft_mod, cuda, case):
    """Every tile the runtime-schedule kernel accepts (narrow and wide; the
    autotune candidates) computes each line with the same arithmetic: forced on
    its pass, each reproduces the first tile's transform bit for bit."""
    import torch
    shape, prec, pass_, E, tiles = case
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    x = torch.empty(shape, dtype=dt, device="cuda")
    a2afft_mod.fill_random(x, 17, "uniform")
    base = a2afft_mod.GridDfft(shape, prec)
    base.set_kernel(pass_, tiles[0], E, KF_MIX)
    ref = x.clone()
    base.forward(ref)
    torch.cuda.synchronize()
    ref_bits = _bits(ref.cpu().numpy())
    for T in tiles[1:]:
        plan = a2afft_mod.GridDfft(shape, prec)
        plan.set_kernel(pass_, T, E, 128)
        assert f"T={T} E={E} " in plan.describe().splitlines()[pass_], plan.describe()
        y = x.clone()
        plan.forward(y)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(y.cpu().numpy()), ref_bits), (T, E)


# (grid shape, precision, pass): lengths with both 3 and 5 as factors, which
# run the per-pass-mapped Stockham kernel by default, contiguous and strided
# (3000-point complex64 columns two to a tile, complex128 one)
MIXL_CASES = [((16, 16, 480), "fp32", 0), ((480, 8, 64), "fp32", 1), ((8, 360, 64), "fp64", 2),
              ((4, 4, 3000), "fp32", 0), ((3000, 2, 16), "fp64", 1), ((3000, 2, 16), "fp32", 1),
              ((16, 16, 120), "fp64", 0),
              ((120, 4, 64), "fp32", 1),
              ((1536, 2, 16), "fp32", 1), ((16, 16, 3072), "fp32", 0), ((8, 1536, 16), "fp64", 2),
              ((16, 16, 720), "fp32", 0), ((720, 4, 16), "fp32", 1), ((8, 900, 16), "fp64", 2),
              ((3600, 2, 8), "fp32", 1), ((4, 4, 3600), "fp64", 0)]


@pytest.mark.parametrize("case", MIXL_CASES, ids=lambda c: f"{c[0]}_{c[1]}_pass{c[2]}")
def test_mixl_tiles_vs_numpy_and_bit_identical(a2afft_mod, cuda, case):
    """The per-pass-mapped kernel: the plan picks it, forward and inverse match
    numpy's fftn in FP64 to the precision's bound, and every tile and flag
    variant it accepts on the pass (the autotune candidates) reproduces the
    default plan bit for bit."""
    import torch
    from a2afft import inputs
    shape, prec, pass_ = case
    n = int(np.prod(shape))
    x = inputs.uniform(23, n).reshape(shape).astype(np.complex128 if prec == "fp64" else np.complex64)
    base = a2afft_mod.GridDfft(shape, prec)
    assert _chosen_flags(base, "zxy"[pass_]) & KF_MIXL, base.describe()
    d = torch.from_numpy(x).cuda()
    base.forward(d)
    torch.cuda.synchronize()
    ref_bits = _bits(d.cpu().numpy())
    lg = np.log2(max(shape))
    tol = 1e-15 * lg if prec == "fp64" else 1.5e-7 * lg
    assert _relerr(d.cpu().numpy(), np.fft.fftn(x.astype(np.complex128))) < tol
    base.inverse(d)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), x.astype(np.complex128)) < 2 * tol
    ran = 0
    for T in (1, 2, 3, 4, 8, 16):
        for E in (16, 24, 32, 48):
            for fl in (1, 5, 0):
                plan = a2afft_mod.GridDfft(shape, prec)
                try:
                    plan.set_kernel(pass_, T, E, KF_MIXL | fl)
                except a2afft_mod.A2AFFTError:
                    continue
                y = torch.from_numpy(x).cuda()
                plan.forward(y)
                torch.cuda.synchronize()
                assert np.array_equal(_bits(y.cpu().numpy()), ref_bits), (T, E, fl)
                ran += 1
    assert ran >= 2, ran   # 3600-point complex128 lines: one line a tile, two flag variants


# (grid shape, precision, pass): multiples of 20 that stay on the runtime
# schedule (4000 complex64, 1280 and 2000 complex128)
RUNTIME_20_CASES = [((4000, 2, 16), "fp32", 1), ((2000, 2, 16), "fp64", 1), ((16, 8, 1280), "fp64", 0),
                    ((4, 4000, 16), "fp32", 2)]


@pytest.mark.parametrize("case", RUNTIME_20_CASES, ids=lambda c: f"{c[0]}_{c[1]}_pass{c[2]}")
def test_runtime_schedule_multiples_of_20(a2afft_mod, cuda, case):
    """Multiples of 20 whose strided passes run faster on the runtime schedule:
    the plan picks that family, and forward and inverse match numpy's fftn in
    FP64 to the precision's bound."""
    import torch
    from a2afft import inputs
    shape, prec, pass_ = case
    n = int(np.prod(shape))
    x = inputs.uniform(29, n).reshape(shape).astype(np.complex128 if prec == "fp64" else np.complex64)
    plan = a2afft_mod.GridDfft(shape, prec)
    assert _chosen_flags(plan, "zxy"[pass_]) & KF_MIX, plan.describe()
    d = torch.from_numpy(x).cuda()
    plan.forward(d)
    torch.cuda.synchronize()
    lg = np.log2(max(shape))
    tol = 1e-15 * lg if prec == "fp64" else 1.5e-7 * lg
    assert _relerr(d.cpu().numpy(), np.fft.fftn(x.astype(np.complex128))) < tol
    plan.inverse(d)
    torch.cuda.synchronize()
    assert _relerr(d.cpu().numpy(), x.astype(np.complex128)) < 2 * tol


# (grid shape, precision, [(pass, T, E)]): the shapes built with half-width
# LDS exchanges (KF_HALFX) on every pass of a grid that has them
HALFX_CASES = [((256, 256, 256), "fp64", [(0, 8, 16), (0, 16, 16), (1, 8, 16), (1, 4, 16), (2, 8, 16)]),
               ((512, 16, 512), "fp64", [(0, 8, 16), (1, 8, 16)]),
               ((512, 16, 512), "fp32", [(0, 16, 16), (1, 16, 32)]),
               ((16, 16, 2048), "fp32", [(0, 2, 16)]),
               ((2048, 4, 32), "fp32", [(1, 8, 32), (1, 8, 64), (1, 16, 64)]),
               ((4, 2048, 32), "fp32", [(2, 16, 64)]),
               ((1024, 4, 64), "fp32", [(1, 16, 32), (1, 16, 64), (1, 32, 64)]),
               ((1024, 4, 64), "fp64", [(1, 8, 16), (1, 8, 32)]),
               ((2048, 4, 32), "fp64", [(1, 8, 32), (1, 4, 16), (1, 4, 32)])]
HALFX_FLAGS = (32 | 1, 32 | 1 | 2, 32 | 1 | 2 | 4, 32 | 8, 32 | 8 | 1)


@pytest.mark.parametrize("case", HALFX_CASES, ids=lambda c: f"{c[0]}_{c[1]}")
def test_halfx_variants_bit_identical(a2afft_mod, cuda, case):
    """Half-width LDS exchanges move the same values in two steps: every such
    variant, forced on its pass, reproduces the default plan bit for bit."""
    import torch
    shape, prec, shapes = case
    dt = torch.complex128 if prec == "fp64" else torch.complex64
    x = torch.empty(shape, dtype=dt, device="cuda")
    a2afft_mod.fill_random(x, 11, "uniform")
    base = a2afft_mod.GridDfft(shape, prec)
    ref = x.clone()
    base.forward(ref)
    torch.cuda.synchronize()
    ref_bits = _bits(ref.cpu().numpy())
    ran = 0
    for (p, T, E) in shapes:
        for flags in HALFX_FLAGS:
            plan = a2afft_mod.GridDfft(shape, prec)
            plan.set_kernel(p, T, E, flags)
            y = x.clone()
            plan.forward(y)
            torch.cuda.synchronize()
            assert np.array_equal(_bits(y.cpu().numpy()), ref_bits), (p, T, E, flags)
            ran += 1
    assert ran == len(shapes) * len(HALFX_FLAGS)


@pytest.mark.parametrize("case", [((512, 16, (1, 4, 4), "c64", "fp32"), 128), ((512, 16, (1, 4, 4), "c128", "fp64"), 64),
                                  ((256, 8, (1, 2, 4), "c128", "fp64"), 0), ((256, 8, (1, 2, 4), "c64", "fp32"), 0),
                                  ((512, 16, (1, 4, 4), "c64", "split_bf16"), 128),
                                  ((512, 16, (1, 4, 4), "c64", "split_fp16"), 128)],
                         ids=lambda c: str(c))
def test_default_chunk_width(a2afft_mod, cuda, case):
    """New plans alternate the x and y passes over windows of at most 256 MiB
    only for grids above 512 MiB on the 512-point half-width-exchange tiles
    or the interleaved-K split kernels; every other plan runs whole passes."""
    (Ng, P, dims, dtype, prec), width = case
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype=dtype)
    plan = a2afft_mod.Dfft(dist, precision=prec)
    assert f"chunk {width} z columns" in plan.describe()
    assert plan.num_launches() == (3 if width == 0 else 1 + 2 * (Ng // width))


# Lines longer than 4096 points: four-step over N = N1 * N2 (both <= 4096).
# Splits: 8192 = 128*64, 12288 = 128*96 (a runtime-schedule pass), 16384 =
# 128*128, 65536 = 256*256, 10000 = 100*100, 8194 = 241*34 (a 241-point
# factor, where chirp-z beats the direct DFT: the whole line runs as chirp-z),
# 2^20 = 1024*1024.
LONG = [8192, 12288, 16384, 65536, 10000, 8194, 1 << 20]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("N", LONG)
def test_long_lines_vs_oracle(a2afft_mod, cuda, N, dtype):
    import torch
    rng = np.random.default_rng(N + 1)
    # two lines for 8194: the oracle sums its 241-point factor directly, O(N^2) per line
    lines = 2 if N == 8194 else (1 << 18) // N + 1
    x = rng.standard_normal((lines, N)) + 1j * rng.standard_normal((lines, N))
    x = x.astype(np.complex128 if dtype == "c128" else np.complex64)
    # Stockham factors grow the error like log2 N; the chirp-z form is bounded by 3 log2 N
    scale = 3 * np.log2(N) if N == 8194 else np.log2(N)
    for direction, fn in ((-1, a2afft_mod.forward_1d_fft), (+1, a2afft_mod.inverse_1d_fft)):
        ref = O.fft_batched(x.astype(np.complex128), direction)
        d = torch.from_numpy(x).cuda()
        fn(d, N)
        torch.cuda.synchronize()
        assert _relerr(d.cpu().numpy(), ref) < TOL_1D[dtype] * scale, direction


def test_long_line_impulse(a2afft_mod, cuda):
    """delta(n - n0) -> exp(-2 pi i k n0 / N), phases reduced mod N exactly."""
    import torch
    N, n0 = 1 << 16, 12345
    x = np.zeros(N, np.complex128)
    x[n0] = 1.0
    d = torch.from_numpy(x).cuda()
    a2afft_mod.forward_1d_fft(d, N)
    torch.cuda.synchronize()
    k = np.arange(N, dtype=np.int64)
    ref = np.exp(-2j * np.pi * ((k * n0) % N) / N)
    assert np.max(np.abs(d.cpu().numpy() - ref)) < 1e-14


@pytest.mark.parametrize("N", [(1 << 24) + 1, 4099 * 4097])
def test_long_line_without_split_is_an_error(a2afft_mod, cuda, N):
    """Above 2^23 a length needs a split into two factors <= 4096; these have
    none (one is above 4096^2, the other has the prime factor 4099)."""
    import torch
    d = torch.zeros(N, dtype=torch.complex64, device="cuda")
    with pytest.raises(a2afft_mod.A2AFFTError, match="neither a product"):
        a2afft_mod.forward_1d_fft(d, N)


# Lines above 4096 with no two-factor split (a prime factor above 4096), and
# lines up to 4096 that only the direct DFT covers once N^2 >= 108 M: Bluestein's
# chirp-z convolution over power-of-two FFTs.  The error of two M-point FFTs
# (M < 4N) plus the chirp products: bounded here by 3 log2(N) times the
# per-length tolerance.
BLUESTEIN = [(4099, 2), (8191, 2), (12297, 1), (251, 8), (509, 4), (2039, 2), (4093, 2), (4095, 2),
             (8186, 1)]   # 8186 = 4093 * 2: a two-factor split whose 4093-point factor goes to chirp-z
_CHIRP_REF = {}


def _chirp_case(N, lines, direction):
    """Input (complex64 values, exact in complex128 too) and the oracle's
    transform of it, computed once per length and direction: the oracle sums
    these lengths directly, O(N^2) per line."""
    key = (N, lines, direction)
    if key not in _CHIRP_REF:
        rng = np.random.default_rng(N)
        x = (rng.standard_normal((lines, N)) + 1j * rng.standard_normal((lines, N))).astype(np.complex64)
        _CHIRP_REF[key] = (x, O.fft_batched(x.astype(np.complex128), direction))
    return _CHIRP_REF[key]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("case", BLUESTEIN, ids=lambda c: f"N{c[0]}")
def test_chirp_z_lines_vs_oracle(a2afft_mod, cuda, case, dtype):
    import torch
    N, lines = case
    for direction, fn in ((-1, a2afft_mod.forward_1d_fft), (+1, a2afft_mod.inverse_1d_fft)):
        x, ref = _chirp_case(N, lines, direction)
        x = x.astype(np.complex128 if dtype == "c128" else np.complex64)
        d = torch.from_numpy(x).cuda()
        fn(d, N)
        torch.cuda.synchronize()
        assert _relerr(d.cpu().numpy(), ref) < TOL_1D[dtype] * 3 * np.log2(N), direction


def test_chirp_z_impulse(a2afft_mod, cuda):
    """delta(n - n0) -> exp(-2 pi i k n0 / N) at the prime N = 8191."""
    import torch
    N, n0 = 8191, 4321
    x = np.zeros(N, np.complex128)
    x[n0] = 1.0
    d = torch.from_numpy(x).cuda()
    a2afft_mod.forward_1d_fft(d, N)
    torch.cuda.synchronize()
    k = np.arange(N, dtype=np.int64)
    ref = np.exp(-2j * np.pi * ((k * n0) % N) / N)
    assert np.max(np.abs(d.cpu().numpy() - ref)) < 1e-13


def test_long_lines_on_two_streams(a2afft_mod, cuda):
    """Two rectangular-split transforms (8192 = 128 * 64, each with its own
    stream-ordered workspace) issued on two streams without synchronizing in
    between: each equals the same transform run alone."""
    import torch
    N = 8192
    rng = np.random.default_rng(77)
    xa = (rng.standard_normal((64, N)) + 1j * rng.standard_normal((64, N))).astype(np.complex64)
    xb = (rng.standard_normal((96, N)) + 1j * rng.standard_normal((96, N))).astype(np.complex64)
    ref_a = torch.from_numpy(xa).cuda()
    ref_b = torch.from_numpy(xb).cuda()
    a2afft_mod.forward_1d_fft(ref_a, N)
    a2afft_mod.forward_1d_fft(ref_b, N)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        a = torch.from_numpy(xa).cuda()
        b = torch.from_numpy(xb).cuda()
        torch.cuda.synchronize()
        with torch.cuda.stream(s1):
            a2afft_mod.forward_1d_fft(a, N)
        with torch.cuda.stream(s2):
            a2afft_mod.forward_1d_fft(b, N)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(a.cpu().numpy()), _bits(ref_a.cpu().numpy())), rep
        assert np.array_equal(_bits(b.cpu().numpy()), _bits(ref_b.cpu().numpy())), rep


@pytest.mark.parametrize("kernel", ["default", "dft"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("case", [((3, 5, 4095), 0, 2), ((4093, 2, 4), 1, 0), ((6, 2039, 4), 2, 1),
                                  ((4, 6, 509), 0, 2), ((251, 3, 8), 1, 0)],
                         ids=lambda c: f"{c[0]}_pass{c[1]}")
def test_prime_lengths_in_plan_passes(a2afft_mod, cuda, case, prec, kernel):
    """Plan passes over lengths no Stockham kernel covers, contiguous (z) and
    strided (x, y): by default the chirp-z pass (lengths where it beats the
    direct DFT, all of these), or the direct DFT forced through set_kernel;
    one pass against the oracle's FFT along that axis."""
    import torch
    shape, pass_, axis = case
    N = shape[axis]
    rng = np.random.default_rng(N + pass_)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    dt = np.complex128 if prec == "fp64" else np.complex64
    x = x.astype(dt)
    plan = a2afft_mod.GridDfft(shape, prec)
    name = "zxy"[pass_]
    if kernel == "dft":
        plan.set_kernel(pass_, 1, N, KF_DFT)
        assert _chosen_flags(plan, name) == KF_DFT
        tol = TOL_1D["c128" if prec == "fp64" else "c64"] * np.sqrt(N)
    else:
        assert _chosen_flags(plan, name) == KF_CZ
        tol = TOL_1D["c128" if prec == "fp64" else "c64"] * 3 * np.log2(N)
    d = torch.from_numpy(x).cuda()
    a2afft_mod.execute_pass(plan, d, pass_, -1)
    torch.cuda.synchronize()
    ref = np.moveaxis(O.fft_batched(np.moveaxis(x.astype(np.complex128), axis, -1), -1), -1, axis)
    assert _relerr(d.cpu().numpy(), ref) < tol
    # inverse of the same pass: back to N times the input (the y pass, last
    # of an inverse transform, also divides by the grid size)
    a2afft_mod.execute_pass(plan, d, pass_, 1)
    torch.cuda.synchronize()
    gain = N / (np.prod(shape) if pass_ == 2 else 1)
    assert _relerr(d.cpu().numpy() / gain, x.astype(np.complex128)) < 2 * tol


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_chirp_3d_fused_staged_numpy(a2afft_mod, cuda, dtype):
    """254^3 (254 = 2 * 127) over dims [1,1,2]: every axis a chirp-z pass.
    Fused and staged (reference stage order, batched chirp-z pencils) agree
    bit for bit, forward and inverse; the forward matches numpy's fftn of the
    stitched grid; a captured graph replays the eager bits."""
    import torch
    from a2afft import inputs
    Ng, dims = 254, (1, 1, 2)
    g = inputs.uniform(11, Ng ** 3).reshape((Ng,) * 3)
    x = inputs.global_to_blocks(g, dims).astype(np.complex128 if dtype == "c128" else np.complex64)
    dist = a2afft_mod.Distribution(Ng, 2, dims=dims, dtype=dtype)
    outs = {}
    for ex in ("staged", "fused"):
        plan = a2afft_mod.Dfft(dist, exec=ex)
        plan.make_plans(dist.empty())
        if ex == "fused":
            for name in "zxy":
                assert _chosen_flags(plan, name) == KF_CZ
        data = torch.from_numpy(x).cuda()
        plan.forward(data)
        torch.cuda.synchronize()
        outs[ex] = data.cpu().numpy()
        plan.inverse(data)
        torch.cuda.synchronize()
        outs[ex + "_inv"] = data.cpu().numpy()
        if ex == "fused":
            for rep in range(3):
                y = torch.from_numpy(x).cuda()
                plan.graph_launch(y, -1)
                torch.cuda.synchronize()
                assert np.array_equal(_bits(y.cpu().numpy()), _bits(outs["fused"])), rep
                z = torch.from_numpy(x).cuda()
                plan.forward(z)   # eager work between replays
                torch.cuda.synchronize()
    assert np.array_equal(_bits(outs["staged"]), _bits(outs["fused"])), "fused != staged (forward)"
    assert np.array_equal(_bits(outs["staged_inv"]), _bits(outs["fused_inv"])), "fused != staged (inverse)"
    tol = TOL_1D[dtype] * 3 * np.log2(Ng) * 2
    ref = np.fft.fftn(g if dtype == "c128" else g.astype(np.complex64).astype(np.complex128))
    out = inputs.blocks_to_global(outs["fused"], dims, g.shape)
    assert _relerr(out, ref) < tol
    inv = inputs.blocks_to_global(outs["fused_inv"], dims, g.shape)
    assert _relerr(inv, inputs.blocks_to_global(x, dims, g.shape).astype(np.complex128)) < 2 * tol


def test_chirp_pass_windows_match_whole(a2afft_mod, cuda):
    """A 509-point strided x pass (chirp-z) over 16-column z windows, and
    every direct-DFT pass forced, against whole passes: same bits per
    kernel choice; chirp-z and direct DFT within rounding of each other."""
    import torch
    shape = (509, 16, 64)
    x = torch.empty(shape, dtype=torch.complex64, device="cuda")
    a2afft_mod.fill_random(x, 29, "uniform")
    whole = a2afft_mod.GridDfft(shape, "fp32")
    win = a2afft_mod.GridDfft(shape, "fp32")
    win.set_chunking(16)
    assert _chosen_flags(whole, "x") == KF_CZ
    a, b = x.clone(), x.clone()
    whole.forward(a)
    win.forward(b)
    assert torch.equal(torch.view_as_real(a), torch.view_as_real(b))
    whole.inverse(a)
    win.inverse(b)
    assert torch.equal(torch.view_as_real(a), torch.view_as_real(b))
    dft = a2afft_mod.GridDfft(shape, "fp32")
    dft.set_kernel(1, 1, 509, KF_DFT)
    c, e = x.clone(), x.clone()
    dft.forward(c)
    whole.forward(e)
    torch.cuda.synchronize()
    ref = np.fft.fftn(x.cpu().numpy().astype(np.complex128))
    assert _relerr(c.cpu().numpy(), ref) < TOL_1D["c64"] * np.sqrt(509) * 2
    assert _relerr(e.cpu().numpy(), ref) < TOL_1D["c64"] * 3 * np.log2(509) * 2


def test_many_windows_match_whole_passes(a2afft_mod, cuda):
    """16 x 16 x 4096 complex128 in 8-column windows: 512 windows, 1025
    launches per transform, every window's x and y passes run (bit-identical
    to whole passes)."""
    import torch
    shape = (16, 16, 4096)
    x = torch.empty(shape, dtype=torch.complex128, device="cuda")
    a2afft_mod.fill_random(x, 23, "uniform")
    whole = a2afft_mod.GridDfft(shape, "fp64")
    win = a2afft_mod.GridDfft(shape, "fp64")
    win.set_chunking(8)
    assert win.num_launches() == 1 + 2 * 512
    a, b = x.clone(), x.clone()
    whole.forward(a)
    win.forward(b)
    assert torch.equal(torch.view_as_real(a), torch.view_as_real(b))
