"""GPU tests of the split-precision passes (radix-16 DFT stages as matrix
products on MFMA with hi + lo operand halves, FP32 accumulation).

The bar is the error against a complex128 reference of the same transform:
rel-L2 <= 1e-5 for the BF16 split (hi*hi + hi*lo + lo*hi keeps about 16 of
FP32's 24 mantissa bits per product) and <= 1e-6 for the FP16 split (about
22 bits, with a power-of-two scale per butterfly for FP16's range).  A wrong
lane layout, twiddle or output position gives O(1) errors, so these bounds
also pin the index arithmetic."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = {"split_bf16": 1e-5, "split_fp16": 1e-6}
GRIDS = [(64, 64, 64), (32, 32, 32), (128, 32, 256), (512, 64, 32), (16, 256, 64), (32, 128, 64),
         (64, 32, 512), (32, 512, 64), (16, 64, 512), (16, 16, 2048), (32, 16, 1024)]


def _rand(torch, shape, seed):
    g = np.random.default_rng(seed)
    x = (g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape)).astype(np.complex64)
    return x, torch.from_numpy(x.reshape(-1).copy()).cuda()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("shape", GRIDS)
def test_split_grid_forward_inverse_vs_numpy(a2afft_mod, cuda, prec, shape):
    import torch
    x, d = _rand(torch, shape, 11)
    plan = a2afft_mod.GridDfft(shape, prec)
    plan.forward(d)
    got = d.cpu().numpy().reshape(shape)
    ref = np.fft.fftn(x.astype(np.complex128))
    err = _rel(got, ref)
    assert err <= TOL[prec], (prec, shape, err)
    plan.inverse(d)
    back = d.cpu().numpy().reshape(shape)
    err_i = _rel(back, x.astype(np.complex128))
    assert err_i <= 2 * TOL[prec], (prec, shape, err_i)


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_split_impulses_give_exact_phases(a2afft_mod, cuda, prec):
    """A unit impulse at (a, b, c) transforms to exp(-2 pi i (ka/nx + lb/ny + mc/nz));
    a permuted lane layout or output index moves the phases."""
    import torch
    shape = (32, 64, 128)
    for pos in [(0, 0, 0), (1, 0, 0), (0, 3, 0), (0, 0, 5), (7, 11, 13)]:
        x = np.zeros(shape, np.complex64)
        x[pos] = 1.0
        d = torch.from_numpy(x.reshape(-1).copy()).cuda()
        a2afft_mod.GridDfft(shape, prec).forward(d)
        got = d.cpu().numpy().reshape(shape)
        k = [np.arange(n) for n in shape]
        ph = (np.exp(-2j * np.pi * k[0] * pos[0] / shape[0])[:, None, None]
              * np.exp(-2j * np.pi * k[1] * pos[1] / shape[1])[None, :, None]
              * np.exp(-2j * np.pi * k[2] * pos[2] / shape[2])[None, None, :])
        assert np.max(np.abs(got - ph)) < 30 * TOL[prec], (prec, pos)


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_split_distribution_plan_matches_fftn(a2afft_mod, cuda, prec):
    """Rank-decomposed block layout (64^3 over dims [1,2,4]): stitched output
    equals fftn of the stitched input."""
    import torch
    dist = a2afft_mod.Distribution(64, 8, dims=(1, 2, 4), dtype="c64")
    buf = dist.empty()
    a2afft_mod.fill_random(buf, 5, "uniform")
    x = buf.cpu().numpy()
    plan = a2afft_mod.Dfft(dist, prec)
    plan.forward(buf)
    from a2afft import inputs
    glob_in = inputs.blocks_to_global(x, dist.dims, (64, 64, 64))
    glob_out = inputs.blocks_to_global(buf.cpu().numpy(), dist.dims, (64, 64, 64))
    err = _rel(glob_out, np.fft.fftn(glob_in.astype(np.complex128)))
    assert err <= TOL[prec], (prec, err)


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_split_close_to_fp32_on_wide_dynamic_range(a2afft_mod, cuda, prec):
    """Inputs spanning 2^-20 .. 2^20: the FP16 split's per-butterfly scaling
    keeps the error relative to the transform's norm."""
    import torch
    shape = (64, 32, 128)
    g = np.random.default_rng(3)
    mag = 2.0 ** g.uniform(-20, 20, shape)
    x = (mag * np.exp(2j * np.pi * g.uniform(0, 1, shape))).astype(np.complex64)
    d = torch.from_numpy(x.reshape(-1).copy()).cuda()
    a2afft_mod.GridDfft(shape, prec).forward(d)
    got = d.cpu().numpy().reshape(shape)
    err = _rel(got, np.fft.fftn(x.astype(np.complex128)))
    assert err <= TOL[prec], (prec, err)


def test_split_rejects_staged_and_short_lines(a2afft_mod, cuda):
    from a2afft import _lib as L
    dist = a2afft_mod.Distribution(32, 8, dtype="c64")
    with pytest.raises(L.A2AFFTError):
        a2afft_mod.Dfft(dist, "split_bf16", exec="staged")
    with pytest.raises(L.A2AFFTError):
        a2afft_mod.GridDfft((8, 64, 64), "split_bf16")


# interleaved-K split-BF16 variants (T=16, E=0): contiguous and strided flags
_MF2_FLAGS = {0: [5, 4, 1], 1: [5, 4, 12, 13, 1, 8, 9, 261, 269, 264, 257], 2: [5, 4, 12, 13, 1, 261, 269]}


def test_split_bf16_variants_same_bits(a2afft_mod, cuda):
    """Every tile order and cache policy of the interleaved-K kernel computes
    the same bits on every pass; the split-K-halves kernel (E = N) stays
    within the BF16 bound of FP64."""
    import torch
    shape = (256, 512, 256)
    x, d0 = _rand(torch, shape, 21)
    ref = np.fft.fftn(x.astype(np.complex128))
    base = a2afft_mod.GridDfft(shape, "split_bf16")
    want = d0.clone()
    base.forward(want)
    assert _rel(want.cpu().numpy().reshape(shape), ref) <= TOL["split_bf16"]
    for p in range(3):
        for fl in _MF2_FLAGS[p]:
            plan = a2afft_mod.GridDfft(shape, "split_bf16")
            plan.set_kernel(p, 16, 0, fl)
            d = d0.clone()
            plan.forward(d)
            assert torch.equal(torch.view_as_real(d), torch.view_as_real(want)), (p, fl)
        n = shape[[2, 0, 1][p]]
        plan = a2afft_mod.GridDfft(shape, "split_bf16")
        plan.set_kernel(p, 16, n, 0)
        d = d0.clone()
        plan.forward(d)
        err = _rel(d.cpu().numpy().reshape(shape), ref)
        assert err <= TOL["split_bf16"], (p, err)


def test_split_bf16_rank_segments_256(a2afft_mod, cuda):
    """256^3 complex64 over dims [1,2,4]: the y and z lines run across rank
    blocks; stitched output equals numpy's fftn within the BF16 bound, and the
    inverse returns the input."""
    import torch
    from a2afft import inputs
    Ng, P, dims = 256, 8, (1, 2, 4)
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    buf = dist.empty()
    a2afft_mod.fill_random(buf, 9, "uniform")
    x = buf.cpu().numpy()
    plan = a2afft_mod.Dfft(dist, "split_bf16")
    plan.forward(buf)
    got = inputs.blocks_to_global(buf.cpu().numpy(), dims, (Ng,) * 3)
    ref = np.fft.fftn(inputs.blocks_to_global(x, dims, (Ng,) * 3).astype(np.complex128))
    assert _rel(got, ref) <= TOL["split_bf16"]
    del got, ref
    plan.inverse(buf)
    assert _rel(buf.cpu().numpy(), x.astype(np.complex128)) <= 2 * TOL["split_bf16"]


# interleaved-K split-FP16 variants (T=16, E=0)
_MF2_F16_FLAGS = {0: [5, 4], 1: [5, 4, 1, 12, 264, 268, 261], 2: [5, 4, 1, 12, 261]}


def test_split_fp16_variants_same_bits(a2afft_mod, cuda):
    """The interleaved-K kernel in FP16 (per-line power-of-two operand scale):
    every tile order and cache policy computes the same bits on every pass,
    within the FP16 bound of FP64; the split-K-halves kernel (E = N, one scale
    per butterfly) stays within the same bound."""
    import torch
    shape = (256, 512, 256)
    x, d0 = _rand(torch, shape, 23)
    ref = np.fft.fftn(x.astype(np.complex128))
    base = a2afft_mod.GridDfft(shape, "split_fp16")
    assert "interleaved-K" in base.describe()
    want = d0.clone()
    base.forward(want)
    assert _rel(want.cpu().numpy().reshape(shape), ref) <= TOL["split_fp16"]
    for p in range(3):
        for fl in _MF2_F16_FLAGS[p]:
            plan = a2afft_mod.GridDfft(shape, "split_fp16")
            plan.set_kernel(p, 16, 0, fl)
            d = d0.clone()
            plan.forward(d)
            assert torch.equal(torch.view_as_real(d), torch.view_as_real(want)), (p, fl)
        n = shape[[2, 0, 1][p]]
        plan = a2afft_mod.GridDfft(shape, "split_fp16")
        plan.set_kernel(p, 16, n, 0)
        d = d0.clone()
        plan.forward(d)
        err = _rel(d.cpu().numpy().reshape(shape), ref)
        assert err <= TOL["split_fp16"], (p, err)


def test_split_fp16_line_scales(a2afft_mod, cuda):
    """Lines of very different magnitude in one 16-line tile (each line scaled
    by 2^k, k from -60 to 60): every line's transform keeps its own relative
    accuracy, on a contiguous and both strided passes, forward and inverse."""
    import torch
    shape = (16, 16, 512)   # 512-point z lines; x and y columns of 16
    g = np.random.default_rng(5)
    x = (g.uniform(-1, 1, shape) + 1j * g.uniform(-1, 1, shape))
    k = g.integers(-60, 61, size=shape[:2])
    x = (x * (2.0 ** k)[:, :, None]).astype(np.complex64)
    d = torch.from_numpy(x.reshape(-1).copy()).cuda()
    plan = a2afft_mod.GridDfft(shape, "split_fp16")
    a2afft_mod.execute_pass(plan, d, 0, -1)   # z pass: the contiguous kernel
    got = d.cpu().numpy().reshape(shape)
    ref = np.fft.fft(x.astype(np.complex128), axis=2)
    err = np.linalg.norm(got - ref, axis=2) / np.linalg.norm(ref, axis=2)
    assert err.max() <= TOL["split_fp16"], err.max()
    # strided 512-point columns: the x pass of a 512 x 16 x 64 grid, each (y, z) line scaled by its own 2^k
    shape2 = (512, 16, 64)
    x2 = g.uniform(-1, 1, shape2) + 1j * g.uniform(-1, 1, shape2)
    k2 = g.integers(-60, 61, size=shape2[1:])
    x2 = (x2 * (2.0 ** k2)[None, :, :]).astype(np.complex64)
    d2 = torch.from_numpy(x2.reshape(-1).copy()).cuda()
    plan2 = a2afft_mod.GridDfft(shape2, "split_fp16")
    a2afft_mod.execute_pass(plan2, d2, 1, -1)
    got2 = d2.cpu().numpy().reshape(shape2)
    ref2 = np.fft.fft(x2.astype(np.complex128), axis=0)
    err2 = np.linalg.norm(got2 - ref2, axis=0) / np.linalg.norm(ref2, axis=0)
    assert err2.max() <= TOL["split_fp16"], err2.max()
    a2afft_mod.execute_pass(plan2, d2, 1, 1)
    back = d2.cpu().numpy().reshape(shape2)
    want = 512 * x2.astype(np.complex128)
    err3 = np.linalg.norm(back - want, axis=0) / np.linalg.norm(want, axis=0)
    assert err3.max() <= 2 * TOL["split_fp16"], err3.max()


@pytest.mark.parametrize("prec", ["split_bf16", "split_fp16"])
def test_split_autotune_and_graph_keep_bits(a2afft_mod, cuda, prec):
    """Autotune of a split plan stays within the interleaved-K family (every
    tile order, cache policy and occupancy computes the same bits), and graph
    replays of the tuned plan equal its eager forward, with eager work on the
    device between replays."""
    import torch
    shape = (128, 512, 256)
    x, d0 = _rand(torch, shape, 29)
    plan = a2afft_mod.GridDfft(shape, prec)
    want = d0.clone()
    plan.forward(want)
    work = d0.clone()
    plan.autotune(work, iters=1)
    assert "interleaved-K" in plan.describe().splitlines()[0]   # the 256-point z pass
    got = d0.clone()
    plan.forward(got)
    assert torch.equal(torch.view_as_real(got), torch.view_as_real(want))
    other = torch.empty_like(d0)
    for rep in range(3):
        got.copy_(d0)
        torch.cuda.synchronize()
        plan.graph_launch(got, -1)
        torch.cuda.synchronize()
        assert torch.equal(torch.view_as_real(got), torch.view_as_real(want)), rep
        other.copy_(d0)
        plan.forward(other)
        torch.cuda.synchronize()
    ref = np.fft.fftn(x.astype(np.complex128))
    assert _rel(want.cpu().numpy().reshape(shape), ref) <= TOL[prec]
