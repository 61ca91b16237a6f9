"""GPU tests at the full sizes of the benchmark workloads.

The oracle cannot transform these grids in a test's time, so the checks are
size-independent: the FP32 result against the library's FP64 result (whose
kernels the oracle pins at smaller sizes), and the FP64 result against closed
forms computed directly from the input -- Parseval's identity, the DC bin and
one off-axis bin as explicit sums in FP64, numpy's fftn where it runs in a few
seconds -- plus bit equality between strategies and launch layouts that must
compute the same bits, and round trips."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    import torch
    if a.dtype == torch.complex128:
        return bool(torch.equal(a.view(torch.float64), b.view(torch.float64)))
    return bool(torch.equal(a.view(torch.float32), b.view(torch.float32)))


def _norm2(t):
    import torch
    return float(torch.sum(torch.abs(t.to(torch.complex128)) ** 2))


def _bin(torch, g, k):
    """sum_n g[n] exp(-2 pi i k.n / N) in FP64, axis by axis (g: global grid, complex128)."""
    out = g
    for axis in (2, 1, 0):
        n = g.shape[axis]
        ph = torch.exp(torch.arange(n, device=g.device, dtype=torch.float64) * (-2j * math.pi * k[axis] / n))
        shape = [1] * out.dim()
        shape[-1] = n
        out = (out * ph.view(shape)).sum(-1)
    return complex(out.item())


def _grf_blocks(a2afft_mod, Ng, dims, seed):
    """The 512^3 GRF P(k) ~ k^-3 (+ 1e-3 noise) of the benchmark, rounded to
    complex64, in block layout over dims (the generator bench.py uses)."""
    import torch
    from a2afft import inputs
    g = inputs.gaussian_random_field_device((Ng,) * 3, seed)
    x = inputs.global_to_blocks(g.to(torch.complex64), dims).contiguous()
    del g
    return x


def _fp64_checked(a2afft_mod, x, Ng, P, dims):
    """FP64 forward of x (complex64 blocks) with Parseval's identity and two
    bins summed explicitly in FP64 checked on it; returns the complex128 result."""
    import torch
    from a2afft import inputs
    n3 = Ng ** 3
    x128 = x.to(torch.complex128)
    ref = x128.clone()
    dist64 = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c128")
    a2afft_mod.Dfft(dist64, "fp64").forward(ref)
    par = _norm2(ref) / (n3 * _norm2(x128))
    assert abs(par - 1.0) < 1e-12, par
    g = inputs.blocks_to_global(x128, dims, (Ng,) * 3)
    yg = inputs.blocks_to_global(ref, dims, (Ng,) * 3)
    scale = math.sqrt(_norm2(ref) / n3)   # RMS of the spectrum
    for k in ((0, 0, 0), (3, 200, 511), (1, 0, 0)):
        want = _bin(torch, g, k)
        got = complex(yg[k].item())
        assert abs(got - want) < 1e-10 * scale, (k, got, want)
    del g, yg, x128
    return ref


def test_c3_512cube_c64_dims144(a2afft_mod, cuda):
    """512^3 complex64 GRF (P(k) ~ k^-3, seed 2, + 1e-3 noise) over 16 virtual
    ranks, dims [1,4,4]: FP32 within 1e-6 of FP64 (rel-L2), the FP64 result
    against Parseval and explicit bins, fused == staged bit for bit, FP32
    round trip."""
    import torch
    Ng, P, dims = 512, 16, (1, 4, 4)
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    x = _grf_blocks(a2afft_mod, Ng, dims, 2).reshape(P, -1)
    ref = _fp64_checked(a2afft_mod, x, Ng, P, dims)
    y = x.clone()
    a2afft_mod.Dfft(dist, "fp32").forward(y)
    rel, mx = a2afft_mod.error_metrics(y, ref)
    print(f"c3 fp32 rel-L2 {rel:.3e} max-abs {mx:.3e}")
    assert rel <= 1e-6, rel
    del ref
    # the reference's stage order computes the same bits
    st = x.clone()
    staged = a2afft_mod.Dfft(dist, "fp32", exec="staged")
    staged.make_plans(dist.empty())
    staged.forward(st)
    assert _same_bits(st, y)
    del st, staged
    # round trip
    a2afft_mod.Dfft(dist, "fp32").inverse(y)
    rel, _ = a2afft_mod.error_metrics(y, x.to(torch.complex128))
    assert rel <= 1e-6, rel


def test_c4_512cube_precision_variants(a2afft_mod, cuda):
    """512^3 complex64 GRF (seed 3) over dims [1,4,4] through FP32 (<= 1e-6),
    split-BF16 (<= 1e-5) and split-FP16 (<= 1e-6), rel-L2 against the FP64
    result, which is itself checked against Parseval and explicit bins."""
    Ng, P, dims = 512, 16, (1, 4, 4)
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    x = _grf_blocks(a2afft_mod, Ng, dims, 3).reshape(P, -1)
    ref = _fp64_checked(a2afft_mod, x, Ng, P, dims)
    for prec, tol in (("fp32", 1e-6), ("split_bf16", 1e-5), ("split_fp16", 1e-6)):
        s = x.clone()
        a2afft_mod.Dfft(dist, prec).forward(s)
        r, mx = a2afft_mod.error_metrics(s, ref)
        print(f"c4 {prec} rel-L2 {r:.3e} max-abs {mx:.3e}")
        assert r <= tol, (prec, r)
        del s


def test_c3_uniform_split_variants(a2afft_mod, cuda):
    """512^3 complex64 U[-1,1) (seed 2) over dims [1,4,4]: the split variants
    within their bounds on a flat spectrum too."""
    import torch
    Ng, P, dims = 512, 16, (1, 4, 4)
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    x = dist.empty()
    a2afft_mod.fill_random(x, 2, "uniform")
    ref = x.to(torch.complex128)
    a2afft_mod.Dfft(a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c128"), "fp64").forward(ref)
    for prec, tol in (("fp32", 1e-6), ("split_bf16", 1e-5), ("split_fp16", 1e-6)):
        s = x.clone()
        a2afft_mod.Dfft(dist, prec).forward(s)
        r, _ = a2afft_mod.error_metrics(s, ref)
        assert r <= tol, (prec, r)
        del s


def test_grf_device_matches_host(a2afft_mod, cuda):
    """The device GRF generator (used by bench.py and the full-size tests)
    equals the host recipe (numpy Philox + scipy FFT) to FP64 rounding."""
    from a2afft import inputs
    for shape, seed in (((32, 32, 32), 2), ((64, 64, 64), 3), ((32, 64, 16), 2)):
        d = inputs.gaussian_random_field_device(shape, seed).cpu().numpy()
        h = inputs.gaussian_random_field(shape, seed)
        rel = np.linalg.norm(d - h) / np.linalg.norm(h)
        assert rel < 1e-13, (shape, rel)


@pytest.mark.parametrize("dims", [(1, 2, 4), (2, 2, 2)])
def test_c2_256cube_c128(a2afft_mod, cuda, dims):
    """256^3 complex128 over 8 virtual ranks: forward against numpy's fftn of
    the stitched grid (rel-L2 <= 1e-14), fused == staged bit for bit, inverse
    (with 1/N) back to the input."""
    import torch
    from a2afft import inputs
    Ng, P = 256, 8
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c128")
    x = dist.empty()
    a2afft_mod.fill_random(x, 1, "normal")
    y = x.clone()
    plan = a2afft_mod.Dfft(dist, "fp64")
    plan.forward(y)
    ref = np.fft.fftn(inputs.blocks_to_global(x.cpu().numpy(), dims, (Ng,) * 3))
    got = inputs.blocks_to_global(y.cpu().numpy(), dims, (Ng,) * 3)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert rel <= 1e-14, rel
    del ref, got
    st = x.clone()
    staged = a2afft_mod.Dfft(dist, "fp64", exec="staged")
    staged.make_plans(dist.empty())
    staged.forward(st)
    assert _same_bits(st, y)
    plan.inverse(y)
    rel, _ = a2afft_mod.error_metrics(y, x)
    assert rel <= 1e-14, rel


def test_c3b_2048_rank_share(a2afft_mod, cuda):
    """One rank's 2048x256x256 complex64 share of 2048^3 over 64 ranks, dims
    [1,8,8]: the batched 2048-point FFT on the pencils the exchange delivers,
    against the FP64 kernels (rel-L2 <= 1e-6) and numpy on sampled pencils;
    the strided x pass on the block gives the same bits as the pencil FFT."""
    import torch
    shape, P, dims = (2048, 256, 256), 64, (1, 8, 8)
    nloc = shape[0] * shape[1] * shape[2]
    block = torch.empty(nloc, dtype=torch.complex64, device="cuda")
    a2afft_mod.fill_random(block, 2, "uniform")
    pencils = torch.empty_like(block)
    a2afft_mod.launch_remap(6, block, pencils, P, nloc, shape, dims, nbuffers=1)   # [lz][ly][lx]
    ref = pencils.to(torch.complex128)
    a2afft_mod.forward_1d_fft(ref, 2048)
    y = pencils.clone()
    a2afft_mod.forward_1d_fft(y, 2048)
    rel, _ = a2afft_mod.error_metrics(y, ref)
    assert rel <= 1e-6, rel
    rows = np.random.default_rng(0).choice(nloc // 2048, 64, replace=False)
    src = pencils.view(-1, 2048)[torch.from_numpy(rows).cuda()].cpu().numpy().astype(np.complex128)
    got = ref.view(-1, 2048)[torch.from_numpy(rows).cuda()].cpu().numpy()
    want = np.fft.fft(src, axis=1)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-14
    del ref
    # the fused x pass on the block layout, transposed the same way, is the pencil FFT
    plan = a2afft_mod.GridDfft(shape, "fp32")
    wb = block.clone()
    a2afft_mod.execute_pass(plan, wb, 1, -1)
    back = torch.empty_like(wb)
    a2afft_mod.launch_remap(6, wb, back, P, nloc, shape, dims, nbuffers=1)
    assert _same_bits(back, y)


@pytest.mark.parametrize("prec,tol", [("split_bf16", 1e-5), ("split_fp16", 1e-6)])
def test_c3b_2048_rank_share_split(a2afft_mod, cuda, prec, tol):
    """The 2048x256x256 rank share's pencil FFT (65536 contiguous lines of
    2048 points, the z pass of a 256x256x2048 grid plan) in split-BF16 and
    split-FP16, on the GRF share (seed 2): within the precision's bound of the
    FP64 line FFTs, and the two radix-16 MFMA stages plus the radix-8 stage
    pinned by numpy on sampled pencils."""
    import torch
    from a2afft import inputs
    shape, P, dims = (2048, 256, 256), 64, (1, 8, 8)
    nloc = shape[0] * shape[1] * shape[2]
    g = inputs.gaussian_random_field_device(shape, 2)
    block = g.to(torch.complex64).reshape(-1)
    del g
    pencils = torch.empty_like(block)
    a2afft_mod.launch_remap(6, block, pencils, P, nloc, shape, dims, nbuffers=1)   # [lz][ly][lx]
    del block
    ref = pencils.to(torch.complex128)
    a2afft_mod.forward_1d_fft(ref, 2048)
    plan = a2afft_mod.GridDfft((256, 256, 2048), prec)
    assert "interleaved-K" in plan.describe().splitlines()[0]
    y = pencils.clone()
    a2afft_mod.execute_pass(plan, y, 0, -1)
    rel, mx = a2afft_mod.error_metrics(y, ref)
    print(f"c3b {prec} pencils rel-L2 {rel:.3e} max-abs {mx:.3e}")
    assert rel <= tol, rel
    rows = np.random.default_rng(1).choice(nloc // 2048, 32, replace=False)
    idx = torch.from_numpy(rows).cuda()
    want = np.fft.fft(pencils.view(-1, 2048)[idx].cpu().numpy().astype(np.complex128), axis=1)
    got = y.view(-1, 2048)[idx].cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= tol
    # inverse pass back to the input (1/N folded in by the caller here: none)
    a2afft_mod.execute_pass(plan, y, 0, 1)
    rel_rt, _ = a2afft_mod.error_metrics(y, pencils.to(torch.complex128) * 2048)
    assert rel_rt <= 2 * tol, rel_rt


def _c5_shapes():
    from a2afft import inputs
    cand = [(32, 32, 32), (64, 64, 64), (128, 128, 128), (256, 256, 256), (128, 256, 64), (256, 64, 512)]
    u = (inputs.uniform(4, 64).real + 1.0) / 2.0
    return [cand[min(5, int(v * 6))] for v in u]


def test_natural_512cube_grid(a2afft_mod, cuda):
    """The plain-grid plan on a natural-order 512^3 complex64 grid (its own
    tiles: 32-column x pass, single-line z pass) against rocFFT's FP64 fftn
    on the same GPU (rel-L2 <= 1e-6), and bit for bit against the same plan
    on the rank-block layout's tiles (16-column x pass, 16-line z pass)."""
    import torch
    sh = (512, 512, 512)
    x = torch.empty(sh, dtype=torch.complex64, device="cuda")
    a2afft_mod.fill_random(x, 29, "uniform")
    plan = a2afft_mod.GridDfft(sh, "fp32")
    desc = plan.describe()
    assert "chosen x T=32 E=32" in desc and "chosen z T=1 E=8" in desc, desc
    y = x.clone()
    plan.forward(y)
    alt = a2afft_mod.GridDfft(sh, "fp32")
    alt.set_kernel(0, 16, 16, 7)
    alt.set_kernel(1, 16, 32, 40)
    z = x.clone()
    alt.forward(z)
    torch.cuda.synchronize()
    assert _same_bits(y, z)
    del z
    ref = torch.fft.fftn(x.to(torch.complex128))
    err = float(torch.linalg.vector_norm(y.to(torch.complex128) - ref) / torch.linalg.vector_norm(ref))
    assert err < 1e-6, err
    plan.inverse(y)
    torch.cuda.synchronize()
    back = float(torch.linalg.vector_norm((y - x).to(torch.complex128)) / torch.linalg.vector_norm(x.to(torch.complex128)))
    assert back < 1e-6, back


@pytest.mark.parametrize("shape,prec,tol", [((2048, 2048, 2048), "fp32", 1e-5), ((2048, 2048, 1024), "fp64", 1e-12)],
                         ids=["2048cube_c64", "2048x2048x1024_c128"])
def test_whole_grid_of_2e32_points_and_more(a2afft_mod, cuda, shape, prec, tol):
    """A whole grid of 2^32 points or more on one device (64 GiB each; the
    passes address it in 64 bits).  A unit impulse at r0 transforms to
    exp(-2 pi i k.r0 / n) at every k, so every output is checked against that
    closed form (slab by slab, in FP64); the inverse returns the impulse: 1 at
    r0 and below `tol` everywhere else."""
    import torch
    nx, ny, nz = shape
    cdt = torch.complex64 if prec == "fp32" else torch.complex128
    need = nx * ny * nz * (8 if prec == "fp32" else 16)
    torch.cuda.empty_cache()   # blocks cached by earlier tests count as used
    free, _ = torch.cuda.mem_get_info()
    if free < need + (32 << 30):
        pytest.skip(f"needs {need >> 30} GiB plus 32 GiB of work space on the device")
    r0 = (1234, 77, 1999 % nz)
    x = torch.zeros(shape, dtype=cdt, device="cuda")
    x[r0] = 1
    plan = a2afft_mod.GridDfft(shape, prec)
    plan.forward(x)
    torch.cuda.synchronize()
    ph = []
    for n, c in zip(shape, r0):
        k = torch.arange(n, device="cuda", dtype=torch.int64)
        ph.append(torch.exp(((k * c) % n).to(torch.float64) * (-2j * math.pi / n)))   # exact phase reduction
    yz = ph[1][:, None] * ph[2][None, :]
    worst = 0.0
    for s in range(0, nx, 64):
        want = ph[0][s:s + 64, None, None] * yz[None]
        worst = max(worst, float((x[s:s + 64].to(torch.complex128) - want).abs().max()))
        del want
    assert worst < tol, worst
    plan.inverse(x)
    torch.cuda.synchronize()
    peak = complex(x[r0].item())
    assert abs(peak - 1) < tol, peak
    x[r0] = 0
    rest = max(float(x[s:s + 64].abs().max()) for s in range(0, nx, 64))
    assert rest < tol, rest


def test_2048cube_over_64_virtual_ranks(a2afft_mod, cuda):
    """The whole grid of the rank-share configuration on one device: 2048^3
    complex64 over 64 virtual ranks, dims [1,8,8] (64 GiB, the reference's
    64-process layout).  An impulse at global r0, placed in its rank's block,
    transforms to exp(-2 pi i k.r0 / N) at every k: each rank's block is
    checked against that closed form in FP64; the inverse returns the
    impulse."""
    import torch
    Ng, P, dims, tol = 2048, 64, (1, 8, 8), 1e-5
    torch.cuda.empty_cache()   # blocks cached by earlier tests count as used
    free, _ = torch.cuda.mem_get_info()
    if free < Ng ** 3 * 8 + (32 << 30):
        pytest.skip("needs 64 GiB plus 32 GiB of work space on the device")
    lx, ly, lz = Ng // dims[0], Ng // dims[1], Ng // dims[2]
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    x = torch.zeros((P, lx, ly, lz), dtype=torch.complex64, device="cuda")
    r0 = (1234, 777, 1999)
    rank0 = ((r0[0] // lx) * dims[1] + r0[1] // ly) * dims[2] + r0[2] // lz   # rank2coords, row-major
    at0 = (rank0, r0[0] % lx, r0[1] % ly, r0[2] % lz)
    x[at0] = 1
    plan = a2afft_mod.Dfft(dist)
    plan.make_plans(None)
    plan.forward(x)
    torch.cuda.synchronize()
    k = torch.arange(Ng, device="cuda", dtype=torch.int64)
    ph = [torch.exp(((k * c) % Ng).to(torch.float64) * (-2j * math.pi / Ng)) for c in r0]
    worst = 0.0
    for r in range(P):
        c0, c1, c2 = r // (dims[1] * dims[2]), (r // dims[2]) % dims[1], r % dims[2]
        yz = ph[1][c1 * ly:(c1 + 1) * ly, None] * ph[2][None, c2 * lz:(c2 + 1) * lz]
        for s in range(0, lx, 512):
            want = ph[0][c0 * lx + s:c0 * lx + s + 512, None, None] * yz[None]
            worst = max(worst, float((x[r, s:s + 512].to(torch.complex128) - want).abs().max()))
            del want
    assert worst < tol, worst
    plan.inverse(x)
    torch.cuda.synchronize()
    peak = complex(x[at0].item())
    assert abs(peak - 1) < tol, peak
    x[at0] = 0
    rest = max(float(x[r].abs().max()) for r in range(P))
    assert rest < tol, rest


def test_2048cube_over_64_virtual_ranks_staged_equals_fused(a2afft_mod, cuda):
    """The reference's stage order (18 launches: transposes, device-local
    all-to-alls, pack / unpack, batched 2048-point FFTs, copy) on the whole
    2048^3 complex64 grid over 64 virtual ranks, dims [1,8,8]: the same bits
    as the fused plan (input, scratch and the fused result take 192 GiB)."""
    import torch
    Ng, P, dims = 2048, 64, (1, 8, 8)
    n = Ng ** 3
    torch.cuda.empty_cache()   # blocks cached by earlier tests count as used
    free, _ = torch.cuda.mem_get_info()
    if free < 3 * n * 8 + (16 << 30):
        pytest.skip("needs 192 GiB plus 16 GiB of work space on the device")
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype="c64")
    x = torch.empty(n, dtype=torch.complex64, device="cuda")
    a2afft_mod.fill_random(x, 31, "uniform")
    y = x.clone()
    fused = a2afft_mod.Dfft(dist)
    fused.make_plans(None)
    fused.forward(y)
    torch.cuda.synchronize()
    staged = a2afft_mod.Dfft(dist, exec="staged")
    scratch = torch.empty_like(x)
    staged.make_plans(scratch)
    staged.forward(x)
    torch.cuda.synchronize()
    assert _same_bits(x, y)
    del staged, scratch, x
    step = 1 << 30
    assert all(bool(torch.isfinite(torch.view_as_real(y[s:s + step])).all()) for s in range(0, n, step))


def test_c5_batch_of_64_grids(a2afft_mod, cuda):
    """The 64-grid batch (shapes drawn with seed 4): every grid equals its
    single-grid plan bit for bit, for the default and a grouped, skewed order;
    one grid of each shape within 6e-7 of FP64."""
    import torch
    shapes = _c5_shapes()
    b = a2afft_mod.Batch(shapes)
    x = b.empty()
    a2afft_mod.fill_random(x, 4, "uniform")
    for group_points, skewed in ((0, False), (1 << 25, True)):
        b.set_grouping(group_points, skewed)
        y = x.clone()
        b.forward(y)
        assert b.status() == 0
        plans = {}
        checked = set()
        for i, s in enumerate(shapes):
            lo, hi = b.offsets[i], b.offsets[i + 1]
            if s not in plans:
                plans[s] = a2afft_mod.GridDfft(s, "fp32")
            plan = plans[s]
            g = x[lo:hi].clone()
            plan.forward(g)
            assert _same_bits(y[lo:hi], g), (i, s, group_points)
            if s not in checked:
                checked.add(s)
                r = x[lo:hi].to(torch.complex128)
                a2afft_mod.GridDfft(s, "fp64").forward(r)
                rel, _ = a2afft_mod.error_metrics(y[lo:hi], r)
                assert rel < 6e-7, (s, rel)
        del y
