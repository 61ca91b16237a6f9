"""GPU tests of a2afft_plan_pattern_copy: the pattern-copy kernel of each
fused pass leaves the data bit-identical, reports a positive time, and is
refused for passes that do not run a fixed-schedule Stockham kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (Ng, P, dims, dtype): contiguous and strided tiles, FP32 and FP64, the
# 512-point shapes of the 512^3 decomposition on a smaller grid
CASES = [(64, 4, (2, 2, 1), "c128"), (256, 8, (1, 2, 4), "c64"), (32, 8, (1, 2, 4), "c64"), (128, 4, (1, 2, 2), "c128")]


def _bits(t):
    a = np.ascontiguousarray(t.cpu().numpy())
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[3]}")
def test_pattern_copy_keeps_data(a2afft_mod, cuda, case):
    import torch
    Ng, P, dims, dtype = case
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype=dtype)
    plan = a2afft_mod.Dfft(dist, exec="fused")
    scratch = dist.empty()
    plan.make_plans(scratch)
    x = dist.empty()
    a2afft_mod.fill_random(x, Ng + P, "uniform")
    ref = x.clone()
    for p in range(3):
        ms = plan.pattern_copy(x, p, 3)
        torch.cuda.synchronize()
        assert ms > 0.0
        assert np.array_equal(_bits(x), _bits(ref)), p


def test_pattern_copy_512_grid(a2afft_mod, cuda):
    """The 512-point pass shapes (E = 16 and 32, strided and contiguous) over
    one 512 x 64 x 64 grid."""
    import torch
    plan = a2afft_mod.GridDfft((512, 64, 64), "fp32")
    x = torch.randn(512 * 64 * 64, dtype=torch.complex64, device="cuda")
    ref = x.clone()
    for p in range(3):
        assert plan.pattern_copy(x, p, 2) > 0.0
    torch.cuda.synchronize()
    assert np.array_equal(_bits(x), _bits(ref))


def test_pattern_copy_refusals(a2afft_mod, cuda):
    """Split-precision plans (MFMA kernels), direct-DFT passes (N = 14) and
    runtime-schedule passes (N = 96) have no pattern copy; bad pass indices
    are argument errors."""
    dist = a2afft_mod.Distribution(64, 4, dims=(2, 2, 1), dtype="c64")
    split = a2afft_mod.Dfft(dist, precision="split_bf16", exec="fused")
    x = dist.empty()
    split.make_plans(dist.empty())
    with pytest.raises(a2afft_mod.A2AFFTError, match="fused FP32/FP64"):
        split.pattern_copy(x, 0)
    d14 = a2afft_mod.Distribution(14, 2, dtype="c64")
    p14 = a2afft_mod.Dfft(d14, exec="fused")
    p14.make_plans(d14.empty())
    with pytest.raises(a2afft_mod.A2AFFTError, match="fixed-schedule"):
        p14.pattern_copy(d14.empty(), 0)
    d96 = a2afft_mod.Distribution(96, 8, dims=(1, 2, 4), dtype="c64")
    p96 = a2afft_mod.Dfft(d96, exec="fused")
    p96.make_plans(d96.empty())
    with pytest.raises(a2afft_mod.A2AFFTError, match="fixed-schedule"):
        p96.pattern_copy(d96.empty(), 0)
    plan = a2afft_mod.Dfft(dist, exec="fused")
    plan.make_plans(dist.empty())
    with pytest.raises(a2afft_mod.A2AFFTError):
        plan.pattern_copy(x, 3)
