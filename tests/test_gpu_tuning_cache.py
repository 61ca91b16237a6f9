"""GPU tests of the on-disk tuning cache (a plan started from a recorded
autotune choice; no reference counterpart, the reference re-plans on every
call, a2affts.cu:15)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def test_autotune_result_is_reused_and_bit_identical(a2afft_mod, cuda, tmp_path):
    import torch
    cache = tmp_path / "tune.txt"
    a2afft_mod.set_tuning_cache(str(cache))
    try:
        shape = (64, 128, 256)
        plan = a2afft_mod.GridDfft(shape, "fp32")
        x = torch.empty(shape, dtype=torch.complex64, device="cuda")
        a2afft_mod.fill_random(x, 7, "uniform")
        work = x.clone()
        plan.autotune(work, iters=2)
        tuned = plan.describe().splitlines()[:3]
        lines = cache.read_text().splitlines()
        assert len(lines) == 1 and "grid=64x128x256" in lines[0] and "prec=1" in lines[0]
        # a new plan with the same key starts from the recorded choice
        again = a2afft_mod.GridDfft(shape, "fp32")
        desc = again.describe()
        assert desc.splitlines()[:3] == tuned
        assert "tuning from cache" in desc
        y1, y2 = x.clone(), x.clone()
        plan.forward(y1)
        again.forward(y2)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(y1.cpu().numpy()), _bits(y2.cpu().numpy()))
        # a second autotune of another key adds a line, the same key replaces its own
        other = a2afft_mod.GridDfft((32, 32, 32), "fp32")
        w2 = torch.empty((32, 32, 32), dtype=torch.complex64, device="cuda")
        other.autotune(w2, iters=1)
        plan.autotune(work, iters=1)
        assert len(cache.read_text().splitlines()) == 2
    finally:
        a2afft_mod.set_tuning_cache(None)


def test_bad_cache_entries_are_ignored(a2afft_mod, cuda, tmp_path):
    cache = tmp_path / "tune.txt"
    plan0 = a2afft_mod.GridDfft((64, 64, 64), "fp32")
    default = plan0.describe().splitlines()[:3]
    a2afft_mod.set_tuning_cache(str(cache))
    try:
        # capture this plan's key by tuning once, then corrupt the recorded kernels
        import torch
        w = torch.empty((64, 64, 64), dtype=torch.complex64, device="cuda")
        plan0.autotune(w, iters=1)
        key = cache.read_text().split(" | ")[0]
        cache.write_text(key + " | z=3,7,0 x=1,1,1 y=5,5,5 chunk=0\n" + "garbage line\n")
        p = a2afft_mod.GridDfft((64, 64, 64), "fp32")
        assert p.describe().splitlines()[:3] == default
        cache.write_text(key + " | z=16\n")
        p = a2afft_mod.GridDfft((64, 64, 64), "fp32")
        assert p.describe().splitlines()[:3] == default
    finally:
        a2afft_mod.set_tuning_cache(None)
