"""GPU tests of plan-level hipGraph replay (a2afft_graph_launch): every replay
equals the eager forward / inverse bit for bit, with other eager work on the
device between replays, and a plan change makes the next call capture again."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(64, 4, (2, 2, 1), "c128", "fused"), (64, 4, (2, 2, 1), "c128", "staged"), (96, 8, (1, 2, 4), "c64", "fused"),
         (14, 2, None, "c64", "fused"), (32, 8, (1, 2, 4), "c64", "staged")]


def _bits(t):
    a = np.ascontiguousarray(t.cpu().numpy())
    return a.view(np.uint64 if a.dtype == np.complex128 else np.uint32)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[3]}_{c[4]}")
def test_plan_graph_replays_match_eager(a2afft_mod, cuda, case):
    import torch
    Ng, P, dims, dtype, ex = case
    dist = a2afft_mod.Distribution(Ng, P, dims=dims, dtype=dtype)
    plan = a2afft_mod.Dfft(dist, exec=ex)
    scratch = dist.empty()
    plan.make_plans(scratch)
    x = dist.empty()
    a2afft_mod.fill_random(x, Ng + P, "uniform")
    ref_f = x.clone()
    plan.forward(ref_f)
    ref_i = ref_f.clone()
    plan.inverse(ref_i)
    torch.cuda.synchronize()
    data = torch.empty_like(x)
    other = torch.empty_like(x)
    for rep in range(3):
        data.copy_(x)
        torch.cuda.synchronize()
        plan.graph_launch(data, -1)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(data), _bits(ref_f)), rep
        other.copy_(x)
        plan.forward(other)   # eager work between replays
        torch.cuda.synchronize()
    plan.graph_launch(data, 1)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(data), _bits(ref_i))
    # a plan change drops the graphs; the next call captures again
    plan.make_plans(scratch)
    data.copy_(x)
    plan.graph_launch(data, -1)
    torch.cuda.synchronize()
    assert np.array_equal(_bits(data), _bits(ref_f))
