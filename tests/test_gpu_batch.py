"""GPU tests of the persistent batched transform (many uneven grids, one launch).

Every grid of a batch must equal, bit for bit, the single-grid plan of its
shape (same per-line arithmetic), and numpy's fftn within the FP32 bound; the
hipGraph replay must equal the eager execution on every replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (64, 64, 64), (128, 256, 64), (16, 32, 64), (64, 16, 128), (32, 64, 16),
          (128, 64, 32), (16, 16, 16)]


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _fill(a2afft_mod, torch, batch, seed):
    buf = torch.empty(batch.total, dtype=torch.complex64, device="cuda")
    a2afft_mod.fill_random(buf, seed, "uniform")
    return buf


@pytest.mark.parametrize("workgroups,xcdq,chunk", [(0, 0, 1), (0, 1, 1), (0, 0, 4), (1, 1, 3), (7, 1, 1),
                                                   (7, 0, 2), (16, 1, 5), (3, 0, 64)])
def test_batch_matches_grid_plans_and_numpy(a2afft_mod, cuda, workgroups, xcdq, chunk):
    import torch
    b = a2afft_mod.Batch(SHAPES)
    b.set_schedule(workgroups, bool(xcdq), chunk)
    x = _fill(a2afft_mod, torch, b, 4)
    x0 = x.cpu().numpy()
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    got = y.cpu().numpy()
    for i, s in enumerate(SHAPES):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        g = torch.from_numpy(x0[lo:hi].copy()).cuda()
        plan = a2afft_mod.GridDfft(s, "fp32")
        plan.forward(g)
        torch.cuda.synchronize()
        assert np.array_equal(_bits(got[lo:hi]), _bits(g.cpu().numpy())), s
        ref = np.fft.fftn(x0[lo:hi].reshape(s).astype(np.complex128))
        err = np.linalg.norm(got[lo:hi].reshape(s) - ref) / np.linalg.norm(ref)
        assert err < 6e-7, (s, err)
    b.inverse(y)
    assert b.status() == 0
    back = y.cpu().numpy()
    assert np.linalg.norm(back - x0) / np.linalg.norm(x0) < 1e-6


@pytest.mark.parametrize("group_points,skewed,xcdq,chunk", [(1, 0, 0, 1), (1, 1, 1, 1), (100000, 1, 0, 2),
                                                            (300000, 0, 1, 1), (2500000, 1, 1, 3), (1 << 40, 1, 0, 1)])
def test_batch_grouped_orders_match_default(a2afft_mod, cuda, group_points, skewed, xcdq, chunk):
    """Group-by-group and skewed phase orders compute the bits of the default
    pass-major order, eagerly and through the graph."""
    import torch
    b = a2afft_mod.Batch(SHAPES)
    x = _fill(a2afft_mod, torch, b, 6)
    ref = x.clone()
    b.forward(ref)
    assert b.status() == 0
    ref = ref.cpu().numpy()
    b.set_grouping(group_points, bool(skewed))
    b.set_schedule(0, bool(xcdq), chunk)
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref))
    for rep in range(2):
        y.copy_(x)
        torch.cuda.synchronize()
        b.graph_launch(y, -1)
        assert b.status() == 0
        assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref)), rep
    b.inverse(y)
    assert b.status() == 0
    x0 = x.cpu().numpy()
    assert np.linalg.norm(y.cpu().numpy() - x0) / np.linalg.norm(x0) < 1e-6


def test_batch_autotune_keeps_bits(a2afft_mod, cuda):
    """The library's autotune picks a schedule among equivalent ones: results
    equal the default schedule's bit for bit, and describe reports it."""
    import torch
    b = a2afft_mod.Batch(SHAPES)
    x = _fill(a2afft_mod, torch, b, 12)
    ref = x.clone()
    b.forward(ref)
    work = b.empty()
    a2afft_mod.fill_random(work, 1, "uniform")
    b.autotune(work, iters=1)
    sched = b.schedule()
    assert 1 <= sched["workgroups"] <= sched["max_workgroups"] and sched["chunk"] in (1, 2, 4)
    assert len(b.describe().splitlines()) > 20   # one timing line per candidate
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref.cpu().numpy()))
    b.graph_launch(y, 1)
    assert b.status() == 0
    x0 = x.cpu().numpy()
    assert np.linalg.norm(y.cpu().numpy() - x0) / np.linalg.norm(x0) < 1e-6


# (N, strided, T, E, pad): every alternative tile shape of the kernel's table
ALT_SHAPES = [(32, 0, 64, 8, 0), (64, 0, 32, 8, 0), (128, 0, 32, 16, 0), (32, 0, 32, 8, 1), (64, 0, 16, 8, 1),
              (128, 0, 16, 16, 1), (32, 1, 32, 8, 0), (64, 1, 32, 8, 0), (128, 1, 32, 16, 0), (64, 1, 16, 8, 1),
              (128, 1, 16, 16, 1), (32, 1, 64, 8, 1), (64, 1, 64, 16, 1), (32, 1, 128, 16, 1), (16, 1, 256, 16, 1),
              (64, 0, 64, 16, 1), (32, 0, 128, 16, 1)]
ALT_SHAPES_BIG = [(256, 0, 16, 16, 0), (512, 0, 8, 16, 0), (256, 0, 8, 16, 1), (512, 0, 4, 16, 1),
                  (256, 1, 16, 16, 0), (512, 1, 8, 16, 0), (256, 1, 8, 16, 1), (512, 1, 4, 16, 1)]


def test_batch_tile_shapes_keep_bits(a2afft_mod, cuda):
    """Every alternative tile shape computes the primary's bits."""
    import torch
    wide = [(16, 32, 256), (64, 64, 256), (32, 128, 256)]   # z extents that the wide strided tiles fit
    for shapes, alts in ((SHAPES + wide, ALT_SHAPES), ([(512, 256, 512), (256, 512, 256)], ALT_SHAPES_BIG)):
        b = a2afft_mod.Batch(shapes)
        x = _fill(a2afft_mod, torch, b, 13)
        ref = x.clone()
        b.forward(ref)
        ref = ref.cpu().numpy()
        for N, strided, T, E, pad in alts:
            b.set_shape(N, strided, T, E, bool(pad))
            y = x.clone()
            b.forward(y)
            assert b.status() == 0
            assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref)), (N, strided, T, E, pad)
            b.set_shape(N, strided)   # back to the primary
    with pytest.raises(a2afft_mod.A2AFFTError):
        b.set_shape(512, 1, 3, 16)   # not in the table


def test_batch_grouping_limits(a2afft_mod, cuda):
    """At most 128 groups; a rejected grouping leaves the previous one in force."""
    import torch
    shapes = [(16, 16, 16)] * 130
    b = a2afft_mod.Batch(shapes)
    x = _fill(a2afft_mod, torch, b, 8)
    ref = x.clone()
    b.forward(ref)
    with pytest.raises(a2afft_mod.A2AFFTError):
        b.set_grouping(1, False)
    with pytest.raises(a2afft_mod.A2AFFTError):
        b.set_grouping(-1, False)
    b.set_grouping(2 * 16 ** 3, True)   # 65 groups
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    assert np.array_equal(_bits(y.cpu().numpy()), _bits(ref.cpu().numpy()))


def test_batch_graph_replays_match_eager(a2afft_mod, cuda):
    import torch
    b = a2afft_mod.Batch(SHAPES)
    x = _fill(a2afft_mod, torch, b, 9)
    eager = x.clone()
    b.forward(eager)
    ref = eager.cpu().numpy()
    data = torch.empty_like(x)
    for rep in range(4):
        data.copy_(x)
        torch.cuda.synchronize()
        b.graph_launch(data, -1)
        assert b.status() == 0
        assert np.array_equal(_bits(data.cpu().numpy()), _bits(ref)), rep
        # eager work between replays
        other = x.clone()
        b.forward(other)
        torch.cuda.synchronize()
    b.graph_launch(data, 1)
    assert b.status() == 0
    assert np.linalg.norm(data.cpu().numpy() - x.cpu().numpy()) / np.linalg.norm(x.cpu().numpy()) < 1e-6


def test_batch_rejects_unsupported_shapes(a2afft_mod, cuda):
    with pytest.raises(a2afft_mod.A2AFFTError):
        a2afft_mod.Batch([(32, 32, 1024)])
    with pytest.raises(a2afft_mod.A2AFFTError):
        a2afft_mod.Batch([(24, 32, 32)])


# (shape, dims): cubic grids over the reference's decompositions and some
# non-cubic blocks, mixed with undecomposed grids
RANK_ITEMS = [((64, 64, 64), (1, 2, 4)), ((32, 32, 32), (2, 2, 1)), ((128, 128, 128), (2, 2, 2)),
              ((64, 64, 64), (1, 1, 1)), ((128, 256, 64), (2, 2, 1)), ((64, 64, 64), (4, 2, 2)),
              ((32, 32, 32), (1, 4, 4)), ((256, 64, 128), (1, 2, 2))]


def test_batch_rank_blocks_match_plans_and_oracle(a2afft_mod, orc, cuda):
    """Grids decomposed into virtual-rank blocks inside the persistent launch:
    every cubic item equals, bit for bit, the fused plan over its distribution
    (same block layout), is within 1e-6 of the oracle's reference-order FP64
    transform at 32^3 and 64^3, and every item (non-cubic included) matches
    numpy's fftn of the stitched grid; the inverse returns the input."""
    import torch
    from a2afft import inputs
    shapes = [s for s, _ in RANK_ITEMS]
    dims = [d for _, d in RANK_ITEMS]
    b = a2afft_mod.Batch(shapes, dims=dims)
    x = _fill(a2afft_mod, torch, b, 12)
    x0 = x.cpu().numpy()
    y = x.clone()
    b.forward(y)
    assert b.status() == 0
    got = y.cpu().numpy()
    for i, (s, d) in enumerate(RANK_ITEMS):
        lo, hi = b.offsets[i], b.offsets[i + 1]
        P = d[0] * d[1] * d[2]
        blocks = x0[lo:hi].reshape(P, -1)
        if s[0] == s[1] == s[2]:
            dist = a2afft_mod.Distribution(s[0], P, dims=d, dtype="c64")
            g = torch.from_numpy(blocks.copy()).cuda()
            a2afft_mod.Dfft(dist, "fp32").forward(g)
            torch.cuda.synchronize()
            assert np.array_equal(_bits(got[lo:hi]), _bits(g.cpu().numpy().reshape(-1))), (s, d)
            if s[0] <= 64:
                rc, od = orc.decompose(s[0], P, d)
                assert rc == 0
                ref = orc.forward(od, blocks.astype(np.complex128))
                err = np.linalg.norm(got[lo:hi].reshape(P, -1) - ref) / np.linalg.norm(ref)
                assert err < 1e-6, (s, d, err)
        glob = inputs.blocks_to_global(got[lo:hi].reshape(P, -1), d, s)
        ref = np.fft.fftn(inputs.blocks_to_global(blocks, d, s).astype(np.complex128))
        err = np.linalg.norm(glob - ref) / np.linalg.norm(ref)
        assert err < 6e-7, (s, d, err)
    b.inverse(y)
    assert b.status() == 0
    back = y.cpu().numpy()
    assert np.linalg.norm(back - x0) / np.linalg.norm(x0) < 1e-6
    # views: rank blocks for decomposed grids
    v = b.views(y)
    assert tuple(v[0].shape) == (8, 64 * 64 * 64 // 8) and tuple(v[3].shape) == (64, 64, 64)


def test_batch_rank_blocks_reject_bad_dims(a2afft_mod, cuda):
    from a2afft import _lib as L
    with pytest.raises(L.A2AFFTError):
        a2afft_mod.Batch([(64, 64, 64)], dims=[(3, 1, 1)])
    with pytest.raises(ValueError):
        a2afft_mod.Batch([(64, 64, 64), (32, 32, 32)], dims=[(1, 1, 1)])
