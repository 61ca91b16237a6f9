"""The reference-format reader/checker (a2afft.stitch) on files written from
the oracle's transform: it must accept a correct result and reject a wrong one."""
import numpy as np

from oracle import oracle as O


def _write(tmp_path, d, steps):
    hdr = np.array([d.P, d.Ng, d.nlocal, *d.dims_t, *d.lgs_t], dtype=np.int32)
    hdr.tofile(tmp_path / "dims")
    for r in range(d.P):
        with open(tmp_path / f"proc{r}", "wb") as f:
            for s in steps:
                np.ascontiguousarray(s[r]).view(np.float64).tofile(f)


def test_stitch_accepts_oracle_and_rejects_corruption(tmp_path):
    from a2afft import stitch
    rc, d = O.decompose(24, 12, None)
    assert rc == 0
    x = O.global_to_blocks(d, O.fill_uniform(3, 24 ** 3).reshape(24, 24, 24))
    f = O.forward(d, x)
    i = O.inverse(d, f)
    _write(tmp_path, d, [x, f, i])
    ok_f, ok_i, rel_f, rel_i = stitch.check(str(tmp_path))
    assert ok_f and ok_i and rel_f < 1e-15 and rel_i < 1e-15
    assert stitch.read_dims(str(tmp_path))["dims"] == d.dims_t
    bad = f.copy()
    bad[[0, 1]] = bad[[1, 0]]                  # two ranks swapped: a data-exchange bug
    _write(tmp_path, d, [x, bad, i])
    ok_f, ok_i, rel_f, _ = stitch.check(str(tmp_path))
    assert not ok_f and ok_i and rel_f > 0.1
