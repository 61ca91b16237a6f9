"""The C++ driver (reference test.cpp behaviour over the C++ class mirror and
the C ABI) on the reference's own integration list, checked with the
reference's own criterion via a2afft.stitch (src/a2afft/tests/run_tests.py:18-21,
stitch.py:89)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "3d-ffts-mpi_amd", "a2afft_driver")

# (P, Ng, blockSize, dims, exec): run_tests.py:18-21 -- (8,8), (8,32), (6,12),
# (4,4) -- plus the C1 and C2 decompositions, a mixed-radix 24^3 over 12 and
# 14^3 over 2 (direct DFT).
CASES = [
    (8, 8, 64, None, "fused"), (8, 8, 64, None, "staged"), (8, 32, 128, None, "fused"), (8, 32, 128, None, "staged"),
    (4, 4, 64, None, "staged"), (4, 4, 64, None, "fused"), (6, 12, 128, None, "fused"), (6, 12, 128, None, "staged"),
    (6, 12, 64, None, "fused"),
    (4, 64, 64, (2, 2, 1), "staged"), (4, 64, 64, (2, 2, 1), "fused"), (8, 32, 64, (1, 2, 4), "fused"),
    (64, 32, 64, (4, 4, 4), "staged"), (12, 24, 64, None, "fused"), (12, 24, 64, None, "staged"),
    (2, 14, 64, None, "fused"), (2, 14, 64, None, "staged"),   # a factor 7: the direct DFT
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"P{c[0]}_Ng{c[1]}_bs{c[2]}_{c[3]}_{c[4]}")
def test_driver_reference_check(cuda, tmp_path, case):
    from a2afft import stitch
    P, Ng, bs, dims, ex = case
    cmd = [DRIVER, "-Ng", str(Ng), "-P", str(P), "-blockSize", str(bs), "-exec", ex, "-out", str(tmp_path)]
    if dims:
        cmd += ["-dims", *map(str, dims)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    fwd_ok, inv_ok, rel_f, rel_i = stitch.check(str(tmp_path))
    assert fwd_ok and inv_ok
    assert rel_f < 1e-14 and rel_i < 1e-14
    assert sorted(os.listdir(tmp_path)) == sorted(["dims"] + [f"proc{r}" for r in range(P)])


def test_unsupported_length_is_an_error(cuda, tmp_path):
    """A line longer than 4096^2 with no split into two kernel lengths, and
    above the chirp-z limit (2^23): a clean error, not a launch."""
    import torch
    import a2afft
    N = (1 << 24) + 1
    d = torch.zeros(N, dtype=torch.complex64, device="cuda")
    with pytest.raises(a2afft.A2AFFTError, match="neither a product of two lengths"):
        a2afft.forward_1d_fft(d, N)
