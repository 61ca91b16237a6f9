"""Reader and checker for the reference-format outputs of the driver.

Reads the `dims` header and the `proc<r>` dumps (the formats of
src/a2afft/distribution/a2adistribution.cpp:101-111 and
src/a2afft/tests/test.cpp:69-125), reassembles the global grid of each step
from the rank blocks (rank2coords, a2adistribution.cpp:14-20) and applies the
reference's pass/fail check, np.fft.fftn(input) ~ forward and input ~ inverse
with numpy's default allclose tolerances (src/a2afft/tests/stitch.py:89),
computed here in complex128 and reported with rel-L2 as well.
"""
import os

import numpy as np


def read_dims(folder):
    v = np.fromfile(os.path.join(folder, "dims"), dtype=np.int32)
    if v.size != 9:
        raise ValueError("dims header must hold 9 int32 values")
    return {"world_size": int(v[0]), "Ng": int(v[1]), "nlocal": int(v[2]), "dims": tuple(int(x) for x in v[3:6]),
            "local_grid_size": tuple(int(x) for x in v[6:9])}


def read_steps(folder, hdr, dtype=np.float64):
    """(world_size, steps, nlocal) complex array of every proc<r> file."""
    out = []
    for r in range(hdr["world_size"]):
        raw = np.fromfile(os.path.join(folder, f"proc{r}"), dtype=dtype)
        c = raw[0::2] + 1j * raw[1::2]
        out.append(c.reshape(-1, hdr["nlocal"]))
    return np.array(out)


def stitch(blocks, hdr):
    """Global (Ng,Ng,Ng) grid from per-rank blocks [P][nlocal]."""
    d0, d1, d2 = hdr["dims"]
    lx, ly, lz = hdr["local_grid_size"]
    Ng = hdr["Ng"]
    return blocks.reshape(d0, d1, d2, lx, ly, lz).transpose(0, 3, 1, 4, 2, 5).reshape(Ng, Ng, Ng)


def check(folder):
    """Returns (forward_ok, inverse_ok, forward_rel_l2, inverse_rel_l2)."""
    hdr = read_dims(folder)
    raw = read_steps(folder, hdr)
    inp, out, inv = (stitch(raw[:, k, :], hdr) for k in range(3))
    ref = np.fft.fftn(inp)
    rel_f = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    rel_i = float(np.linalg.norm(inv - inp) / np.linalg.norm(inp))
    return bool(np.allclose(ref, out)), bool(np.allclose(inp, inv)), rel_f, rel_i
