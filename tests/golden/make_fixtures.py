"""Writes tests/golden/remap_twins.npz from the reference's own C remap twins.

The twins (src/pencils_demo/a2a_to_pencils/a2a_to_pencils.c and
src/pencils_demo/change_fast_axis/change_fast_axis.c) are int-typed copies of
the CUDA index kernels (src/a2afft/distribution/kernels/a2apencils.cu,
a2achangefastaxis.cu).  `make -C oracle ref` compiles them, in place, into
oracle/_ref/libref_twins.so; this script feeds them iota and stores, per
configuration and kernel, the destination index of every source element
(perm[src] = dst).  Run from the repository root:

    make -C oracle && python tests/golden/make_fixtures.py

The fixture is data only: index permutations plus the configuration they
belong to.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# (Ng, P, dims): the reference's test list under MPI_Dims_create
# (src/a2afft/tests/run_tests.py:18-21) and the decompositions the benchmark
# configurations name, at sizes that keep the fixture small.
CONFIGS = [
    (8, 8, None), (32, 8, None), (12, 6, None), (4, 4, None),
    (32, 4, (2, 2, 1)), (32, 8, (1, 2, 4)), (32, 16, (1, 4, 4)), (64, 64, (1, 8, 8)), (16, 16, (4, 2, 2)),
]


def main():
    if O.ref_twins() is None:
        raise SystemExit("oracle/_ref/libref_twins.so is not built: run `make -C oracle ref`")
    out = {}
    for Ng, P, dims in CONFIGS:
        rc, d = O.decompose(Ng, P, dims)
        assert rc == 0, (Ng, P, dims, rc)
        tag = f"Ng{Ng}_P{P}_{d.dims[0]}x{d.dims[1]}x{d.dims[2]}"
        out[f"{tag}__cfg"] = np.array([Ng, P, *d.dims_t, *d.lgs_t, d.nlocal], dtype=np.int64)
        for kind, name in enumerate(O.REMAP_NAMES):
            perm = O.twin_perm(kind, d)
            dt = np.uint16 if d.nlocal <= 65536 else np.uint32
            out[f"{tag}__{name}"] = perm.astype(dt)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "remap_twins.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
