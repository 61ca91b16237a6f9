"""CPU tests of the C ABI (include/a2afft.h) and the host logic behind it.

No kernel runs here: the library must load and export every declared entry
point, and the decomposition entry points (pure host code) must reproduce the
reference's a2aDistribution fields, preconditions and `dims` file.
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "a2afft.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(a2afft_\w+)\s*\(", text)))


def test_header_compiles_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "a2afft.h"\nint main(void){ a2afft_dist d = 0; (void)d; return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c", str(src),
                    "-o", str(tmp_path / "t.o")], check=True)


def test_library_exports_every_declared_symbol(a2afft_mod):
    names = _declared()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", a2afft_mod.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(a2afft_\w+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(a2afft_mod.lib(), n)


def test_python_bindings_cover_header(a2afft_mod):
    from a2afft import _lib
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_dims_create(a2afft_mod):
    table = {4: (2, 2, 1), 6: (3, 2, 1), 8: (2, 2, 2), 12: (3, 2, 2), 16: (4, 2, 2), 64: (4, 4, 4), 7: (7, 1, 1)}
    for P, dims in table.items():
        assert a2afft_mod.dims_create(P) == dims


def test_dist_fields_and_preconditions(a2afft_mod):
    A = a2afft_mod
    d = A.Distribution(256, 8, dims=(1, 2, 4), dtype="c128")
    assert d.local_grid_size == (256, 128, 64) and d.nlocal == 256 * 128 * 64 and d.dims == (1, 2, 4)
    assert d.reference_domain
    assert d.coords(5) == ((0, 1, 1), (0, 128, 64))
    d = A.Distribution(8, 8)                           # MPI_Dims_create rule
    assert d.dims == (2, 2, 2)
    assert not A.Distribution(32, 64, dims=(4, 4, 4)).reference_domain
    with pytest.raises(A.A2AFFTError) as e:
        A.Distribution(9, 4)
    assert e.value.status == -2                        # Ng % dims (a2adistribution.cpp:61-63)
    with pytest.raises(A.A2AFFTError) as e:
        A.Distribution(10, 4)
    assert e.value.status == -4                        # (lx*ly) % P (a2adistribution.cpp:78)
    with pytest.raises(A.A2AFFTError) as e:
        A.Distribution(8, 8, dims=(2, 2, 4))
    assert e.value.status == -3
    with pytest.raises(A.A2AFFTError):
        A.Distribution(8, 8, blockSize=0)


def test_write_dims_matches_reference_layout(a2afft_mod, tmp_path):
    """int32 world_size, Ng, nlocal, dims[3], local_grid_size[3] (a2adistribution.cpp:101-111),
    the header tests/stitch.py:59-66 reads."""
    d = a2afft_mod.Distribution(32, 8, dims=(1, 2, 4))
    path = tmp_path / "dims"
    d.write_dims(path)
    v = np.fromfile(path, dtype=np.int32)
    assert v.tolist() == [8, 32, 32 * 16 * 8, 1, 2, 4, 32, 16, 8]


def test_status_strings(a2afft_mod):
    L = a2afft_mod.lib()
    for s in (0, 1, -1, -2, -3, -4, -5, -6, -7, -8, -9):
        assert L.a2afft_status_string(s)
    assert L.a2afft_version() >= 100


def test_host_generator_matches_oracle():
    from a2afft import inputs
    from oracle import oracle as O
    assert np.array_equal(inputs.uniform(7, 5000), O.fill_uniform(7, 5000))
    assert np.array_equal(inputs.uniform(7, 100, start=4900), O.fill_uniform(7, 5000)[4900:])
    assert np.abs(inputs.normal(7, 5000) - O.fill_normal(7, 5000)).max() < 1e-14
    c0, c1, c2, c3 = inputs.philox4x32_10(np.array([0x243f6a88]), np.array([0x85a308d3]), np.array([0x13198a2e]),
                                          np.array([0x03707344]), 0xa4093822, 0x299f31d0)
    assert (int(c0[0]), int(c1[0]), int(c2[0]), int(c3[0])) == (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_block_layout_helpers_match_oracle():
    from a2afft import inputs
    from oracle import oracle as O
    rc, d = O.decompose(16, 16, (4, 2, 2))
    g = np.arange(16 ** 3).reshape(16, 16, 16)
    b = inputs.global_to_blocks(g, d.dims_t)
    assert np.array_equal(b, O.global_to_blocks(d, g.astype(np.int64)))
    assert np.array_equal(inputs.blocks_to_global(b, d.dims_t, g.shape), g)


def test_grf_spectrum_and_normalisation():
    """P(k) ~ k^-3 field: unit RMS before the 1e-3 noise, no k=0 power."""
    from a2afft import inputs
    f = inputs.gaussian_random_field((32, 32, 32), seed=2, noise=0.0)
    assert abs(np.sqrt(np.mean(np.abs(f) ** 2)) - 1) < 1e-12
    F = np.fft.fftn(f)
    assert abs(F[0, 0, 0]) < 1e-9 * np.abs(F).max()
    k = np.fft.fftfreq(32, 1 / 32)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    lo = np.mean(np.abs(F[(k2 >= 1) & (k2 < 4)]) ** 2)
    hi = np.mean(np.abs(F[(k2 >= 64) & (k2 < 100)]) ** 2)
    assert 50 < lo / hi < 2000                          # ~ (8/1.5)^3 ~ 150 for k^-3
