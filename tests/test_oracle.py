"""CPU tests that pin the oracle (oracle/a2afft_oracle.c) before it is trusted.

Each oracle function is checked against something it does not restate:
  - remaps: golden permutations from the reference's own compiled C twins
    (tests/golden/remap_twins.npz; live twins too when oracle/_ref is built),
    plus invariants (bijection, inverse pairs) and the global-ID trace;
  - the y-return defect of src/a2afft/distribution/kernels/a2apencils.cu:197:
    the reference formula is a bijection exactly on dims[1]==2 or dims[0]==1;
  - MPI_Dims_create: the MPI-standard table;
  - FFT: a long-double naive DFT (brute force), numpy's pocketfft, and known
    answers (delta, constant, single tone), Parseval, linearity, round trip;
  - 3D transform: np.fft.fftn of the stitched grid, the reference's own check
    (src/a2afft/tests/stitch.py:89), on the reference's own test list;
  - Philox4x32-10: the Random123 known-answer vectors.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "remap_twins.npz")


def _golden():
    z = np.load(GOLDEN, allow_pickle=False)
    cfgs = sorted({k.split("__")[0] for k in z.files})
    return z, cfgs


def test_dims_create_table():
    # MPI standard: balanced, non-increasing (a2adistribution.cpp:59)
    table = {1: (1, 1, 1), 2: (2, 1, 1), 4: (2, 2, 1), 6: (3, 2, 1), 8: (2, 2, 2), 12: (3, 2, 2),
             16: (4, 2, 2), 24: (4, 3, 2), 27: (3, 3, 3), 32: (4, 4, 2), 64: (4, 4, 4), 128: (8, 4, 4)}
    for P, dims in table.items():
        assert O.dims_create(P) == dims, P


def test_decompose_preconditions():
    assert O.decompose(64, 4, None)[0] == 0
    assert O.decompose(9, 4, None)[0] == -1           # Ng % dims (a2adistribution.cpp:61-63)
    assert O.decompose(10, 4, None)[0] == -3          # (lx*ly) % P (:78)
    assert O.decompose(32, 64, (1, 8, 8))[0] == -5    # (ly*lz) % P (:78-80)
    assert O.decompose(8, 8, (2, 2, 3))[0] == -6      # dims product != P
    rc, d = O.decompose(256, 8, (1, 2, 4))
    assert rc == 0 and d.lgs_t == (256, 128, 64) and d.nlocal == 256 * 128 * 64


def test_golden_fixture_present():
    z, cfgs = _golden()
    assert len(cfgs) >= 9


@pytest.mark.parametrize("tag", _golden()[1])
def test_oracle_remaps_match_reference_twins(tag):
    """perm[src] = dst from the compiled reference twins vs the oracle."""
    z, _ = _golden()
    cfg = z[f"{tag}__cfg"]
    Ng, P, d0, d1, d2 = (int(v) for v in cfg[:5])
    rc, d = O.decompose(Ng, P, (d0, d1, d2))
    assert rc == 0
    for kind, name in enumerate(O.REMAP_NAMES):
        gold = z[f"{tag}__{name}"].astype(np.int64)
        assert np.array_equal(O.remap_perm(kind, d), gold), name
        live = None
        if O.ref_twins() is not None:
            live = O.twin_perm(kind, d)
            assert np.array_equal(live, gold), f"{name}: fixture stale"


LIVE_CONFIGS = [(48, 24, None), (64, 32, None), (24, 12, None), (64, 8, (1, 4, 2)), (128, 16, (1, 4, 4))]


@pytest.mark.parametrize("cfg", LIVE_CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_oracle_remaps_match_live_twins(cfg):
    if O.ref_twins() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    for kind in range(9):
        if kind == O.Y_PENCILS_TO_Y_A2A and not O.y_return_in_reference_domain(d):
            continue
        assert np.array_equal(O.remap_perm(kind, d), O.twin_perm(kind, d)), O.REMAP_NAMES[kind]


INV_CONFIGS = [(8, 8, None), (32, 8, (1, 2, 4)), (32, 64, (4, 4, 4)), (16, 8, (2, 4, 1)), (27, 27, (3, 3, 3)),
               (16, 16, (4, 2, 2))]


@pytest.mark.parametrize("cfg", INV_CONFIGS, ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_remaps_are_bijections_with_inverse_pairs(cfg):
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    ident = np.arange(d.nlocal)
    perms = [O.remap_perm(k, d) for k in range(9)]
    for k, p in enumerate(perms):
        assert np.array_equal(np.sort(p), ident), O.REMAP_NAMES[k]
    for a, b in ((0, 1), (2, 3), (4, 5)):          # pencils_to_a2a inverts a2a_to_pencils
        assert np.array_equal(perms[b][perms[a]], ident), (O.REMAP_NAMES[a], O.REMAP_NAMES[b])


def test_y_return_reference_formula_domain():
    """a2apencils.cu:197 gives rank mp + s + (s/rpt)*rpt instead of
    mp + s%rpt + (s/rpt)*rpt*d1: a bijection exactly when d1 == 2 or d0 == 1."""
    checked = 0
    for P in (2, 4, 6, 8, 12, 16, 24, 27, 32, 64):
        for dims in {O.dims_create(P), (1,) + O.dims_create(P)[:2][::-1], O.dims_create(P)[::-1]}:
            if np.prod(dims) != P:
                continue
            Ng = 4 * int(np.lcm.reduce(dims)) * max(dims)
            rc, d = O.decompose(Ng, P, dims)
            if rc or d.nlocal > 300000:
                continue
            ref = O.remap_index(O.Y_PENCILS_TO_Y_A2A, d, ref_formula=True)
            inside = O.y_return_in_reference_domain(d)
            bij = (ref.min() >= 0 and ref.max() < d.nlocal and
                   np.array_equal(np.sort(ref), np.arange(d.nlocal)))
            assert bij == inside, (P, dims)
            if inside:
                assert np.array_equal(ref, O.remap_index(O.Y_PENCILS_TO_Y_A2A, d))
            checked += 1
    assert checked >= 10


@pytest.mark.parametrize("cfg", [(8, 8, None), (12, 6, None), (16, 16, (4, 2, 2)), (32, 8, (1, 2, 4))],
                         ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_global_id_trace(cfg):
    """The reference's own debugging method (src/pencils_demo/helpers/helpers.c:37-55):
    seed each cell with its global ID; after each get-stage every pencil must
    be one complete global line in natural order; the return stages restore
    the block layout."""
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    gid = np.arange(Ng ** 3, dtype=np.int64).reshape((Ng,) * 3)
    b1 = O.global_to_blocks(d, gid).copy()
    b2 = np.zeros_like(b1)
    start = b1.copy()

    def check_pencils(buf, axis):
        lines = buf.reshape(-1, Ng)
        coords = np.stack(np.unravel_index(lines, (Ng,) * 3))
        assert np.array_equal(coords[axis], np.broadcast_to(np.arange(Ng), lines.shape))
        other = [a for a in range(3) if a != axis]
        assert (coords[other[0]] == coords[other[0]][:, :1]).all()
        assert (coords[other[1]] == coords[other[1]][:, :1]).all()
        key = np.sort(coords[other[0]][:, 0] * Ng + coords[other[1]][:, 0])
        assert np.array_equal(key, np.arange(Ng * Ng))

    O.stage(O.GET_Z, d, b1, b2)
    check_pencils(b1, 2)
    O.stage(O.RETURN_Z, d, b1, b2)
    assert np.array_equal(b1, start)
    O.stage(O.GET_X, d, b1, b2)
    check_pencils(b2, 0)
    O.stage(O.RETURN_X, d, b1, b2)
    O.stage(O.GET_Y, d, b1, b2)
    check_pencils(b1, 1)
    b2[:] = b1                                       # launch_fast_copy_fft (a2adfft.cpp:79)
    O.stage(O.RETURN_Y, d, b1, b2)
    assert np.array_equal(b1, start)


def test_alltoall_closed_form():
    rc, d = O.decompose(32, 8, (1, 2, 4))
    P, n = d.P, d.nlocal
    send = np.arange(P * n, dtype=np.int64).reshape(P, n)
    recv = O.alltoall(d, send)
    ns = n // P
    for s in range(P):
        for q in range(P):
            assert np.array_equal(recv[s, q * ns:(q + 1) * ns], send[q, s * ns:(s + 1) * ns])


# ---- FFT ---------------------------------------------------------------

@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7, 8, 12, 14, 16, 30, 32, 49, 60, 63, 64])
def test_fft_matches_naive_dft(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    for sign in (-1, 1):
        ref = np.stack([O.dft_naive(x[i], sign) for i in range(3)])
        got = O.fft_batched(x, sign)
        assert np.abs(got - ref).max() <= 4e-15 * np.abs(ref).max() * np.log2(max(N, 2)), (N, sign)


@pytest.mark.parametrize("N", [128, 256, 512, 1024, 2048, 4096, 96, 384])
def test_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((5, N)) + 1j * rng.standard_normal((5, N))
    assert np.linalg.norm(O.fft_batched(x, -1) - np.fft.fft(x)) <= 1e-15 * np.log2(N) * np.linalg.norm(np.fft.fft(x))
    assert np.linalg.norm(O.fft_batched(x, 1) - N * np.fft.ifft(x)) <= 1e-15 * np.log2(N) * N * \
        np.linalg.norm(np.fft.ifft(x))


@pytest.mark.parametrize("N", [67, 97, 194, 1009])
def test_fft_large_prime_factor_matches_numpy(N):
    """Lengths with a prime factor above 64 go through the O(N^2) sum; numpy's
    pocketfft (Bluestein for large primes) is the independent check."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N)) + 1j * rng.standard_normal((3, N))
    ref = np.fft.fft(x)
    assert np.linalg.norm(O.fft_batched(x, -1) - ref) <= 1e-15 * np.sqrt(N) * np.linalg.norm(ref)
    assert np.linalg.norm(O.fft_batched(x, 1) - N * np.fft.ifft(x)) <= 1e-15 * np.sqrt(N) * np.linalg.norm(ref)


def test_fft_known_answers_and_identities():
    N, k0 = 512, 7
    x = np.zeros((3, N), dtype=np.complex128)
    x[0, 0] = 1
    x[1] = 1
    x[2] = np.exp(2j * np.pi * k0 * np.arange(N) / N)
    y = O.fft_batched(x, -1)
    assert np.allclose(y[0], 1, atol=0, rtol=0)
    e = np.zeros(N)
    e[0] = N
    assert np.abs(y[1] - e).max() < 1e-12
    e = np.zeros(N)
    e[k0] = N
    assert np.abs(y[2] - e).max() < 1e-11
    rng = np.random.default_rng(1)
    a = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    b = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    fa, fb = O.fft_batched(a), O.fft_batched(b)
    assert np.allclose(O.fft_batched(2.5 * a - 1j * b), 2.5 * fa - 1j * fb, rtol=0, atol=1e-12)   # linearity
    assert abs(np.vdot(fa, fa).real - N * np.vdot(a, a).real) < 1e-10 * N * np.vdot(a, a).real      # Parseval
    assert np.abs(O.fft_batched(fa, 1) / N - a).max() < 1e-14                                       # round trip


def test_twiddles_exact_at_quadrants():
    for M in (4, 8, 16, 512):
        assert O.twiddle(0, M) == complex(1, 0)
        assert O.twiddle(M // 4, M) == complex(0, -1)
        assert O.twiddle(M // 2, M) == complex(-1, 0)
        assert O.twiddle(3 * M // 4, M, +1) == complex(0, -1)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    for M in (512, 2048, 12):
        for m in range(1, M):
            w = O.twiddle(m, M)
            a = -2 * pi * np.longdouble(m) / np.longdouble(M)
            # rounded once from long double: within half a double ulp of the long-double value
            assert abs(w.real - float(np.cos(a))) <= 1.2e-16 and abs(w.imag - float(np.sin(a))) <= 1.2e-16, (M, m)


# ---- 3D transform in the reference stage order ----------------------------

@pytest.mark.parametrize("cfg", [(8, 8, None), (32, 8, None), (12, 6, None), (4, 4, None),
                                 (16, 16, (4, 2, 2)), (32, 64, (4, 4, 4)), (16, 16, (1, 4, 4))],
                         ids=lambda c: f"Ng{c[0]}_P{c[1]}_{c[2]}")
def test_3d_matches_fftn(cfg):
    """The reference's integration configs (run_tests.py:18-21) and more."""
    Ng, P, dims = cfg
    rc, d = O.decompose(Ng, P, dims)
    assert rc == 0
    g = O.fill_uniform(Ng * 7 + P, Ng ** 3).reshape((Ng,) * 3)
    f = O.forward(d, O.global_to_blocks(d, g))
    ref = np.fft.fftn(g)
    assert np.linalg.norm(O.blocks_to_global(d, f) - ref) <= 1e-15 * np.linalg.norm(ref) * 3
    inv = O.blocks_to_global(d, O.inverse(d, f))
    assert np.linalg.norm(inv - g) <= 1e-15 * np.linalg.norm(g) * 3


def test_config_c1_round_trip():
    """C1: 64^3 complex128, U[-1,1) seed 0, dims [2,2,1]: round trip <= 1e-13."""
    rc, d = O.decompose(64, 4, (2, 2, 1))
    g = O.fill_uniform(0, 64 ** 3).reshape((64,) * 3)
    x = O.global_to_blocks(d, g)
    f = O.forward(d, x)
    ref = np.fft.fftn(g)
    assert np.linalg.norm(O.blocks_to_global(d, f) - ref) / np.linalg.norm(ref) < 1e-14
    back = O.inverse(d, f)
    assert np.linalg.norm(back - x) / np.linalg.norm(x) < 1e-13


# ---- generator ---------------------------------------------------------------

def test_philox_known_answers():
    """Random123 kat_vectors for philox4x32_10."""
    assert O.philox((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert O.philox((0xffffffff,) * 4, (0xffffffff, 0xffffffff)) == (0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                      0x6d5451fd)
    assert O.philox((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)) == \
        (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_uniform_and_normal_statistics():
    u = O.fill_uniform(3, 200000)
    assert u.real.min() >= -1 and u.real.max() < 1
    assert abs(u.real.mean()) < 0.01 and abs(u.real.var() - 1 / 3) < 0.01
    n = O.fill_normal(3, 200000)
    assert abs(n.real.mean()) < 0.01 and abs(n.real.std() - 1) < 0.01 and abs(n.imag.std() - 1) < 0.01
