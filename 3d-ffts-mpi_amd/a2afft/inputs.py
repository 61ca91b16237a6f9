"""Seeded synthetic inputs and block-layout helpers.

Inputs are defined by the global C-order index g = (x*Ny + y)*Nz + z, so they
do not depend on the decomposition.  The generator is Philox4x32-10 (Salmon,
Moraes, Dror, Shaw, SC'11) with key = (seed lo, seed hi) and counter =
(g lo, g hi, stream, 0); stream 0 draws uniforms, stream 1 normals.
  uniform:  re = ((w0<<32 | w1) >> 11) * 2^-52 - 1,  im from (w2, w3)   in [-1, 1)
  normal:   Box-Muller on u1 = (bits53(w0,w1) + 1) * 2^-53 in (0, 1] and
            u2 = bits53(w2,w3) * 2^-53 in [0, 1)
The reference's own test input (src/a2afft/tests/test.cpp:12-17) draws from an
unseeded rand() identically on every rank with the imaginary part 0, so it is
not reproducible per position and is not used.
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)
_CHUNK = 1 << 22


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32-10 over uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) & _MASK for c in (c0, c1, c2, c3))
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for rnd in range(10):
        if rnd:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c0
        p1 = _M1 * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n1 = p1 & _MASK
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        n3 = p0 & _MASK
        c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def _words(seed, g, stream):
    g = np.asarray(g, dtype=np.uint64)
    return philox4x32_10(g & _MASK, g >> np.uint64(32), np.full_like(g, stream), np.zeros_like(g),
                         seed & 0xFFFFFFFF, seed >> 32)


def _bits53(hi, lo):
    return ((hi << np.uint64(32)) | lo) >> np.uint64(11)


def uniform(seed, n, start=0):
    """n complex128 values, re/im i.i.d. U[-1,1), for global indices start..start+n-1."""
    out = np.empty(n, dtype=np.complex128)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        w0, w1, w2, w3 = _words(seed, np.arange(start + a, start + b, dtype=np.uint64), 0)
        out.real[a:b] = _bits53(w0, w1).astype(np.float64) * 2.0 ** -52 - 1.0
        out.imag[a:b] = _bits53(w2, w3).astype(np.float64) * 2.0 ** -52 - 1.0
    return out


def normal(seed, n, start=0):
    """n complex128 values, re/im i.i.d. N(0,1) by Box-Muller."""
    out = np.empty(n, dtype=np.complex128)
    for a in range(0, n, _CHUNK):
        b = min(n, a + _CHUNK)
        w0, w1, w2, w3 = _words(seed, np.arange(start + a, start + b, dtype=np.uint64), 1)
        u1 = (_bits53(w0, w1) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
        u2 = _bits53(w2, w3).astype(np.float64) * 2.0 ** -53
        r = np.sqrt(-2.0 * np.log(u1))
        out.real[a:b] = r * np.cos(2.0 * np.pi * u2)
        out.imag[a:b] = r * np.sin(2.0 * np.pi * u2)
    return out


def gaussian_random_field(shape, seed, slope=-3.0, noise=1e-3, workers=None):
    """Gaussian random field with power spectrum P(k) ~ |k|^slope (|k| >= 1).

    White noise W(k) from `seed` times A(k) = |k|^(slope/2) (A(0) = 0) on the
    integer wavenumbers k = N*fftfreq(N) of each axis; inverse FFT in FP64;
    scaled to unit RMS; plus noise*(U + iU) from the same seed; complex128.
    """
    import scipy.fft as sfft
    nx, ny, nz = shape
    n = nx * ny * nz
    w = normal(seed, n).reshape(shape)
    kx = np.fft.fftfreq(nx, 1.0 / nx)[:, None, None]
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None]
    kz = np.fft.fftfreq(nz, 1.0 / nz)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    with np.errstate(divide="ignore"):
        amp = np.where(k2 >= 1.0, k2 ** (slope / 4.0), 0.0)
    w *= amp
    del amp, k2
    f = sfft.ifftn(w, workers=workers, overwrite_x=True)
    f /= np.sqrt(np.mean(np.abs(f) ** 2))
    f += noise * uniform(seed, n).reshape(shape)
    return f


def gaussian_random_field_device(shape, seed, slope=-3.0, noise=1e-3, device="cuda"):
    """gaussian_random_field computed on the GPU: the same recipe and the same
    Philox draws (the library's fill_random), the inverse FFT by the library's
    FP64 grid plan.  Returns a complex128 device tensor of `shape` (global C
    order).  It agrees with the host generator to FP64 rounding (the two
    inverse FFTs round differently), which the tests check at 32^3 and 64^3."""
    import torch
    from . import core
    if isinstance(shape, int):
        shape = (shape,) * 3
    shape = tuple(int(s) for s in shape)
    n = shape[0] * shape[1] * shape[2]
    w = torch.empty(n, dtype=torch.complex128, device=device)
    core.fill_random(w, seed, "normal")
    # integer wavenumbers k = N * fftfreq(N) of each axis
    ks = [torch.arange(m, dtype=torch.float64, device=device) for m in shape]
    ks = [torch.where(k < (m + 1) // 2, k, k - m) for k, m in zip(ks, shape)]
    k2 = (ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2).reshape(-1)
    amp = torch.where(k2 >= 1.0, k2.clamp(min=1.0) ** (slope / 4.0), torch.zeros_like(k2))
    w *= amp
    del amp, k2
    w = w.reshape(shape)
    core.GridDfft(shape, "fp64").inverse(w)
    rms = torch.sqrt(torch.mean(w.real ** 2 + w.imag ** 2))
    w /= rms
    u = torch.empty_like(w)
    core.fill_random(u, seed, "uniform")
    w += noise * u
    del u
    return w


# ---- block layout ----------------------------------------------------------

def global_to_blocks(g, dims):
    """(Nx,Ny,Nz) global grid -> (P, nlocal) rank blocks, rank-major (a2adistribution.cpp:14-42)."""
    d0, d1, d2 = dims
    nx, ny, nz = g.shape
    lx, ly, lz = nx // d0, ny // d1, nz // d2
    v = g.reshape(d0, lx, d1, ly, d2, lz)
    if isinstance(v, np.ndarray):
        return np.ascontiguousarray(v.transpose(0, 2, 4, 1, 3, 5)).reshape(d0 * d1 * d2, lx * ly * lz)
    return v.permute(0, 2, 4, 1, 3, 5).contiguous().reshape(d0 * d1 * d2, lx * ly * lz)


def blocks_to_global(b, dims, shape):
    """Inverse of global_to_blocks (tests/stitch.py:28-47)."""
    d0, d1, d2 = dims
    nx, ny, nz = shape
    lx, ly, lz = nx // d0, ny // d1, nz // d2
    v = b.reshape(d0, d1, d2, lx, ly, lz)
    if isinstance(v, np.ndarray):
        return np.ascontiguousarray(v.transpose(0, 3, 1, 4, 2, 5)).reshape(nx, ny, nz)
    return v.permute(0, 3, 1, 4, 2, 5).contiguous().reshape(nx, ny, nz)
