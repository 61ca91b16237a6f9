"""The allocation-free error reduction (a2afft_error_metrics_noalloc) against
a2afft_error_metrics through the C ABI: two calls in a row, the second on a
smaller input, each bit-identical to the allocating entry point's result."""
import ctypes

import pytest


def _call(lib, name, y, ref, n):
    import torch
    rel, mx = ctypes.c_double(), ctypes.c_double()
    code = 0 if y.dtype == torch.complex128 else 1
    st = getattr(lib, name)(ctypes.c_void_p(y.data_ptr()), code, ctypes.c_void_p(ref.data_ptr()), n,
                            ctypes.byref(rel), ctypes.byref(mx),
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, (name, st, lib.a2afft_last_error())
    return rel.value, mx.value


def test_error_metrics_noalloc_rejects_oversize(a2afft_mod):
    """Beyond 2^36 elements the entry point returns A2AFFT_E_SIZE before any
    device work (host-side check, no GPU needed)."""
    lib = a2afft_mod.lib()
    rel, mx = ctypes.c_double(), ctypes.c_double()
    st = lib.a2afft_error_metrics_noalloc(ctypes.c_void_p(64), 1, ctypes.c_void_p(128), (1 << 36) + 1,
                                          ctypes.byref(rel), ctypes.byref(mx), None)
    assert st == -9, st   # A2AFFT_E_SIZE
    assert b"2^36" in lib.a2afft_last_error()


@pytest.mark.gpu
def test_error_metrics_noalloc_matches(a2afft_mod, cuda):
    import torch
    lib = a2afft_mod.lib()
    assert a2afft_mod.C128 == 0 and a2afft_mod.C64 == 1
    for n, dt, seed in ((1 << 22, torch.complex64, 3), (12345, torch.complex128, 5), (1000, torch.complex64, 7)):
        ref = torch.empty(n, dtype=torch.complex128, device="cuda")
        a2afft_mod.fill_random(ref, seed, "normal")
        noise = torch.empty(n, dtype=torch.complex128, device="cuda")
        a2afft_mod.fill_random(noise, seed + 1, "uniform")
        y = (ref + 1e-6 * noise).to(dt)
        want = _call(lib, "a2afft_error_metrics", y, ref, n)
        got = _call(lib, "a2afft_error_metrics_noalloc", y, ref, n)
        print(f"n={n} {dt}: {got}")
        assert got == want, (n, got, want)
        assert 0 < got[0] < 1e-5
