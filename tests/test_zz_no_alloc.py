"""The FFT entry point that allocates nothing (msosm_k_fft_c2c_noalloc): its
output equals msosm_k_fft_c2c's bit for bit (the same kernels, with the two-
pass scratch taken from a fixed device array instead of a per-call buffer),
over two calls in a row (the second smaller), and sizes beyond the fixed
scratch are refused before any launch."""
import ctypes

import numpy as np
import pytest

import msosm_gpu_amd as ms

pytestmark = pytest.mark.gpu

SCRATCH_ELEMS = 1 << 22


def _call(fn, inp, out, n, batch, idist, odist, direction):
    return fn(ctypes.c_void_p(inp.data_ptr()), ctypes.c_void_p(out.data_ptr()), n, batch, idist, odist, direction,
              None)


def _pair(device, log2n, batch, direction, seed, gap=0):
    """Run both entry points on the same input; return (reference, no-alloc)
    outputs as numpy arrays and the FP64 numpy transform."""
    import torch
    L = ms.lib()
    n = 1 << log2n
    idist, odist = n + gap, n + 2 * gap
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal((batch, idist)) + 1j * rng.standard_normal((batch, idist))).astype(np.complex64)
    xd = torch.from_numpy(x).to(device)
    ref = torch.full((batch, odist), 7.0 + 0j, dtype=torch.complex64, device=device)
    got = torch.full((batch, odist), 7.0 + 0j, dtype=torch.complex64, device=device)
    torch.cuda.synchronize()
    print(f"no-alloc fft: n=2^{log2n} batch={batch} direction={direction}", flush=True)
    ms.check(_call(L.msosm_k_fft_c2c, xd, ref, n, batch, idist, odist, direction))
    ms.check(_call(L.msosm_k_fft_c2c_noalloc, xd, got, n, batch, idist, odist, direction))
    torch.cuda.synchronize()
    xs = x[:, :n].astype(np.complex128)
    exact = np.fft.fft(xs, axis=1) if direction < 0 else np.fft.ifft(xs, axis=1) * n
    return ref.cpu().numpy(), got.cpu().numpy(), exact


def test_noalloc_fft_matches_allocating_entry_point(device):
    # first call: the whole scratch (4 x 2^20), forward; second call: a
    # smaller two-pass transform with gaps between rows, inverse
    for log2n, batch, direction, gap in ((20, 4, -1, 0), (16, 3, +1, 64)):
        ref, got, exact = _pair(device, log2n, batch, direction, seed=log2n, gap=gap)
        n = 1 << log2n
        np.testing.assert_array_equal(got, ref)
        # the gap between output rows is left untouched by both
        assert np.all(got[:, n:] == 7.0)
        err = np.linalg.norm(got[:, :n] - exact) / np.linalg.norm(exact)
        assert err < 2e-6, err


def test_noalloc_fft_single_pass_sizes(device):
    # up to 2^14 points the transform needs no scratch at all
    for log2n, batch in ((14, 8), (10, 5)):
        ref, got, _ = _pair(device, log2n, batch, -1, seed=100 + log2n)
        np.testing.assert_array_equal(got, ref)


def test_noalloc_fft_rejects_sizes_beyond_scratch(device):
    import torch
    L = ms.lib()
    # 3 x 2^21 > 2^22 complex samples: refused on the host, nothing launched
    # (the buffer is large enough for the transform all the same)
    x = torch.zeros(3 << 21, dtype=torch.complex64, device=device)
    rc = L.msosm_k_fft_c2c_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()), 1 << 21, 3,
                                   1 << 21, 1 << 21, -1, None)
    assert rc == 1
    assert b"exceeds the fixed scratch" in L.msosm_last_error()
    assert 3 * (1 << 21) > SCRATCH_ELEMS
