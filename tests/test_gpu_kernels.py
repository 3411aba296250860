"""GPU parity of the element kernels and the hand-written FFT (C ABI calls on
torch-allocated device memory) against the oracle and numpy."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu


def _dev(a, device):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def test_unpack_exhaustive(device):
    import torch
    u = np.arange(65536, dtype=np.uint16)
    pairs = np.stack([u, u[::-1]], axis=1).reshape(-1)
    src = _dev(pairs.view(np.int32), device)
    dst = torch.empty(65536, dtype=torch.complex64, device=device)
    ms.uint16_to_complex(dst, src, 65536)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dst.cpu().numpy(), oracle.unpack_u16(pairs))


def test_pack_bit_exact(device):
    import torch
    rng = np.random.default_rng(1)
    v = (rng.standard_normal(100000) * 20000).astype(np.float32)
    v[:8] = [0.0, 1.9, -1.9, 40000.0, -40000.0, np.nan, 32767.9, -32767.5]
    c = v[0::2] + 1j * v[1::2]
    c = c.astype(np.complex64)
    src = _dev(c, device)
    dst = torch.empty(c.size, dtype=torch.int32, device=device)
    ms.complex_to_uint16(dst, src, c.size)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dst.cpu().numpy().view(np.uint16), oracle.pack_u16(c))


def test_intensity_bit_exact(device):
    import torch
    rng = np.random.default_rng(2)
    a = ((rng.standard_normal(50000) + 1j * rng.standard_normal(50000)) * 300).astype(np.complex64)
    b = ((rng.standard_normal(50000) + 1j * rng.standard_normal(50000)) * 300).astype(np.complex64)
    out = torch.empty(50000, dtype=torch.float32, device=device)
    ms.intensity(_dev(a, device), _dev(b, device), out, 50000)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), oracle.intensity(a, b))


@pytest.mark.parametrize("log2n", list(range(1, 17)) + [18, 20])
def test_fft_c2c_vs_numpy(device, log2n):
    import torch
    n = 1 << log2n
    batch = max(1, min(64, (1 << 21) // n))
    rng = np.random.default_rng(log2n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    xd = _dev(x, device)
    out = torch.empty_like(xd)
    print(f"fft n=2^{log2n} batch={batch}", flush=True)
    ms.fft_c2c(xd, out, n, batch, direction=-1)
    torch.cuda.synchronize()
    ref = np.fft.fft(x.astype(np.complex128), axis=1)
    assert rel_l2(out.cpu().numpy(), ref) < 2e-6
    ms.fft_c2c(xd, out, n, batch, direction=+1)
    torch.cuda.synchronize()
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) * n
    assert rel_l2(out.cpu().numpy(), ref) < 2e-6


def test_fft_overlapped_windows(device):
    # idist = n/2: consecutive transforms read half-overlapping windows (the
    # forward plan of the reference engines, msosm_gpu_batch.cpp:41-48)
    import torch
    n, count = 1024, 9
    rng = np.random.default_rng(5)
    buf = (rng.standard_normal((count + 1) * n // 2) + 1j * rng.standard_normal((count + 1) * n // 2))
    buf = buf.astype(np.complex64)
    out = torch.empty((count, n), dtype=torch.complex64, device=device)
    ms.fft_c2c(_dev(buf, device), out, n, count, idist=n // 2, odist=n)
    torch.cuda.synchronize()
    ref = np.stack([np.fft.fft(buf[b * n // 2:b * n // 2 + n].astype(np.complex128)) for b in range(count)])
    assert rel_l2(out.cpu().numpy(), ref) < 2e-6


def test_fft_two_pass_overlapped(device):
    import torch
    n, count = 1 << 16, 3
    rng = np.random.default_rng(6)
    buf = (rng.standard_normal((count + 1) * n // 2) + 1j * rng.standard_normal((count + 1) * n // 2))
    buf = buf.astype(np.complex64)
    out = torch.empty((count, n), dtype=torch.complex64, device=device)
    ms.fft_c2c(_dev(buf, device), out, n, count, idist=n // 2, odist=n)
    torch.cuda.synchronize()
    ref = np.stack([np.fft.fft(buf[b * n // 2:b * n // 2 + n].astype(np.complex128)) for b in range(count)])
    assert rel_l2(out.cpu().numpy(), ref) < 2e-6


def test_multiply_delay_scatter(device):
    # gpu_kernel.cu:160-197 layout [j/count][dm][j%count][N]
    import torch
    N, count, ndm, index = 64, 4, 3, 4
    rng = np.random.default_rng(7)
    d = rng.integers(0, 5, (ndm, N)).astype(np.int32)
    dbs = 8  # roundup(K + count - 1, count) with K = 5
    a = ((rng.standard_normal((count, N)) + 1j * rng.standard_normal((count, N))) * 100).astype(np.complex64)
    b = (rng.standard_normal((ndm, N)) + 1j * rng.standard_normal((ndm, N))).astype(np.complex64)
    ring = torch.zeros((dbs // count, ndm, count, N), dtype=torch.complex64, device=device)
    ms.multiply_delay_batch(_dev(a, device), _dev(b, device), N, ring, _dev(d, device), dbs, index, count, ndm)
    torch.cuda.synchronize()
    got = ring.cpu().numpy()
    exp = np.zeros_like(got)
    for dm in range(ndm):
        for c in range(count):
            for n in range(N):
                j = (index + c + d[dm, n]) % dbs
                exp[j // count, dm, j % count, n] = a[c, n].astype(np.complex128) * b[dm, n] / N
    assert np.max(np.abs(got - exp)) <= 1e-6 * np.max(np.abs(exp))
    assert np.array_equal(got != 0, exp != 0)


def test_discard_and_multiply(device):
    import torch
    M, count, ndm = 256, 3, 2
    rng = np.random.default_rng(8)
    a = (rng.standard_normal((ndm * count, 2 * M)) + 1j * rng.standard_normal((ndm * count, 2 * M)))
    a = a.astype(np.complex64)
    out = torch.empty((ndm * count, M), dtype=torch.complex64, device=device)
    ms.discard_batch(_dev(a, device), out, M, count, ndm)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), a[:, M:] / np.float32(2 * M))
    b = (rng.standard_normal((ndm, 2 * M)) + 1j * rng.standard_normal((ndm, 2 * M))).astype(np.complex64)
    x = a[:count]
    prod = torch.empty((ndm, count, 2 * M), dtype=torch.complex64, device=device)
    ms.multiply_batch(_dev(x, device), _dev(b, device), 2 * M, prod, count, ndm)
    torch.cuda.synchronize()
    exp = (x[None].astype(np.complex128) / (2 * M)) * b[:, None, :]
    assert rel_l2(prod.cpu().numpy(), exp) < 1e-6


@pytest.mark.parametrize("log2n", [22, 24, 26, 28])
def test_fft_c2c_large_closed_form(device, log2n):
    """Two-pass transforms of 2^22 .. 2^28 points (column passes of 2^11 ..
    2^14 points, the last with 32 points per thread) on two impulses, whose
    spectrum W^(n0 k) + 0.5i W^(n1 k) is known in closed form: 2^16 sampled
    bins against it in FP64 (phases from integer arithmetic mod N), both
    directions."""
    import torch
    n = 1 << log2n
    n0, n1 = (n // 3) | 1, n - 12345
    x = torch.zeros(n, dtype=torch.complex64, device=device)
    x[n0] = 1.0
    x[n1] = 0.5j
    out = torch.empty_like(x)
    rng = np.random.default_rng(log2n)
    ks = np.unique(np.concatenate([rng.integers(0, n, 1 << 16), [0, 1, n // 2, n - 1]]))
    kd = torch.from_numpy(ks).to(device)
    for direction in (-1, +1):
        print(f"fft n=2^{log2n} direction={direction}", flush=True)
        ms.fft_c2c(x, out, n, 1, direction=direction)
        torch.cuda.synchronize()
        got = out[kd].cpu().numpy().astype(np.complex128)
        ph0 = (n0 * ks.astype(object)) % n
        ph1 = (n1 * ks.astype(object)) % n
        e0 = np.exp(direction * 2j * np.pi * ph0.astype(np.float64) / n)
        e1 = np.exp(direction * 2j * np.pi * ph1.astype(np.float64) / n)
        ref = e0 + 0.5j * e1
        err = rel_l2(got, ref)
        print(f"  rel L2 on {ks.size} bins {err:.2e}", flush=True)
        assert err < 2e-6
    del x, out
    torch.cuda.empty_cache()
