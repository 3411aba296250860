"""The delay-line accumulate as per-bin complex GEMMs on the matrix cores
(FORM_HANKEL_*): equality with the one-hot gather form on the same operands
up to the rounding of the stored products, the error of each operand form
against the FP64 same-plan oracle, many DM tiles with their delay bands, and
the refusal of plans whose delay span does not fit."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle
from test_gpu_engine import make_input, oracle_out, rel_l2, run_engine

pytestmark = pytest.mark.gpu

MULTI = (64e6, [0.0, 0.4, 3.0, 7.5, 12.0, 18.25, 24.0, 29.0], 1368e6)


def test_operand_forms_vs_fp64(device):
    """FP16 operands within the mixed gate; BF16 hi+lo about 2^-16 relative.
    (Single BF16 operands, 2^-8 relative rounding, miss the 1e-3 budget --
    which is why the split form exists: the engine refuses that form, and
    test_gemm_kernel_vs_numpy shows its GEMM error.)"""
    bw, dms, f0 = MULTI
    u = make_input(16 * 8192, 99)
    refs = [oracle_out(bw, dm, f0, 8192, u) for dm in dms]
    errs = {}
    for form in (ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16X3):
        print(f"form {form}", flush=True)
        y = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=form), u, device=device)
        errs[form] = max(rel_l2(y[i], refs[i]) for i in range(len(dms)))
    print("max rel L2 vs FP64:", errs)
    assert errs[ms.FORM_HANKEL_F16] < 1e-3
    assert errs[ms.FORM_HANKEL_BF16X3] < 5e-5


def test_equals_gather_form(device):
    """Same FP16 operands as the fused gather form: the only difference is
    the FP16 rounding of the stored products (<= 2^-11 relative each)."""
    bw, dms, f0 = MULTI
    u = make_input(16 * 8192, 5)
    a = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_HANKEL_F16), u, device=device)
    b = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_FUSED_F16), u, device=device)
    for i in range(len(dms)):
        assert rel_l2(a[i], b[i]) < 4e-4, (i, rel_l2(a[i], b[i]))
    c = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_HANKEL_BF16X3), u, device=device)
    d = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_FUSED_F32), u, device=device)
    for i in range(len(dms)):
        assert rel_l2(c[i], d[i]) < 5e-5, (i, rel_l2(c[i], d[i]))


def test_many_trials_delay_bands(device):
    """256 trials on [0, 100] at 64 MHz (K up to 152): 16 DM tiles whose
    delay bands span one or two 16-wide k steps, trials past the last full
    tile, two calls so the Hankel blocks reach into the previous call."""
    bw, f0 = 64e6, 1368e6
    dms = list(np.linspace(0, 100, 250))
    u = make_input(2 * 16 * 8192, 77)
    p16 = ms.Plan(bw, dms, f0, fftpoint=16384, count=16, form=ms.FORM_HANKEL_F16)
    assert p16.K_max == 152
    a = run_engine(p16, u, device=device)
    b = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=16, form=ms.FORM_FUSED_F16), u, device=device)
    worst = max(rel_l2(a[i], b[i]) for i in range(len(dms)))
    print(f"max rel L2 Hankel FP16 vs gather FP16 over {len(dms)} trials: {worst:.2e}")
    assert worst < 4e-4
    for i in (0, 97, 249):
        assert rel_l2(a[i], oracle_out(bw, dms[i], f0, 8192, u)) < 1e-3


def test_refuses_delay_span_beyond_lds(device):
    p = ms.Plan(64e6, [300.0], 1368e6, fftpoint=8192, count=2, form=ms.FORM_HANKEL_F16)
    assert p.K_max == 909
    with pytest.raises(ms.MsosmError):
        ms.Engine(p)


def _half2(z):
    return np.stack([z.real, z.imag], axis=-1).astype(np.float16)


@pytest.mark.parametrize("form", [ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16X3, ms.FORM_HANKEL_BF16])
def test_gemm_kernel_vs_numpy(device, form):
    """The GEMM kernel alone on random operands: 19 trials (a partial 16-trial
    tile), 20 blocks (a partial 16-block tile), random delays (bands of up to
    three 32-deep steps), blocks before the stream start read as zero."""
    import torch
    log2n, count, ndm, kmax, g0 = 7, 20, 19, 37, 30
    N, rx = 1 << log2n, kmax - 1 + count
    rng = np.random.default_rng(form)
    ring = (rng.standard_normal((rx, N)) + 1j * rng.standard_normal((rx, N))) * 3
    H = np.exp(2j * np.pi * rng.random((ndm, N))) * (0.5 + rng.random((ndm, N)))
    d = rng.integers(0, kmax, (ndm, N)).astype(np.uint16)
    d[0, 0] = kmax - 1
    f16 = form == ms.FORM_HANKEL_F16
    if f16:
        ring_h, H_h = _half2(ring), _half2(H)
        ring = ring_h[..., 0].astype(np.float64) + 1j * ring_h[..., 1]
        H = H_h[..., 0].astype(np.float64) + 1j * H_h[..., 1]
        ring_d, H_d = torch.from_numpy(ring_h).to(device), torch.from_numpy(H_h).to(device)
        Y = torch.zeros((ndm, count, N, 2), dtype=torch.float16, device=device)
    else:
        ring_d = torch.from_numpy(ring.astype(np.complex64)).to(device)
        H_d = torch.from_numpy(H.astype(np.complex64)).to(device)
        ring = ring.astype(np.complex64).astype(np.complex128)
        H = H.astype(np.complex64).astype(np.complex128)
        Y = torch.zeros((ndm, count, N), dtype=torch.complex64, device=device)
    d_d = torch.from_numpy(d.view(np.int16)).to(device)
    ms.hankel_delay_gemm(ring_d, H_d, d_d, Y, form, log2n, count, ndm, g0, rx, kmax)
    torch.cuda.synchronize()
    exp = np.zeros((ndm, count, N), complex)
    n = np.arange(N)
    for t in range(ndm):
        for b in range(count):
            a = g0 + b - d[t].astype(np.int64)
            ok = a >= 0
            exp[t, b, ok] = H[t, ok] * ring[np.mod(a, rx)[ok], n[ok]]
    if f16:
        # exact FP16 products summed in FP32, then one FP16 rounding per component
        y = Y.cpu().numpy().astype(np.float64)
        got = y[..., 0] + 1j * y[..., 1]
        err = np.abs(got - exp) / np.maximum(np.abs(exp), 1e-30)
        print(f"form {form}: max relative element error {err[exp != 0].max():.2e}")
        assert err[exp != 0].max() < 2.0 ** -10
    else:
        got = Y.cpu().numpy().astype(np.complex128)
        err = rel_l2(got, exp)
        print(f"form {form}: relative L2 error {err:.2e}")
        assert err < (5e-5 if form == ms.FORM_HANKEL_BF16X3 else 2.0 ** -7)
    assert np.all(got[exp == 0] == 0)


@pytest.mark.parametrize("y_f16", [True, False])
def test_inverse_discard_vs_numpy(device, y_f16):
    import torch
    log2n, units = 10, 5
    N = 1 << log2n
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((units, N)) + 1j * rng.standard_normal((units, N))
    if y_f16:
        Yh = _half2(Y)
        Y = Yh[..., 0].astype(np.float64) + 1j * Yh[..., 1]
        Yd = torch.from_numpy(Yh).to(device)
    else:
        Y = Y.astype(np.complex64).astype(np.complex128)
        Yd = torch.from_numpy(Y.astype(np.complex64)).to(device)
    out = torch.empty((units, N // 2), dtype=torch.complex64, device=device)
    ms.inverse_discard(Yd, out, y_f16, log2n, units, scale=0.25)
    torch.cuda.synchronize()
    ref = 0.25 * np.fft.ifft(Y, axis=1)[:, N // 2:] * N
    assert rel_l2(out.cpu().numpy(), ref) < 2e-6
