"""The reference drivers' simulated signal on the GPU: the dispersed pulse
period against an FP64 numpy restatement of the same steps, and the pulse
coming back as the two unit impulses once the engine dedisperses it (the
simulator's dispersion filter and the engine's dedispersion filter are built
from the same float-rounded chirp grid, with opposite signs)."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu

BW, DM, F0, PERIOD = 16e6, 10.0, 1392e6, 0.005


def _numpy_period(bw, dm, f0, period):
    Np, Nd, rng, fftpoint = ms.refsim_info(bw, dm, f0, period)
    order = fftpoint // 2
    h1 = np.conj(oracle.chirp(bw, dm, f0, order, oracle.CHIRP_REF_F32W))  # exp(+i phi), IDFT input order
    buf = np.zeros(fftpoint, complex)
    buf[order:] = np.fft.ifft(h1)
    H = np.fft.fft(buf)
    acc = np.zeros(Np, complex)
    for i in range(rng):
        x = np.zeros(fftpoint, complex)
        x[i * Np + Np // 2] = 1
        x[i * Np + Np // 2 + 2] = 1
        acc += np.fft.ifft(np.fft.fft(x) * H)[:Np]
    return acc


def test_period_matches_fp64_restatement(device):
    sig, u16, power = ms.refsim_generate(BW, DM, F0, PERIOD, repeat=3)
    Np = ms.refsim_info(BW, DM, F0, PERIOD)[0]
    ref = _numpy_period(BW, DM, F0, PERIOD)
    err = np.linalg.norm(sig[:Np] - ref) / np.linalg.norm(ref)
    print(f"dispersed period vs FP64: rel L2 {err:.2e}, power {power:.4f} dB")
    assert err < 1e-5
    np.testing.assert_array_equal(sig[Np:2 * Np], sig[:Np])
    ref_power = 10 * np.log10(np.sum(np.abs(ref) ** 2) / Np)
    assert abs(power - ref_power) < 1e-3
    v = sig.view(np.float32)
    exp_u16 = np.where(v == 0, 0, (v * np.float32(32767) + np.float32(32768)).astype(np.uint16))
    np.testing.assert_array_equal(u16, exp_u16)


def test_engine_undoes_the_dispersion(device):
    Np, Nd, rng, _ = ms.refsim_info(BW, DM, F0, PERIOD)
    repeat = 8
    _, u16, _ = ms.refsim_generate(BW, DM, F0, PERIOD, repeat=repeat)
    M = 512
    p = ms.Plan(BW, DM, F0, fftpoint=2 * M, count=(repeat * Np) // M, form=ms.FORM_FUSED_F32)
    n = p.count * p.M
    eng = ms.Engine(p)
    eng.filter_u16(u16[:2 * n].view(np.uint32))
    y = np.abs(eng.output_c64()[0].reshape(-1))
    eng.close()
    # five whole periods after the filter transient: per period two impulses
    # 2 samples apart, each close to 32767 / 2 (uint16 scale, output scale
    # 1/2; the multi-segment filter is an approximation of the full-band
    # inverse and the dispersed signal was truncated to uint16, so a few per
    # cent are lost), and nothing else of note
    steady = y[2 * Np:7 * Np]
    peaks = np.flatnonzero(steady > 0.5 * 16383.5)
    assert peaks.size == 10, peaks
    gaps = np.diff(peaks)
    assert np.all(gaps[0::2] == 2) and np.all(gaps[1::2] == Np - 2), gaps
    rel = steady[peaks] / 16383.5
    rest = np.delete(steady, peaks).max() / 16383.5
    print(f"impulses at {rel.min():.3f}..{rel.max():.3f} of 32767/2, largest other sample {rest:.4f}")
    assert np.all((rel > 0.9) & (rel < 1.02)), rel
    assert rest < 0.1
