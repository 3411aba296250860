"""The reference drivers' simulated signal, host side: its geometry (period,
dispersion length, spread and transform size) against the closed forms of
simulated_complex_signal.cpp:21-25,66-68, and its noise against the
sequence that this C++ standard library's default_random_engine and
normal_distribution<float> give in the reference's order
(tests/golden/gen_awgn.cpp wrote the fixtures).  No GPU calls."""
import math
import os

import numpy as np
import pytest

import msosm_gpu_amd as ms

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("bw,dm,f0,period", [(16e6, 10.0, 1392e6, 0.005), (128e6, 75.0, 1e9, 0.032768),
                                             (64e6, 57.3, 1368e6, 0.008)])
def test_geometry_closed_forms(bw, dm, f0, period):
    Np, Nd, rng, fftpoint = ms.refsim_info(bw, dm, f0, period)
    fs = np.float32(bw)
    assert Np == int(np.float32(period) * fs)
    kdm = float(np.float32(4.15e15))
    nd = math.floor(kdm * float(np.float32(dm)) * (1 / float(np.float32(f0)) ** 2 - 1 / float(np.float32(f0) + fs) ** 2)
                    * float(fs))
    assert abs(Nd - nd) <= 1
    assert rng == math.ceil(np.float32(Nd) / np.float32(Np))
    assert fftpoint == 2 * 2 ** math.ceil(math.log2(rng * Np))


@pytest.mark.parametrize("snr", ["10", "-3.5"])
def test_noise_sequence(snr):
    ref = np.loadtxt(os.path.join(GOLDEN, f"awgn_snr{snr}.txt"))
    got = ms.refsim_awgn(float(snr), ref.shape[0])
    np.testing.assert_array_equal(got.real, ref[:, 1].astype(np.float32))
    np.testing.assert_array_equal(got.imag, ref[:, 2].astype(np.float32))


def test_noise_statistics():
    snr = 6.0
    z = ms.refsim_awgn(snr, 200000)
    var = 10 ** (-snr / 10)
    assert abs(np.mean(np.abs(z) ** 2) / var - 1) < 0.01
    assert abs(np.mean(z.real)) < 0.01 and abs(np.mean(z.imag)) < 0.01
