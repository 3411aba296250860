"""Pins the FP64 network oracle (oracle/dmnet_ref.py) with central finite
differences of the loss for every parameter tensor and the input, and its
front end with direct computations."""
import numpy as np
import pytest

from oracle import dmnet_ref as R


def _net(rng, cin=5, width=4, kw=3, dils=(1, 2), nout=7, head="pool"):
    p = {}
    c = cin
    for i, _ in enumerate(dils):
        p[f"conv{i}_w"] = rng.standard_normal((width, c, kw)) * 0.5
        p[f"conv{i}_b"] = rng.standard_normal(width) * 0.1
        c = width
    if head == "conv":
        p["head_w"] = rng.standard_normal((1, width, kw)) * 0.5
        p["head_b"] = rng.standard_normal(1) * 0.1
    else:
        p["fc_w"] = rng.standard_normal((nout, 2 * width)) * 0.5
        p["fc_b"] = rng.standard_normal(nout) * 0.1
    return R.Net(p, dils, head=head)


@pytest.mark.parametrize("head", ["pool", "conv"])
def test_gradients_finite_difference(head):
    rng = np.random.default_rng(0)
    B, T = 3, 11
    # the convolutional head's logits run along the time axis (nout = T)
    net = _net(rng, head=head, nout=7 if head == "pool" else T)
    nout = 7 if head == "pool" else T
    x = rng.standard_normal((B, 5, T))
    tgt = rng.random((B, nout))
    tgt /= tgt.sum(axis=1, keepdims=True)
    logits = net.forward(x)
    loss, gl = R.softmax_xent(logits, tgt)
    grads, gx = net.backward(gl)
    eps = 1e-6
    for name, arr in net.p.items():
        flat = arr.reshape(-1)
        for idx in rng.choice(flat.size, size=min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = net.loss(x, tgt)
            flat[idx] = old - eps
            lm = net.loss(x, tgt)
            flat[idx] = old
            num = (lp - lm) / (2 * eps)
            assert abs(num - grads[name].reshape(-1)[idx]) < 1e-6 + 1e-4 * abs(num), (name, idx)
    xf = x.reshape(-1)
    for idx in rng.choice(xf.size, size=12, replace=False):
        old = xf[idx]
        xf[idx] = old + eps
        lp = net.loss(x, tgt)
        xf[idx] = old - eps
        lm = net.loss(x, tgt)
        xf[idx] = old
        assert abs((lp - lm) / (2 * eps) - gx.reshape(-1)[idx]) < 1e-6


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 3, 20))
    w = rng.standard_normal((4, 3, 5))
    b = rng.standard_normal(4)
    y = R.conv1d_fwd(x, w, b, 3, False)
    for bb in range(2):
        for o in range(4):
            for t in range(20):
                s = b[o]
                for c in range(3):
                    for k in range(5):
                        ts = t + 3 * (k - 2)
                        if 0 <= ts < 20:
                            s += w[o, c, k] * x[bb, c, ts]
                assert abs(y[bb, o, t] - s) < 1e-12


def test_spectrogram_and_incoherent():
    rng = np.random.default_rng(2)
    # a tone in channel bin k lands in channel (k + 128) mod 256
    n = 256 * 16 * 3
    t = np.arange(n)
    x = np.exp(2j * np.pi * 5 * t / 256)
    S = R.spectrogram(x)
    assert S.shape == (256, 3)
    assert np.argmax(S[:, 0]) == (5 + 128) % 256
    S = rng.random((4, 30))
    d = np.array([[0, 1, 2, 3], [3, 2, 1, 0]])
    out = R.incoherent(S, d, 25)
    assert abs(out[0, 4] - (S[0, 4] + S[1, 5] + S[2, 6] + S[3, 7]) / 2) < 1e-12


def test_rows_fold_closed_forms():
    # a periodic impulse train lands in the bin of its phase; a constant row
    # folds to the constant; bins no sample reaches read 0
    T, period, nbins = 10000, 40.0, 16
    x = np.zeros((3, T))
    x[0, 5::40] = 1.0                       # phase 5/40 -> bin 2
    x[1] = 3.25
    x[2, :] = np.arange(T) % 40 < 20        # first half of each period
    y = R.rows_fold(x, period, nbins)
    # bin 2 holds phases [5, 7.5) / 40: samples 5, 6, 7 of each period
    assert y[0].argmax() == 2 and np.isclose(y[0, 2], 1.0 / 3) and np.sum(y[0] > 0) == 1
    np.testing.assert_allclose(y[1], 3.25)
    np.testing.assert_allclose(y[2], [1.0] * 8 + [0.0] * 8)
    z = R.rows_fold(np.ones((1, 3)), 100.0, 8)   # three samples, phases 0, .01, .02
    assert z[0, 0] == 1.0 and np.all(z[0, 1:] == 0)


def test_harmonics_closed_forms():
    # a pure phase-shifted cosine at harmonic h has amplitude A nb / 2 at h
    # only (/ sqrt(nb)); a shift of the profile leaves every amplitude alone;
    # a constant has none
    nb, K = 64, 16
    b = np.arange(nb)
    prof = np.stack([2.0 * np.cos(2 * np.pi * 3 * b / nb + 0.7), np.full(nb, 5.0),
                     np.exp(-0.5 * ((b - 20) / 2.0) ** 2)])
    prof = np.vstack([prof, np.roll(prof[2], 17)])
    h = R.harmonics(prof, K)
    assert h.shape == (K, 4)
    np.testing.assert_allclose(h[2, 0], 2.0 * nb / 2 / np.sqrt(nb), rtol=1e-12)
    assert np.all(np.delete(h[:, 0], 2) < 1e-12)
    assert np.all(h[:, 1] < 1e-12)
    np.testing.assert_allclose(h[:, 2], h[:, 3], rtol=1e-12)


def test_incoherent_fold_equals_folding_the_dedispersed_series():
    # channels periodic with period P = nb * q samples and delays that are
    # whole multiples of q: folding each channel and shifting by its delay in
    # phase bins equals folding the incoherently dedispersed series
    rng = np.random.default_rng(5)
    nchan, nb, q = 6, 16, 5
    P = nb * q
    f = rng.standard_normal((nchan, P))
    T = 40 * P
    S = np.tile(f, (1, T // P))
    d = q * rng.integers(0, 30, (3, nchan))                  # samples
    Tout = T - 10 * P                                        # a whole number of periods
    assert d.max() <= T - Tout
    series = R.incoherent(S, d, Tout)                        # sum_c S[c][t + d] / sqrt(nchan)
    direct = R.rows_fold(series, P, nb)
    Sf = R.rows_fold(S, P, nb)
    np.testing.assert_allclose(R.incoherent_fold(Sf, d * 1.0, float(P)), direct, rtol=1e-12, atol=1e-12)
