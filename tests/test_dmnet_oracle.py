"""Pins the FP64 network oracle (oracle/dmnet_ref.py) with central finite
differences of the loss for every parameter tensor and the input, and its
front end with direct computations."""
import numpy as np

from oracle import dmnet_ref as R


def _net(rng, cin=5, width=4, kw=3, dils=(1, 2), nout=7):
    p = {}
    c = cin
    for i, _ in enumerate(dils):
        p[f"conv{i}_w"] = rng.standard_normal((width, c, kw)) * 0.5
        p[f"conv{i}_b"] = rng.standard_normal(width) * 0.1
        c = width
    p["fc_w"] = rng.standard_normal((nout, 2 * width)) * 0.5
    p["fc_b"] = rng.standard_normal(nout) * 0.1
    return R.Net(p, dils)


def test_gradients_finite_difference():
    rng = np.random.default_rng(0)
    net = _net(rng)
    B, T = 3, 11
    x = rng.standard_normal((B, 5, T))
    tgt = rng.random((B, 7))
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
