"""FP64 numpy restatement of the DM-proposal network and its front end.

TEST INFRASTRUCTURE ONLY: the checker for the HIP forward/backward kernels
(msosm_gpu_amd/csrc/learn.hip).  The network has no counterpart in the
reference (which searches its DM grid exhaustively); what pins this file is a
central finite-difference check of every gradient (tests/test_dmnet_oracle.py)
and, for the front end, direct sums / numpy FFTs.

Layouts match the kernels: activations [B][C][T], conv weights
[Cout][Cin][KW] ("same" padding, dilation), linear weights [O][I].
"""
import numpy as np


def conv1d_fwd(x, w, b, dil, relu):
    B, Cin, T = x.shape
    Cout, _, KW = w.shape
    half = KW // 2
    xp = np.pad(x, ((0, 0), (0, 0), (half * dil, half * dil)))
    y = np.zeros((B, Cout, T)) + b[None, :, None]
    for k in range(KW):
        y += np.einsum("oc,bct->bot", w[:, :, k], xp[:, :, k * dil:k * dil + T])
    return np.maximum(y, 0) if relu else y


def conv1d_bwd(x, w, y, gy, dil, relu):
    B, Cin, T = x.shape
    Cout, _, KW = w.shape
    half = KW // 2
    g = gy * (y > 0) if relu else gy
    xp = np.pad(x, ((0, 0), (0, 0), (half * dil, half * dil)))
    gw = np.zeros_like(w)
    for k in range(KW):
        gw[:, :, k] = np.einsum("bot,bct->oc", g, xp[:, :, k * dil:k * dil + T])
    gb = g.sum(axis=(0, 2))
    gp = np.pad(g, ((0, 0), (0, 0), (half * dil, half * dil)))
    gx = np.zeros_like(x)
    for k in range(KW):
        # x[t'] contributes to y[t' - dil (k - half)]
        gx += np.einsum("oc,bot->bct", w[:, :, k], gp[:, :, (KW - 1 - k) * dil:(KW - 1 - k) * dil + T])
    return gx, gw, gb


def pool_fwd(x):
    am = x.argmax(axis=2)
    feat = np.concatenate([x.max(axis=2), x.mean(axis=2)], axis=1)
    return feat, am


def pool_bwd(gfeat, am, shape):
    B, C, T = shape
    gx = np.repeat((gfeat[:, C:] / T)[:, :, None], T, axis=2)
    bi, ci = np.meshgrid(np.arange(B), np.arange(C), indexing="ij")
    gx[bi, ci, am] += gfeat[:, :C]
    return gx


def linear_fwd(x, w, b):
    return x @ w.T + b[None, :]


def linear_bwd(x, w, gy):
    return gy @ w, gy.T @ x, gy.sum(axis=0)


def softmax_xent(logits, target):
    m = logits.max(axis=1, keepdims=True)
    lse = m + np.log(np.exp(logits - m).sum(axis=1, keepdims=True))
    lp = logits - lse
    B = logits.shape[0]
    loss = -(target * lp).sum() / B
    return loss, (np.exp(lp) - target) / B


class Net:
    """params: dict of float64 arrays conv{i}_w, conv{i}_b and either fc_w,
    fc_b (head "pool": max+mean pooling over time, linear layer) or head_w,
    head_b (head "conv": a final one-channel convolution whose time axis is
    the logits)."""

    def __init__(self, params, dils, head="pool"):
        self.p = params
        self.dils = dils
        self.head = head

    def forward(self, x):
        cache = [x]
        h = x
        for i, d in enumerate(self.dils):
            h = conv1d_fwd(h, self.p[f"conv{i}_w"], self.p[f"conv{i}_b"], d, True)
            cache.append(h)
        if self.head == "conv":
            y = conv1d_fwd(h, self.p["head_w"], self.p["head_b"], 1, False)
            self.cache = (cache, y)
            return y[:, 0, :]
        feat, am = pool_fwd(h)
        logits = linear_fwd(feat, self.p["fc_w"], self.p["fc_b"])
        self.cache = (cache, feat, am)
        return logits

    def backward(self, glogits):
        g = {}
        if self.head == "conv":
            cache, y = self.cache
            gh, g["head_w"], g["head_b"] = conv1d_bwd(cache[-1], self.p["head_w"], y, glogits[:, None, :], 1, False)
        else:
            cache, feat, am = self.cache
            gfeat, g["fc_w"], g["fc_b"] = linear_bwd(feat, self.p["fc_w"], glogits)
            gh = pool_bwd(gfeat, am, cache[-1].shape)
        for i in reversed(range(len(self.dils))):
            gh, g[f"conv{i}_w"], g[f"conv{i}_b"] = conv1d_bwd(cache[i], self.p[f"conv{i}_w"], cache[i + 1], gh,
                                                              self.dils[i], True)
        return g, gh

    def loss(self, x, target):
        return softmax_xent(self.forward(x), target)[0]


def spectrogram(x, nchan=256, navg=16):
    """Power spectrogram [nchan][T] of complex samples x, channel 0 lowest."""
    T = x.size // (nchan * navg)
    X = np.fft.fft(x[:T * nchan * navg].reshape(T, navg, nchan), axis=2)
    p = (np.abs(X) ** 2).sum(axis=1)               # [T][nchan]
    return np.fft.fftshift(p, axes=1).T            # channel c = bin (c + nchan/2) mod nchan


def rows_fold(x, period, nbins):
    """Per row, the mean of the samples in each of nbins phase bins of
    frac(t / period); 0 for an empty bin."""
    T = x.shape[-1]
    ph = np.arange(T) / period
    b = np.minimum(((ph - np.floor(ph)) * nbins).astype(np.int64), nbins - 1)
    cnt = np.bincount(b, minlength=nbins).astype(np.float64)
    out = np.stack([np.bincount(b, weights=r, minlength=nbins) for r in x.reshape(-1, T)])
    out = np.where(cnt > 0, out / np.maximum(cnt, 1), 0.0)
    return out.reshape(x.shape[:-1] + (nbins,))


def incoherent_fold(Sf, delay_s, period):
    """Trial profiles from folded channels: out[k][b] = sum_c Sf[c][(b + s)
    mod nbins] / sqrt(nchan), s = rint(delay_s[k][c] / period * nbins)."""
    nchan, nb = Sf.shape
    out = np.zeros((delay_s.shape[0], nb))
    b = np.arange(nb)
    for k in range(delay_s.shape[0]):
        for c in range(nchan):
            s = int(np.rint(float(delay_s[k, c]) / period * nb)) % nb
            out[k] += Sf[c, (b + s) % nb]
    return out / np.sqrt(nchan)


def harmonics(prof, K):
    """|DFT of each folded profile row| at harmonics 1..K / sqrt(nbins),
    harmonic-major: [K][rows]."""
    nb = prof.shape[-1]
    F = np.fft.fft(prof, axis=-1)[..., 1:K + 1]
    return (np.abs(F) / np.sqrt(nb)).T


def incoherent(S, delay, Tout):
    nchan = S.shape[0]
    out = np.zeros((delay.shape[0], Tout))
    for k in range(delay.shape[0]):
        for c in range(nchan):
            out[k] += S[c, delay[k, c]:delay[k, c] + Tout]
    return out / np.sqrt(nchan)
