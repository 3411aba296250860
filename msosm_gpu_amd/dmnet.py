"""Learned DM proposal: a small 1D-convolution network that ranks DM trials
from a cheap incoherent spectrogram, so only the top-k trials get coherent
dedispersion (the full grid stays available as fallback and judge).

Pipeline (every arithmetic step is a HIP kernel of libmsosm.so; torch only
allocates device memory):
  uint16 baseband -> 256-channel power spectrogram at 4096 samples per bin
  -> per-channel standardisation -> incoherent dedispersion to n_coarse DM
  rows -> row standardisation -> 3 dilated conv layers (ReLU) -> max and
  mean pooling over time -> linear layer -> logits over the fine DM grid.
Weights are FP16 for the kernels with FP32 master copies for Adam.

Training data are simulated pulsars (msosm_sim_* kernels); the label of an
example is the trial with the highest folded S/N found by the plain coherent
method (engine + phase fold) over a local neighbourhood of the fine grid
around the true DM.
"""
import ctypes
import math

import numpy as np

from . import (CHIRP_REF_F32W, FORM_FUSED_F32, Engine, Plan, _ptr, check, fft_c2c, fold_phase, lib)

KDM = 4.15e15


def _p(x):
    return ctypes.c_void_p(_ptr(x))


def snr_of_profile(prof, hits):
    """Folded S/N of a power profile: (peak - median) / (1.4826 MAD) of the
    per-bin mean power."""
    m = prof / np.maximum(hits, 1)
    med = np.median(m, axis=-1, keepdims=True)
    mad = np.median(np.abs(m - med), axis=-1, keepdims=True) * 1.4826
    return ((m.max(axis=-1, keepdims=True) - med) / np.maximum(mad, 1e-30))[..., 0]


class Frontend:
    """Spectrogram -> incoherent dedispersion to n_coarse DM rows."""

    def __init__(self, bw, f0, dm_max, n_coarse, n_samples, device, nchan=256, navg=16):
        import torch
        self.bw, self.f0, self.nchan, self.navg = bw, f0, nchan, navg
        self.T = n_samples // (nchan * navg)
        self.tbin = nchan * navg / bw
        self.dms = np.linspace(0.0, dm_max, n_coarse)
        fc = f0 + np.arange(nchan) * (bw / nchan)          # channel c -> offset c*bw/nchan
        ftop = fc[-1]
        d = np.rint(KDM * self.dms[:, None] * (1 / fc[None, :] ** 2 - 1 / ftop ** 2) / self.tbin).astype(np.int32)
        self.delay_host = d
        self.Tout = self.T - int(d.max())
        if self.Tout < 64:
            raise ValueError("input too short for the DM range")
        self.delay = torch.from_numpy(d).to(device)
        self.S = torch.empty((nchan, self.T), dtype=torch.float32, device=device)
        self.device = device

    def __call__(self, u16_dev, out):
        """u16_dev: packed input (T*nchan*navg samples); out: [n_coarse][Tout]."""
        n = self.T * self.nchan * self.navg
        L = lib()
        check(L.msosm_spectrogram_u16(_p(u16_dev), n, self.nchan, self.navg, _p(self.S), None))
        check(L.msosm_rows_standardize(_p(self.S), self.nchan, self.T, None))
        check(L.msosm_incoherent_dedisperse(_p(self.S), self.nchan, self.T, _p(self.delay), len(self.dms),
                                            self.Tout, _p(out), None))
        check(L.msosm_rows_standardize(_p(out), len(self.dms), self.Tout, None))


class DMNet:
    def __init__(self, c_in, n_out, width=64, kw=9, dils=(1, 2, 4), seed=0, device="cuda:0"):
        import torch
        self.c_in, self.n_out, self.width, self.kw, self.dils = c_in, n_out, width, kw, tuple(dils)
        g = torch.Generator().manual_seed(seed)
        self.shapes = {}
        c = c_in
        for i in range(len(dils)):
            self.shapes[f"conv{i}_w"] = (width, c, kw)
            self.shapes[f"conv{i}_b"] = (width,)
            c = width
        self.shapes["fc_w"] = (n_out, 2 * width)
        self.shapes["fc_b"] = (n_out,)
        self.p, self.p16, self.g, self.m, self.v = {}, {}, {}, {}, {}
        for k, s in self.shapes.items():
            if k.endswith("_w"):
                fan_in = int(np.prod(s[1:]))
                w = torch.randn(s, generator=g) * math.sqrt(2.0 / fan_in)
            else:
                w = torch.zeros(s)
            self.p[k] = w.to(device).contiguous()
            self.g[k] = torch.zeros(s, device=device)
            self.m[k] = torch.zeros(s, device=device)
            self.v[k] = torch.zeros(s, device=device)
            if k.endswith("_w"):
                self.p16[k] = torch.empty(s, dtype=torch.float16, device=device)
                check(lib().msosm_f32_to_f16(_p(self.p[k]), _p(self.p16[k]), self.p[k].numel(), None))
        self.step_no = 0
        self.device = device

    @property
    def n_params(self):
        return int(sum(np.prod(s) for s in self.shapes.values()))

    def forward(self, x):
        """x: [B][c_in][T] FP32 device tensor -> logits [B][n_out]."""
        import torch
        B, C, T = x.shape
        L = lib()
        self.acts = [x]
        h = x
        for i, d in enumerate(self.dils):
            y = torch.empty((B, self.width, T), dtype=torch.float32, device=self.device)
            check(L.msosm_conv1d_fwd(_p(h), _p(self.p16[f"conv{i}_w"]), _p(self.p[f"conv{i}_b"]), _p(y), B,
                                     h.shape[1], self.width, T, self.kw, d, 1, None))
            self.acts.append(y)
            h = y
        self.feat = torch.empty((B, 2 * self.width), dtype=torch.float32, device=self.device)
        self.amax = torch.empty((B, self.width), dtype=torch.int32, device=self.device)
        check(L.msosm_pool_fwd(_p(h), B, self.width, T, _p(self.feat), _p(self.amax), None))
        logits = torch.empty((B, self.n_out), dtype=torch.float32, device=self.device)
        check(L.msosm_linear_fwd(_p(self.feat), _p(self.p16["fc_w"]), _p(self.p["fc_b"]), _p(logits), B,
                                 2 * self.width, self.n_out, None))
        return logits

    def backward(self, glogits, want_input_grad=False):
        import torch
        L = lib()
        for t in self.g.values():
            t.zero_()
        B = glogits.shape[0]
        T = self.acts[0].shape[2]
        gfeat = torch.empty_like(self.feat)
        check(L.msosm_linear_bwd(_p(self.feat), _p(self.p16["fc_w"]), _p(glogits), _p(gfeat), _p(self.g["fc_w"]),
                                 _p(self.g["fc_b"]), B, 2 * self.width, self.n_out, None))
        gh = torch.empty_like(self.acts[-1])
        check(L.msosm_pool_bwd(_p(gfeat), _p(self.amax), B, self.width, T, _p(gh), None))
        for i in reversed(range(len(self.dils))):
            x, y = self.acts[i], self.acts[i + 1]
            gx = torch.empty_like(x) if (i > 0 or want_input_grad) else None
            check(L.msosm_conv1d_bwd(_p(x), _p(self.p16[f"conv{i}_w"]), _p(y), _p(gh), _p(gx) if gx is not None
                                     else None, _p(self.g[f"conv{i}_w"]), _p(self.g[f"conv{i}_b"]), B, x.shape[1],
                                     self.width, T, self.kw, self.dils[i], 1, None))
            gh = gx
        return gh

    def loss_and_grad(self, x, target):
        import torch
        logits = self.forward(x)
        loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        gl = torch.empty_like(logits)
        check(lib().msosm_softmax_xent(_p(logits), _p(target), logits.shape[0], self.n_out, _p(loss), _p(gl), None))
        self.backward(gl)
        return loss

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step_no += 1
        for k in self.p:
            check(lib().msosm_adam(_p(self.p[k]), _p(self.p16[k]) if k in self.p16 else None, _p(self.g[k]),
                                   _p(self.m[k]), _p(self.v[k]), self.p[k].numel(), lr, b1, b2, eps, self.step_no,
                                   None))

    def host_params(self):
        """FP64 copies with the weights rounded to FP16 as the kernels read them."""
        out = {}
        for k in self.p:
            src = self.p16[k] if k in self.p16 else self.p[k]
            out[k] = src.float().cpu().numpy().astype(np.float64)
        return out


# ---------------------------------------------------------------------------
def simulate(n, bw, f0, dm, period_s, duty, snr, seed, device, scale=64.0):
    """Packed uint16 baseband (device tensor) with a dispersed pulsar: noise
    power 1, period-averaged pulsar power snr (see msosm_sim_* kernels)."""
    import torch
    L = lib()
    x = torch.empty(n, dtype=torch.complex64, device=device)
    X = torch.empty_like(x)
    sigma = duty / 2.3548
    gmean = sigma * math.sqrt(2 * math.pi) * math.erf(0.5 / (sigma * math.sqrt(2)))
    amp = snr / max(gmean, 1e-12)
    check(L.msosm_sim_pulsar(_p(x), n, period_s * bw, duty, amp, seed & 0xFFFFFFFF, None))
    fft_c2c(x, X, n, 1, direction=-1)
    check(L.msosm_sim_disperse(_p(X), n, bw, f0, dm, +1, None))
    fft_c2c(X, x, n, 1, direction=+1)
    u = torch.empty(n, dtype=torch.int32, device=device)
    check(L.msosm_sim_quantize(_p(x), n, 1.0 / n, (seed * 2654435761) & 0xFFFFFFFF, scale, _p(u), None))
    return u


def folded_snr(u16_dev, n, bw, f0, dms, period_s, fftpoint=16384, count=256, nbins=64, device="cuda:0"):
    """Plain coherent method: dedisperse u16 at every DM in dms (engine, FP32
    fused form), phase-fold each call's output, return the folded S/N per DM."""
    import torch
    plan = Plan(bw, np.asarray(dms, np.float32), f0, fftpoint=fftpoint, count=count, form=FORM_FUSED_F32,
                chirp=CHIRP_REF_F32W)
    eng = Engine(plan, stream=torch.cuda.current_stream())
    per = plan.count * plan.M
    nd = len(dms)
    prof = torch.zeros((nd, nbins), dtype=torch.float32, device=device)
    hits = torch.zeros(nbins, dtype=torch.float32, device=device)
    for c in range(n // per):
        eng.filter_u16(u16_dev[c * per:(c + 1) * per])
        fold_phase(eng.output_ptr(), per, period_s * bw, nbins, prof, hits, ndm=nd, stride=per, n0=c * per,
                   stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    eng.close()
    return snr_of_profile(prof.cpu().numpy().astype(np.float64), hits.cpu().numpy().astype(np.float64)[None, :])


def soft_target(n_out, idx, sigma=2.0):
    k = np.arange(n_out)
    t = np.exp(-0.5 * ((k - idx) / sigma) ** 2)
    return (t / t.sum()).astype(np.float32)


def topk(logits, k):
    return np.argsort(-logits, axis=-1)[..., :k]


def make_dataset(n_examples, n_samples, bw, f0, dm_max, fine_dms, frontend, period_range, duty_range, snr_range,
                 seed, device, label_halfwidth=8, fftpoint=16384, count=None):
    """Simulate pulsars, run the front end, label each with the plain coherent
    method over fine trials within +-label_halfwidth of the true DM."""
    import torch
    rng = np.random.Generator(np.random.Philox(seed))
    X = torch.empty((n_examples, len(frontend.dms), frontend.Tout), dtype=torch.float32, device=device)
    meta = []
    count = count or max(1, n_samples // (fftpoint // 2) // 4)
    for i in range(n_examples):
        dm = float(rng.uniform(0, dm_max))
        P = float(np.exp(rng.uniform(np.log(period_range[0]), np.log(period_range[1]))))
        duty = float(rng.uniform(*duty_range))
        snr = float(rng.uniform(*snr_range))
        sim_seed = int(rng.integers(1, 2 ** 31))
        u = simulate(n_samples, bw, f0, dm, P, duty, snr, sim_seed, device)
        frontend(u, X[i])
        j0 = int(np.argmin(np.abs(fine_dms - dm)))
        idx = np.arange(max(0, j0 - label_halfwidth), min(len(fine_dms), j0 + label_halfwidth + 1))
        s = folded_snr(u, n_samples, bw, f0, fine_dms[idx], P, fftpoint=fftpoint, count=count, device=device)
        meta.append(dict(dm=dm, period=P, duty=duty, snr=snr, label=int(idx[int(np.argmax(s))]),
                         label_snr=float(s.max()), seed=sim_seed))
    return X, meta


def train(net, X, labels, epochs=10, batch=16, lr=2e-3, sigma=2.0, seed=0):
    import torch
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    T = torch.from_numpy(np.stack([soft_target(net.n_out, l, sigma) for l in labels])).to(X.device)
    losses = []
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        for s in range(0, n - batch + 1, batch):
            idx = torch.from_numpy(order[s:s + batch]).to(X.device)
            loss = net.loss_and_grad(X.index_select(0, idx).contiguous(), T.index_select(0, idx).contiguous())
            net.adam_step(lr=lr)
            tot += float(loss.item())
        losses.append(tot / max(1, n // batch))
    return losses
