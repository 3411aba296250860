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

from . import (CHIRP_REF_F32W, FORM_FUSED_F32, Engine, FftPlan, Plan, _ptr, check, fold_phase, lib)

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
    """Spectrogram -> incoherent dedispersion to DM rows (n_coarse trials on
    [0, dm_max], or the trials `dms`).  Without fold_bins the rows are
    standardised time series [trials][time].  With fold_bins the trials'
    series are instead folded at the pulsar's period (known to the search, as
    to the folding judge) into fold_bins phase bins at unit noise -- computed
    by folding each channel and shifting it by its dispersion delay in phase
    bins, which is the same linear map -- giving [trials][phase]; with
    harmonics as well, each folded profile becomes the amplitudes of its first
    `harmonics` harmonics (independent of the pulse phase), laid out
    [harmonics][trials] so that a convolution runs along the trial axis."""

    def __init__(self, bw, f0, dm_max, n_coarse, n_samples, device, nchan=256, navg=16, fold_bins=0, harmonics=0,
                 dms=None):
        import torch
        self.bw, self.f0, self.nchan, self.navg = bw, f0, nchan, navg
        self.fold_bins, self.harmonics = fold_bins, harmonics
        if harmonics and not fold_bins:
            raise ValueError("harmonic features need folded rows (fold_bins)")
        self.T = n_samples // (nchan * navg)
        self.tbin = nchan * navg / bw
        self.dms = np.linspace(0.0, dm_max, n_coarse) if dms is None else np.asarray(dms, np.float64)
        fc = f0 + np.arange(nchan) * (bw / nchan)          # channel c -> offset c*bw/nchan
        ftop = fc[-1]
        delay_s = KDM * self.dms[:, None] * (1 / fc[None, :] ** 2 - 1 / ftop ** 2)
        d = np.rint(delay_s / self.tbin).astype(np.int32)
        self.delay_host = d
        self.Tfull = self.T - int(d.max())
        self.Tout = len(self.dms) if harmonics else (fold_bins if fold_bins else self.Tfull)
        # network input [channels][steps]: trials x time, trials x phase, or
        # harmonics x trials
        self.shape = (harmonics, len(self.dms)) if harmonics else (len(self.dms), self.Tout)
        if self.Tfull < 64:
            raise ValueError("input too short for the DM range")
        self.delay = torch.from_numpy(d).to(device)
        self.delay_s = torch.from_numpy(delay_s.astype(np.float32)).to(device)
        self.S = torch.empty((nchan, self.T), dtype=torch.float32, device=device)
        self.Sf = torch.empty((nchan, fold_bins), dtype=torch.float32, device=device) if fold_bins else None
        self.prof = (torch.empty((len(self.dms), fold_bins), dtype=torch.float32, device=device) if harmonics
                     else None)
        self.device = device

    def __call__(self, u16_dev, out, period_s=None):
        """u16_dev: packed input (T*nchan*navg samples); out: [n_coarse][Tout].
        period_s: the pulsar's period (needed with fold_bins)."""
        n = self.T * self.nchan * self.navg
        L = lib()
        check(L.msosm_spectrogram_u16(_p(u16_dev), n, self.nchan, self.navg, _p(self.S), None))
        check(L.msosm_rows_standardize(_p(self.S), self.nchan, self.T, None))
        if self.fold_bins:
            if period_s is None:
                raise ValueError("a folding front end needs the period")
            check(L.msosm_rows_fold(_p(self.S), self.nchan, self.T, period_s / self.tbin, self.fold_bins,
                                    math.sqrt(self.T / self.fold_bins), _p(self.Sf), None))
            prof = self.prof if self.harmonics else out
            check(L.msosm_incoherent_fold(_p(self.Sf), self.nchan, self.fold_bins, _p(self.delay_s), len(self.dms),
                                          period_s, _p(prof), None))
            if self.harmonics:
                check(L.msosm_rows_harmonics(_p(prof), len(self.dms), self.fold_bins, self.harmonics, _p(out), None))
            return
        check(L.msosm_incoherent_dedisperse(_p(self.S), self.nchan, self.T, _p(self.delay), len(self.dms),
                                            self.Tout, _p(out), None))
        check(L.msosm_rows_standardize(_p(out), len(self.dms), self.Tout, None))


class DMNet:
    """Dilated 1D convolutions (ReLU) followed by either head "pool" (max and
    mean over time, a linear layer to n_out logits) or head "conv" (a final
    one-channel convolution: the time axis is the trial axis and its values
    are the logits, n_out = the input's length)."""

    def __init__(self, c_in, n_out, width=64, kw=9, dils=(1, 2, 4), seed=0, device="cuda:0", head="pool"):
        import torch
        self.c_in, self.n_out, self.width, self.kw, self.dils = c_in, n_out, width, kw, tuple(dils)
        self.head = head
        g = torch.Generator().manual_seed(seed)
        self.shapes = {}
        c = c_in
        for i in range(len(dils)):
            self.shapes[f"conv{i}_w"] = (width, c, kw)
            self.shapes[f"conv{i}_b"] = (width,)
            c = width
        if head == "conv":
            self.shapes["head_w"] = (1, width, kw)
            self.shapes["head_b"] = (1,)
        else:
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
        if self.head == "conv":
            if T != self.n_out:
                raise ValueError(f"the convolutional head gives one logit per input step ({T}), not {self.n_out}")
            self.hy = torch.empty((B, 1, T), dtype=torch.float32, device=self.device)
            check(L.msosm_conv1d_fwd(_p(h), _p(self.p16["head_w"]), _p(self.p["head_b"]), _p(self.hy), B, self.width,
                                     1, T, self.kw, 1, 0, None))
            return self.hy.view(B, T)
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
        gh = torch.empty_like(self.acts[-1])
        if self.head == "conv":
            gl = glogits.contiguous()
            check(L.msosm_conv1d_bwd(_p(self.acts[-1]), _p(self.p16["head_w"]), _p(self.hy), _p(gl), _p(gh),
                                     _p(self.g["head_w"]), _p(self.g["head_b"]), B, self.width, 1, T, self.kw, 1, 0,
                                     None))
        else:
            gfeat = torch.empty_like(self.feat)
            check(L.msosm_linear_bwd(_p(self.feat), _p(self.p16["fc_w"]), _p(glogits), _p(gfeat),
                                     _p(self.g["fc_w"]), _p(self.g["fc_b"]), B, 2 * self.width, self.n_out, None))
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
_fft_plans = {}


def _fft_plan(n):
    """One cached whole-signal FFT plan per length (its scratch is allocated
    once, not per transform)."""
    if n not in _fft_plans:
        _fft_plans[n] = FftPlan(n)
    return _fft_plans[n]


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
    plan = _fft_plan(n)
    check(L.msosm_sim_pulsar(_p(x), n, period_s * bw, duty, amp, seed & 0xFFFFFFFF, None))
    plan(x, X, direction=-1)
    check(L.msosm_sim_disperse(_p(X), n, bw, f0, dm, +1, None))
    plan(X, x, direction=+1)
    u = torch.empty(n, dtype=torch.int32, device=device)
    check(L.msosm_sim_quantize(_p(x), n, 1.0 / n, (seed * 2654435761) & 0xFFFFFFFF, scale, _p(u), None))
    return u


def power_for_snr(target_snr, n, duty, nbins=64):
    """Period-averaged pulse power (noise power 1) whose folded profile of n
    samples in nbins phase bins peaks target_snr noise sigmas above the
    baseline: the peak-bin excess is the pulse peak power amp, the baseline
    noise of a bin mean is 1/sqrt(n/nbins) (exponentially distributed power).
    A fold of only part of the samples measures sqrt(part) times this."""
    sigma = duty / 2.3548
    gmean = sigma * math.sqrt(2 * math.pi) * math.erf(0.5 / (sigma * math.sqrt(2)))
    return target_snr / math.sqrt(n / nbins) * gmean


class Judge:
    """Plain coherent method as judge: dedisperse at a set of trials of a fixed
    fine DM grid (engine, FP32 fused form unless another form is given) and
    phase-fold every call's output; engines are cached per trial range (the
    max_engines most recently used) so repeated calls only launch kernels."""

    def __init__(self, bw, f0, fine_dms, n_samples, fftpoint=16384, count=32, nbins=64, device="cuda:0",
                 form=FORM_FUSED_F32, max_engines=None):
        from collections import OrderedDict
        self.bw, self.f0, self.fine = bw, f0, np.asarray(fine_dms, np.float32)
        self.n, self.fftpoint, self.count, self.nbins, self.device = n_samples, fftpoint, count, nbins, device
        self.form, self.max_engines = form, max_engines
        self.engines = OrderedDict()
        self.trials_run = 0
        # Fold only output past the longest filter transient of the grid (the
        # first K-1 blocks of an engine ramp up from zero history, which would
        # modulate the folded baseline): the same samples for every trial, as
        # the reference's fold discards Nd samples (msosm_psrdada.cpp:111-112).
        top = Plan(bw, [float(self.fine.max())], f0, fftpoint=fftpoint, count=count, form=FORM_FUSED_F32)
        self.skip = top.K_max * top.M
        if self.skip >= n_samples:
            raise ValueError("input shorter than the filter transient of the largest DM")

    def _engine(self, lo, hi):
        key = (lo, hi)
        if key in self.engines:
            self.engines.move_to_end(key)
        else:
            if self.max_engines is not None:
                while len(self.engines) >= self.max_engines:
                    self.engines.popitem(last=False)[1].close()
            plan = Plan(self.bw, self.fine[lo:hi], self.f0, fftpoint=self.fftpoint, count=self.count,
                        form=self.form, chirp=CHIRP_REF_F32W)
            self.engines[key] = Engine(plan)
        return self.engines[key]

    def snr_range(self, u16_dev, period_s, lo, hi):
        """Folded S/N of fine trials lo..hi-1."""
        import torch
        eng = self._engine(lo, hi)
        eng.reset()
        per = eng.plan.count * eng.plan.M
        nd = hi - lo
        prof = torch.zeros((nd, self.nbins), dtype=torch.float32, device=self.device)
        hits = torch.zeros(self.nbins, dtype=torch.float32, device=self.device)
        torch.cuda.synchronize()  # input and zeroed profiles ready before the engine's stream reads them
        for c in range(self.n // per):
            eng.filter_u16(u16_dev[c * per:(c + 1) * per])
            o = max(0, self.skip - c * per)
            if o < per:
                fold_phase(eng.output_ptr() + 8 * o, per - o, period_s * self.bw, self.nbins, prof, hits, ndm=nd,
                           stride=per, n0=c * per + o, stream=eng.stream)
        eng.synchronize()
        self.trials_run += nd
        return snr_of_profile(prof.cpu().numpy().astype(np.float64), hits.cpu().numpy().astype(np.float64)[None, :])

    def snr_trials(self, u16_dev, period_s, idx, chunk=128):
        """Folded S/N of arbitrary fine trials idx (one engine per chunk of
        consecutive trials of the sorted index set)."""
        idx = np.asarray(idx)
        out = np.empty(len(idx))
        order = np.argsort(idx)
        srt = idx[order]
        i = 0
        while i < len(srt):
            j = i + 1
            while j < len(srt) and srt[j] == srt[j - 1] + 1 and j - i < chunk:
                j += 1
            out[order[i:j]] = self.snr_range(u16_dev, period_s, int(srt[i]), int(srt[j - 1]) + 1)
            i = j
        return out

    def snr_full(self, u16_dev, period_s, chunk=128):
        out = np.empty(len(self.fine))
        for lo in range(0, len(self.fine), chunk):
            hi = min(len(self.fine), lo + chunk)
            out[lo:hi] = self.snr_range(u16_dev, period_s, lo, hi)
        return out

    def close(self):
        for e in self.engines.values():
            e.close()
        self.engines = {}


def soft_target(n_out, idx, sigma=2.0):
    k = np.arange(n_out)
    t = np.exp(-0.5 * ((k - idx) / sigma) ** 2)
    return (t / t.sum()).astype(np.float32)


def topk(logits, k):
    return np.argsort(-logits, axis=-1)[..., :k]


def draw_pulsars(n_examples, dm_max, period_range, duty_range, snr_range, seed):
    """Seeded pulsar parameters: DM uniform, period log-uniform, duty cycle
    uniform, folded S/N uniform, plus a simulation seed per example."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    for _ in range(n_examples):
        dm = float(rng.uniform(0, dm_max))
        P = float(np.exp(rng.uniform(np.log(period_range[0]), np.log(period_range[1]))))
        duty = float(rng.uniform(*duty_range))
        snr = float(rng.uniform(*snr_range))
        out.append(dict(dm=dm, period=P, duty=duty, snr=snr, seed=int(rng.integers(1, 2 ** 31))))
    return out


def make_dataset(pulsars, n_samples, bw, f0, frontend, judge, label_halfwidth=8, label_smooth=0, nbins=64,
                 device="cuda:0", progress=None, label_stride=32, log=print):
    """Simulate each pulsar, run the front end, and label it with the plain
    coherent method: the fine trial with the highest folded S/N within
    +-label_halfwidth trials of the true DM, after a running mean over
    +-label_smooth trials (the S/N of neighbouring trials differs mostly by
    noise when the pulse is wide compared with their smearing difference).
    With progress = k, log() receives a line every k pulsars."""
    import torch
    X = torch.empty((len(pulsars),) + frontend.shape, dtype=torch.float32, device=device)
    nf = len(judge.fine)
    labels = []
    for i, m in enumerate(pulsars):
        u = simulate(n_samples, bw, f0, m["dm"], m["period"], m["duty"],
                     power_for_snr(m["snr"], n_samples, m["duty"], nbins), m["seed"], device)
        frontend(u, X[i], m["period"])
        j0 = int(np.argmin(np.abs(judge.fine - m["dm"])))
        # trial windows on a fixed stride keep the number of cached engines small
        width = min(nf, 2 * label_halfwidth + label_stride)
        lo = min(max(0, (j0 - label_halfwidth) // label_stride * label_stride), nf - width)
        s = judge.snr_range(u, m["period"], lo, lo + width)
        if label_smooth > 0:
            ker = np.ones(2 * label_smooth + 1)
            s = np.convolve(s, ker, "same") / np.convolve(np.ones_like(s), ker, "same")
        a, b = max(j0 - label_halfwidth, lo), min(j0 + label_halfwidth + 1, lo + width)
        labels.append(a + int(np.argmax(s[a - lo:b - lo])))
        if progress and (i + 1) % progress == 0:
            log(f"  simulated and labelled {i + 1}/{len(pulsars)}")
    return X, np.array(labels)


def train(net, X, labels, epochs=10, batch=16, lr=2e-3, sigma=2.0, seed=0, cosine=False, progress=False, log=print):
    """Adam on soft Gaussian targets (sigma in fine trials) around each
    label; with cosine the learning rate decays to zero over the run.  With
    progress, log() receives each epoch's mean loss."""
    import torch
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    T = torch.from_numpy(np.stack([soft_target(net.n_out, l, sigma) for l in labels])).to(X.device)
    losses = []
    total = epochs * (n // batch)
    it = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        steps = 0
        for s in range(0, n - batch + 1, batch):
            idx = torch.from_numpy(order[s:s + batch]).to(X.device)
            loss = net.loss_and_grad(X.index_select(0, idx).contiguous(), T.index_select(0, idx).contiguous())
            net.adam_step(lr=lr * (0.5 * (1 + math.cos(math.pi * it / total)) if cosine else 1.0))
            it += 1
            tot += float(loss.item())
            steps += 1
        losses.append(tot / max(1, steps))
        if progress:
            log(f"  epoch {ep + 1}/{epochs}: loss {losses[-1]:.4f}")
    return losses


class Proposer:
    """Learned DM proposal in front of the plain coherent method: front end ->
    network -> top-k fine trials -> coherent dedispersion + fold of those
    trials only; if the best proposed S/N is below fallback_snr the full grid
    is searched."""

    def __init__(self, net, frontend, judge, k=16, fallback_snr=None):
        self.net, self.frontend, self.judge, self.k, self.fallback_snr = net, frontend, judge, k, fallback_snr

    def search(self, u16_dev, period_s):
        import torch
        x = torch.empty((1,) + self.frontend.shape, dtype=torch.float32, device=self.frontend.device)
        self.frontend(u16_dev, x[0], period_s)
        logits = self.net.forward(x).cpu().numpy()[0]
        idx = topk(logits, self.k)
        s = self.judge.snr_trials(u16_dev, period_s, idx)
        best = int(np.argmax(s))
        if self.fallback_snr is not None and s[best] < self.fallback_snr:
            full = self.judge.snr_full(u16_dev, period_s)
            j = int(np.argmax(full))
            return j, float(full[j]), len(self.judge.fine), True
        return int(idx[best]), float(s[best]), self.k, False
