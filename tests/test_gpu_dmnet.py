"""Learned DM proposal on the GPU: every forward/backward kernel against the
FP64 oracle (oracle/dmnet_ref.py, weights rounded to FP16 as the kernels read
them), the front end against numpy, and an end-to-end test of what the
trained network gains over the same trial budget without it."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
from msosm_gpu_amd import dmnet
from oracle import dmnet_ref as R

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_network_kernels_match_oracle(device):
    import torch
    net = dmnet.DMNet(c_in=24, n_out=40, width=32, kw=9, dils=(1, 2, 4), seed=3, device=device)
    rng = np.random.default_rng(4)
    B, T = 3, 150
    x = rng.standard_normal((B, 24, T)).astype(np.float32)
    tgt = rng.random((B, 40)).astype(np.float32)
    tgt /= tgt.sum(axis=1, keepdims=True)
    xd = torch.from_numpy(x).to(device)
    td = torch.from_numpy(tgt).to(device)
    loss = net.loss_and_grad(xd, td)
    torch.cuda.synchronize()
    ref = R.Net(net.host_params(), (1, 2, 4))
    logits = ref.forward(x.astype(np.float64))
    rloss, gl = R.softmax_xent(logits, tgt.astype(np.float64))
    grads, _ = ref.backward(gl)
    assert abs(float(loss.item()) - rloss) < 1e-4 * abs(rloss)
    for i, a in enumerate(net.acts[1:]):
        assert rel(a.cpu().numpy(), ref.cache[0][i + 1]) < 1e-5
    for k in grads:
        assert rel(net.g[k].cpu().numpy(), grads[k]) < 1e-4, k


def test_input_gradient_and_pool(device):
    import torch
    net = dmnet.DMNet(c_in=8, n_out=16, width=16, kw=5, dils=(1, 3), seed=5, device=device)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((2, 8, 90)).astype(np.float32)
    xd = torch.from_numpy(x).to(device)
    logits = net.forward(xd)
    gl = torch.from_numpy(rng.standard_normal((2, 16)).astype(np.float32)).to(device)
    gx = net.backward(gl, want_input_grad=True)
    torch.cuda.synchronize()
    ref = R.Net(net.host_params(), (1, 3))
    rl = ref.forward(x.astype(np.float64))
    _, rgx = ref.backward(gl.cpu().numpy().astype(np.float64))
    assert rel(logits.cpu().numpy(), rl) < 1e-5
    assert rel(gx.cpu().numpy(), rgx) < 1e-4


def test_frontend_matches_numpy(device):
    import torch
    rng = np.random.default_rng(7)
    n = 256 * 16 * 200
    u = np.clip(np.round(32768 + 700 * rng.standard_normal(2 * n)), 1, 65535).astype(np.uint16)
    fe = dmnet.Frontend(16e6, 1392e6, 30.0, 8, n, device)
    ud = torch.from_numpy(u.view(np.int32)).to(device)
    out = torch.empty((8, fe.Tout), dtype=torch.float32, device=device)
    fe(ud, out)
    torch.cuda.synchronize()
    x = u[0::2].astype(np.float64) - 32768 + 1j * (u[1::2].astype(np.float64) - 32768)
    S = R.spectrogram(x)
    S = (S - S.mean(axis=1, keepdims=True)) / S.std(axis=1, keepdims=True)
    ref = R.incoherent(S, fe.delay_host.astype(np.int64), fe.Tout)
    ref = (ref - ref.mean(axis=1, keepdims=True)) / ref.std(axis=1, keepdims=True)
    assert rel(out.cpu().numpy(), ref) < 1e-4


def test_simulated_pulsar_peaks_at_true_dm(device):
    bw, f0, n = 64e6, 1368e6, 1 << 22
    dm_true, P = 40.0, 0.005
    u = dmnet.simulate(n, bw, f0, dm_true, P, 0.05, 2e-3, 11, device)
    dms = np.array([0.0, 20.0, 35.0, 39.0, 40.0, 41.0, 45.0, 60.0, 90.0], np.float32)
    s = dmnet.folded_snr(u, n, bw, f0, dms, P, count=128, device=device)
    assert dms[int(np.argmax(s))] in (39.0, 40.0, 41.0), s
    assert s.max() > 8


def test_trained_network_gain(device):
    """Train on simulated pulsars (labels from the plain coherent method), then
    propose the top-k of a 256-trial grid for held-out pulsars and compare the
    best folded S/N found with the full-grid best -- against the same budget
    of k trials spread evenly (no network) and an untrained network."""
    import torch
    bw, f0, n = 64e6, 1368e6, 1 << 22
    dm_max, n_fine, n_coarse, k = 100.0, 256, 32, 8
    fine = np.linspace(0, dm_max, n_fine).astype(np.float32)
    fe = dmnet.Frontend(bw, f0, dm_max, n_coarse, n, device)
    kw = dict(period_range=(0.002, 0.01), duty_range=(0.02, 0.08), snr_range=(8e-4, 3e-3), device=device)
    Xtr, mtr = dmnet.make_dataset(480, n, bw, f0, dm_max, fine, fe, seed=101, **kw)
    Xte, mte = dmnet.make_dataset(24, n, bw, f0, dm_max, fine, fe, seed=202, **kw)
    net = dmnet.DMNet(c_in=n_coarse, n_out=n_fine, width=48, kw=9, dils=(1, 2, 4), seed=0, device=device)
    untrained = net.forward(Xte).cpu().numpy()
    losses = dmnet.train(net, Xtr, [m["label"] for m in mtr], epochs=12, batch=16, lr=2e-3)
    assert losses[-1] < 0.8 * losses[0], losses
    logits = net.forward(Xte).cpu().numpy()
    rec_net, rec_even, rec_untr = [], [], []
    even = np.linspace(0, n_fine - 1, k).round().astype(int)
    for i, m in enumerate(mte):
        u = dmnet.simulate(n, bw, f0, m["dm"], m["period"], m["duty"], m["snr"], m["seed"], device)
        full = dmnet.folded_snr(u, n, bw, f0, fine, m["period"], count=128, device=device)
        best = full.max()
        rec_net.append(full[dmnet.topk(logits[i], k)].max() / best)
        rec_even.append(full[even].max() / best)
        rec_untr.append(full[dmnet.topk(untrained[i], k)].max() / best)
    rec_net, rec_even, rec_untr = map(np.array, (rec_net, rec_even, rec_untr))
    print(f"recovered S/N fraction with {k} of {n_fine} trials: network median {np.median(rec_net):.3f} "
          f"(>=0.9 in {np.mean(rec_net >= 0.9):.2f}), evenly spread {np.median(rec_even):.3f}, "
          f"untrained {np.median(rec_untr):.3f}")
    assert np.mean(rec_net >= 0.9) >= 0.75
    assert np.median(rec_net) > np.median(rec_even) + 0.1
    assert np.median(rec_net) > np.median(rec_untr) + 0.1
