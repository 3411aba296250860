"""Parity at the full sizes of the three GPU workloads, through checks that do
not need an FP64 run over the whole stream: output blocks late in the stream
against the FP64 oracle run on just the input those blocks depend on (an
output block g reads input blocks g-K-1 .. g, and the oracle is invariant to
a shift of the stream by whole blocks), agreement between forms for every
trial, and graph replay against eager execution.

  - single trial at production size: 2^28 samples, 128 MHz, DM 100, FFT 2^16
    (K = 152), reference-structured FP32 form
  - many trials: 2^26 samples, 64 MHz, 256 trials on [0, 100], FFT 2^14,
    FP16 (vector ALU and matrix-core inverse) and FP32 fused forms
  - wide delay span: DM up to 500 at 64 MHz, FFT 2^14 (K = 757), fused forms
  - 24 independent streams of different bandwidth (8-64 MHz), length
    (2^20-2^25 samples), trial count (1-64, DM 0-300) and autotuned FFT size,
    each phase-folded into 512 bins, in one hipGraph
"""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def device_input(L, seed, device):
    """2L offset-binary uint16 codes in [1, 65535] packed in L int32 words."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    lo = torch.randint(1, 65536, (L,), generator=g, device=device, dtype=torch.int32)
    hi = torch.randint(1, 65536, (L,), generator=g, device=device, dtype=torch.int32)
    return lo | (hi << 16)


def oracle_blocks(d_in, M, K, b0, nb, bw, dm, f0):
    """FP64 oracle output blocks b0 .. b0+nb of the stream in d_in, from the
    input blocks they depend on."""
    s0 = max(0, b0 - K - 1)
    u = d_in[s0 * M:(b0 + nb) * M].cpu().numpy().view(np.uint16)
    x = oracle.unpack_u16(u).astype(np.complex128)
    d, _, _ = oracle.delay_points(bw, dm, f0, M)
    H = oracle.dedisp_params(bw, dm, f0, M)
    return oracle.msosm_run(x, M, d, H, b0=b0 - s0, nb=nb)


@pytest.mark.timeout(300)
def test_production_single_trial(device):
    import torch
    bw, dm, f0 = 128e6, 100.0, 1336e6
    L = 1 << 28
    p = ms.Plan(bw, dm, f0, fftpoint=65536, count=32, form=ms.FORM_PLAIN_F32)
    assert (p.M, p.K_max) == (32768, 152)
    n = p.count * p.M
    print(f"single trial: {L} samples, M {p.M}, K {p.K_max}, {L // n} calls", flush=True)
    d_in = device_input(L, 0x5EED0002, device)
    eng = ms.Engine(p, stream=torch.cuda.current_stream())
    for c in range(L // n):
        eng.filter_u16(d_in[c * n:(c + 1) * n])
    out = eng.output_c64()[0]  # last call: blocks (L/M - count) .. L/M
    eng.close()
    b_last = L // p.M - p.count
    for k in (0, p.count - 2):
        ref = oracle_blocks(d_in, p.M, p.K_max, b_last + k, 2, bw, dm, f0)
        err = rel_l2(out[k:k + 2].reshape(-1), ref)
        print(f"  blocks {b_last + k}..{b_last + k + 1}: rel L2 vs FP64 {err:.2e}", flush=True)
        assert err < 1e-5


@pytest.mark.timeout(300)
def test_many_trials_full_stream(device):
    import torch
    bw, f0 = 64e6, 1368e6
    dms = np.linspace(0.0, 100.0, 256).astype(np.float32)
    L = 1 << 26
    d_in = device_input(L, 0x5EED0003, device)
    outs = {}
    for form in (ms.FORM_FUSED_F16, ms.FORM_FUSED_F16_MFMA, ms.FORM_FUSED_F32):
        p = ms.Plan(bw, dms, f0, fftpoint=16384, count=64, form=form)
        n = p.count * p.M
        eng = ms.Engine(p, stream=torch.cuda.current_stream())
        for c in range(L // n):
            eng.filter_u16(d_in[c * n:(c + 1) * n])
        y = torch.empty((len(dms), p.count, p.M), dtype=torch.complex64, device=device)
        eng.output_c64(y)
        torch.cuda.synchronize()
        eng.close()
        outs[form] = y
        print(f"form {form}: {L // n} calls done", flush=True)
    a, b = outs[ms.FORM_FUSED_F16], outs[ms.FORM_FUSED_F32]
    per_dm = (torch.linalg.vector_norm((a - b).reshape(len(dms), -1), dim=1) /
              torch.linalg.vector_norm(b.reshape(len(dms), -1), dim=1)).cpu().numpy()
    print(f"FP16 vs FP32 fused, all 256 trials: max rel L2 {per_dm.max():.2e}", flush=True)
    assert per_dm.max() < 1e-3
    c = outs[ms.FORM_FUSED_F16_MFMA]
    per_dm = (torch.linalg.vector_norm((c - b).reshape(len(dms), -1), dim=1) /
              torch.linalg.vector_norm(b.reshape(len(dms), -1), dim=1)).cpu().numpy()
    print(f"FP16 matrix-core vs FP32 fused, all 256 trials: max rel L2 {per_dm.max():.2e}", flush=True)
    assert per_dm.max() < 1e-3
    M, K = 8192, p.K_max
    b_last = L // M - p.count
    for i in (0, 146, 255):
        ref = oracle_blocks(d_in, M, K, b_last + 62, 2, bw, float(dms[i]), f0)
        got = b[i, 62:64].cpu().numpy().reshape(-1)
        err = rel_l2(got, ref)
        print(f"  DM {dms[i]:.2f}: FP32 fused vs FP64, blocks {b_last + 62}..: {err:.2e}", flush=True)
        assert err < 1e-5


@pytest.mark.timeout(300)
def test_wide_delay_span_dm500(device):
    """DM up to 500 at 64 MHz with FFT 2^14: delays of up to K - 1 = 756
    blocks, 2^24 samples, 64 trials; FP16 (vector ALU
    and matrix-core inverse) against FP32 fused for every trial, FP32 against
    the FP64 oracle on the last blocks."""
    import torch
    bw, f0 = 64e6, 1368e6
    dms = np.linspace(0.0, 500.0, 64).astype(np.float32)
    L = 1 << 24
    d_in = device_input(L, 0x5EED0005, device)
    outs = {}
    for form in (ms.FORM_FUSED_F32, ms.FORM_FUSED_F16, ms.FORM_FUSED_F16_MFMA):
        p = ms.Plan(bw, dms, f0, fftpoint=16384, count=64, form=form)
        assert (p.M, p.K_max) == (8192, 757)
        n = p.count * p.M
        eng = ms.Engine(p, stream=torch.cuda.current_stream())
        for c in range(L // n):
            eng.filter_u16(d_in[c * n:(c + 1) * n])
        y = torch.empty((len(dms), p.count, p.M), dtype=torch.complex64, device=device)
        eng.output_c64(y)
        torch.cuda.synchronize()
        eng.close()
        outs[form] = y
        print(f"form {form}: {L // n} calls, ring {p.R} slots", flush=True)
    b = outs[ms.FORM_FUSED_F32]
    for form in (ms.FORM_FUSED_F16, ms.FORM_FUSED_F16_MFMA):
        a = outs[form]
        per_dm = (torch.linalg.vector_norm((a - b).reshape(len(dms), -1), dim=1) /
                  torch.linalg.vector_norm(b.reshape(len(dms), -1), dim=1)).cpu().numpy()
        print(f"form {form} vs FP32 fused, 64 trials: max rel L2 {per_dm.max():.2e}", flush=True)
        assert per_dm.max() < 1e-3
    M, K = 8192, 757
    b_last = L // M - 64
    for i in (0, 37, 63):
        ref = oracle_blocks(d_in, M, K, b_last + 62, 2, bw, float(dms[i]), f0)
        err = rel_l2(b[i, 62:64].cpu().numpy().reshape(-1), ref)
        print(f"  DM {dms[i]:.1f}: FP32 fused vs FP64, blocks {b_last + 62}..: {err:.2e}", flush=True)
        assert err < 1e-5


@pytest.mark.timeout(600)
def test_uneven_batch_24_streams(device):
    import torch
    rng = np.random.Generator(np.random.Philox(0x5EED0004))
    nbins = 512
    streams = []
    for s in range(24):
        bw = float(rng.choice([8e6, 16e6, 32e6, 64e6]))
        f0 = 1400e6 - bw / 2
        L = 1 << int(rng.integers(20, 26))
        ndm = int(rng.integers(1, 65))
        dms = np.sort(rng.uniform(0, 300, ndm)).astype(np.float32)
        period = float(np.exp(rng.uniform(np.log(1e-3), np.log(1.0))))
        tuned = ms.autotune(bw, dms, f0, min_fftpoint=1024, max_fftpoint=65536, target_samples=min(L, 1 << 19),
                            calls=1, form=ms.FORM_FUSED_F16)
        fp, form = tuned["fftpoint"], tuned["form"]
        M = fp // 2
        count = max(1, min(64, L // M))
        streams.append(dict(bw=bw, f0=f0, L=L, dms=dms, period=period * bw, fp=fp, form=form, count=count,
                            tuning=tuned["tuning"]))
        print(f"stream {s}: {bw / 1e6:.0f} MHz, 2^{int(np.log2(L))} samples, {ndm} trials, FFT {fp} "
              f"(form {form}), P {period * 1e3:.1f} ms", flush=True)
    engines, inputs = [], []
    for k, st in enumerate(streams):
        p = ms.Plan(st["bw"], st["dms"], st["f0"], fftpoint=st["fp"], count=st["count"], form=st["form"])
        engines.append(ms.Engine(p))
        engines[-1].set_tuning(**st["tuning"])
        inputs.append(device_input(st["L"], 1000 + k, device))
        st["plan"] = p
    torch.cuda.synchronize()
    batch = ms.Batch(engines, inputs, [st["L"] for st in streams], [st["period"] for st in streams], nbins=nbins)
    batch.launch()
    batch.synchronize()
    print(f"graph with {batch.graph_nodes} nodes replayed", flush=True)
    worst = 0.0
    for k, st in enumerate(streams):
        prof, hits = batch.profiles(k)
        p = st["plan"]
        eng = ms.Engine(p, stream=torch.cuda.current_stream())
        n = p.count * p.M
        rp = torch.zeros((len(st["dms"]), nbins), dtype=torch.float32, device=device)
        rh = torch.zeros(nbins, dtype=torch.float32, device=device)
        for c in range(st["L"] // n):
            eng.filter_u16(inputs[k][c * n:(c + 1) * n])
            ms.fold_phase(eng.output_ptr(), n, st["period"], nbins, rp, rh, ndm=len(st["dms"]), stride=n, n0=c * n,
                          stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(hits, rh.cpu().numpy())
        rpn = rp.cpu().numpy()
        worst = max(worst, float(np.max(np.abs(prof - rpn) / np.maximum(np.abs(rpn), 1e-30))))
        np.testing.assert_array_equal(engines[k].output_c64(), eng.output_c64())
        eng.close()
        # the last call's last block of the highest trial against FP64
        b_last = st["L"] // p.M - 1
        ref = oracle_blocks(inputs[k], p.M, p.K_max, b_last, 1, st["bw"], float(st["dms"][-1]), st["f0"])
        got = engines[k].output_c64()[-1, -1]
        err = rel_l2(got, ref)
        assert err < (1e-3 if st["form"] == ms.FORM_FUSED_F16 else 1e-5), (k, err)
    print(f"all 24 streams: graph = eager (outputs bit-exact, profiles within {worst:.1e}), last blocks vs FP64 ok")
    assert worst < 1e-5
    batch.close()
    for e in engines:
        e.close()
