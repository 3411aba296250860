"""Uneven multi-stream batch in one hipGraph (engines with different
bandwidths, DM trials, FFT sizes and lengths, each call phase-folded) against
the same engines run eagerly, and the FFT-size autotuner."""
import numpy as np
import pytest

import msosm_gpu_amd as ms

pytestmark = pytest.mark.gpu

STREAMS = [
    # bw, f0, dms, fftpoint, count, calls, form, period (samples)
    (8e6, 1396e6, [30.0], 1024, 8, 6, ms.FORM_FUSED_F16, 1777.7),
    (16e6, 1392e6, [10.0, 25.0, 60.0], 2048, 4, 5, ms.FORM_FUSED_F32, 4321.0),
    (32e6, 1384e6, [5.0, 150.0], 4096, 3, 4, ms.FORM_PLAIN_F32, 10000.5),
    (64e6, 1368e6, list(np.linspace(0, 40, 7)), 16384, 2, 3, ms.FORM_FUSED_F16, 64000.25),
]
# the matrix-core and Hankel forms and the single-segment method (FFT size
# 0: the OSM planner's own) in one graph
STREAMS_MORE = [
    (64e6, 1368e6, [0.0, 12.0, 33.0], 16384, 2, 3, ms.FORM_FUSED_F16_MFMA, 50000.5),
    (32e6, 1384e6, [4.0, 20.0, 45.0, 80.0], 8192, 3, 4, ms.FORM_HANKEL_F16, 9000.75),
    (16e6, 1392e6, [7.0, 30.0], 4096, 5, 3, ms.FORM_HANKEL_BF16X3, 3333.3),
    (16e6, 1392e6, [0.5, 2.0], 0, 3, 3, ms.FORM_FUSED_F16, 2500.0, ms.METHOD_OSM),
]


def _plan(st):
    bw, f0, dms, fp, count, calls, form, per = st[:8]
    method = st[8] if len(st) > 8 else ms.METHOD_MSOSM
    return ms.Plan(bw, dms, f0, fftpoint=fp, count=count, form=form, method=method)


def _inputs(device, streams):
    import torch
    out = []
    for k, st in enumerate(streams):
        n = st[4] * _plan(st).M * st[5]
        rng = np.random.Generator(np.random.Philox(100 + k))
        u = np.clip(np.round(32768 + 800 * rng.standard_normal(2 * n)), 1, 65535).astype(np.uint16)
        out.append((u, torch.from_numpy(u.view(np.int32)).to(device)))
    return out


def _eager(device, streams, k, dev_in, nbins):
    import torch
    bw, f0, dms, fp, count, calls, form, per = streams[k][:8]
    p = _plan(streams[k])
    e = ms.Engine(p)
    n = count * p.M
    prof = torch.zeros((len(dms), nbins), dtype=torch.float32, device=device)
    hits = torch.zeros(nbins, dtype=torch.float32, device=device)
    for c in range(calls):
        e.filter_u16(dev_in[c * n:(c + 1) * n])
        e.synchronize()
        ms.fold_phase(e.output_ptr(), n, per, nbins, prof, hits, ndm=len(dms), stride=n, n0=c * n)
        torch.cuda.synchronize()
    last = e.output_c64()
    e.close()
    return prof.cpu().numpy(), hits.cpu().numpy(), last


@pytest.mark.parametrize("streams", [STREAMS, STREAMS_MORE], ids=["fused_plain", "mfma_hankel_osm"])
def test_batch_graph_matches_eager(device, streams):
    import torch
    nbins = 64
    ins = _inputs(device, streams)
    engines, samples, periods = [], [], []
    for st in streams:
        p = _plan(st)
        engines.append(ms.Engine(p))
        samples.append(st[4] * p.M * st[5])
        periods.append(st[7])
    batch = ms.Batch(engines, [d for _, d in ins], samples, periods, nbins=nbins)
    assert batch.graph_nodes > sum(s[5] for s in streams)
    ref = [_eager(device, streams, k, ins[k][1], nbins) for k in range(len(streams))]
    for replay in range(3):
        batch.launch()
        batch.synchronize()
        for k in range(len(streams)):
            prof, hits = batch.profiles(k)
            rprof, rhits, rlast = ref[k]
            np.testing.assert_array_equal(hits, rhits)
            assert np.max(np.abs(prof - rprof) / np.maximum(np.abs(rprof), 1e-30)) < 1e-5, (replay, k)
            np.testing.assert_array_equal(engines[k].output_c64(), rlast)
        # eager work between replays
        x = torch.randn(1 << 20, device=device)
        torch.cuda.synchronize()
        assert float((x * 0).sum()) == 0.0
    batch.close()
    for e in engines:
        e.close()


def test_autotune_picks_a_valid_size(device):
    bw, f0, dms = 64e6, 1368e6, [20.0, 80.0, 300.0]
    best, ns, form = ms.autotune_fftpoint(bw, dms, f0, min_fftpoint=1024, max_fftpoint=16384,
                                          target_samples=1 << 19, calls=2)
    mmin = max(ms.dm_model(bw, d, f0)[1] for d in dms)
    assert best & (best - 1) == 0 and 2 * mmin <= best <= 16384
    assert ns > 0 and form == ms.FORM_FUSED_F16
    again = ms.autotune_fftpoint(bw, dms, f0, min_fftpoint=1024, max_fftpoint=16384, target_samples=1 << 19,
                                 calls=2)
    assert again == (best, ns, form)  # cached


def test_autotune_above_single_workgroup_sizes(device):
    # sizes above 2^14 run the four-step fused form; above 2^21 the
    # reference-structured form
    bw, f0 = 128e6, 1336e6
    best, ns, form = ms.autotune_fftpoint(bw, [100.0], f0, min_fftpoint=32768, max_fftpoint=65536,
                                          target_samples=1 << 20, calls=1)
    assert best in (32768, 65536) and form == ms.FORM_FUSED_F16 and ns > 0
    r = ms.autotune(bw, [100.0], f0, min_fftpoint=1 << 22, max_fftpoint=1 << 22, target_samples=1 << 22, calls=1,
                    form=ms.FORM_FUSED_F32)
    assert r["fftpoint"] == 1 << 22 and r["form"] == ms.FORM_PLAIN_F32


def test_autotune_launch_shapes(device):
    """The tuner times the launch shapes of each size: blocks per workgroup
    for the 2^12..2^14 inverses, tiles per group for the Hankel GEMM; the
    chosen shapes run and give the rule's output bit for bit."""
    import torch
    bw, f0 = 64e6, 1368e6
    dms = list(np.linspace(0, 60, 40))
    for form in (ms.FORM_FUSED_F16, ms.FORM_HANKEL_F16):
        r = ms.autotune(bw, dms, f0, min_fftpoint=8192, max_fftpoint=16384, target_samples=1 << 18, calls=2,
                        form=form)
        print(f"form {form}: {r}", flush=True)
        assert r["form"] == form and r["fftpoint"] in (8192, 16384)
        assert r["candidates"] >= 6
        p = ms.Plan(bw, dms, f0, fftpoint=r["fftpoint"], count=16, form=form)
        n = p.count * p.M
        rng = np.random.Generator(np.random.Philox(5))
        u = np.clip(np.round(32768 + 800 * rng.standard_normal(4 * n)), 1, 65535).astype(np.uint16)
        dev = torch.from_numpy(u.view(np.int32)).to(device)
        outs = []
        for tuned in (False, True):
            e = ms.Engine(p)
            if tuned:
                e.set_tuning(**r["tuning"])
            for c in range(2):
                e.filter_u16(dev[c * n:(c + 1) * n])
            outs.append(e.output_c64())
            e.close()
        np.testing.assert_array_equal(outs[0], outs[1])
