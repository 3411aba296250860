#!/usr/bin/env python
"""Throughput of the MS-OSM coherent-dedispersion engine on one MI355X.

Workload (multi-DM batched precision configuration): 2^26 complex samples of
synthetic 64 MHz baseband (f0 = 1368 MHz, a pulsar at DM 57.3 with P = 8 ms),
256 DM trials evenly spaced on [0, 100] pc cm^-3, FFT size 2^14 (M = 8192,
K = 1..152 segments), count = 64 blocks per call.  One step = one pass of
the whole input through all 256 trials (128 calls), input resident in HBM,
output written per call to the engine's complex64 output buffer.

The headline value is the FP16-storage fused form; the plain FP32 form (the
reference's product-ring structure) is timed beside it for the ratio, and
the matrix-core (per-bin Hankel GEMM) forms for comparison.  Every form is
checked against the FP64 same-plan oracle on a block range.  Per-kernel
rooflines: the dominant kernel of the headline form, the forward FFT, the
plain form's inverse FFT, the GEMM kernel (useful and issued MFMA flops) and
the phase fold over the output.  HBM traffic per launch and the matrix-core
counters come from hardware counters collected live: three rocprofv3 passes
over a child process that replays a short slice of the same workload.

Metric: DM-trial * Msamples / s = num_dms * input_samples / time / 1e6.
"""
import argparse
import csv
import glob
import json
import os
import re
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trial·Msamples/s on 1 MI355X, mixed vs FP32 plain at ≤1e-3 rel err"
UNIT = "DM-trial·Msamples/s"
HBM_PEAK_GBS = 8000.0
MFMA_F16_PEAK_TFLOPS = 2500.0
FORM_NAMES = ("plain_f32", "mixed_f16", "mixed_f16_mfma", "hankel_f16", "hankel_bf16x3")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2-samples", type=int, default=26)
    ap.add_argument("--num-dms", type=int, default=256)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--fftpoint", type=int, default=16384)
    ap.add_argument("--forms", default=",".join(FORM_NAMES),
                    help="comma list of forms to time (mixed_f16 is the headline and always runs)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison")
    ap.add_argument("--fold-bins", type=int, default=512)
    ap.add_argument("--pmc-workload", action="store_true",
                    help="short run for hardware-counter passes: a few calls of each --pmc-forms form, phase "
                         "folds of the last output and a 1 GiB device copy (the counter calibration), no timing")
    ap.add_argument("--pmc-forms", default="mixed_f16", help="comma list of forms the counter run drives")
    ap.add_argument("--pmc-single128", action="store_true",
                    help="with --pmc-workload: drive the 128 MHz single-trial workload instead (DM 100, f0 1336 MHz, "
                         "FFT --fftpoint, --count blocks per call) to count its kernels")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (traffic and matrix-core counters are then null)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the FP64 check of every form (and the comparison with the plain FP32 output)")
    ap.add_argument("--cpu-log2-samples", type=int, default=23)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the short runs of the other workloads (single trial at 16 and 128 MHz, 24 streams)")
    ap.add_argument("--only-config", choices=("single16", "single128", "streams24", "science"),
                    help="run only this one of the other workloads and print its results as one JSON line (no "
                         "headline; for profiling a workload on its own)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine shards over DM trials with no exchange beyond an input "
              "broadcast; this benchmark runs on one GPU only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import msosm_gpu_amd as ms
    from msosm_gpu_amd import simulate

    dev = torch.device("cuda:0")
    ms.lib()
    L = 1 << args.log2_samples
    bw, f0 = 64e6, 1368e6
    dms = np.linspace(0.0, 100.0, args.num_dms).astype(np.float32)
    if args.pmc_workload:
        if args.pmc_single128:
            bw, f0, dms = 128e6, 1336e6, np.array([100.0], np.float32)
        pmc_workload(ms, torch, dev, bw, f0, dms, args)
        return
    if args.only_config:
        print(json.dumps({"other_configs": other_configs(ms, torch, dev, only=args.only_config)}), flush=True)
        return
    t0 = time.time()
    u = simulate.pulsar_u16(L, bw, f0, 57.3, 0.008, duty=0.05, snr=0.1, seed=0x5EED0003)
    log(f"input: {L} samples generated in {time.time() - t0:.1f}s")
    d_in = torch.from_numpy(u.view(np.int32)).to(dev)
    stream = torch.cuda.current_stream()
    results = {}
    checks = {}
    mixed_rows = None
    fp64 = {}
    fp64_ref = None
    plain_ref = None
    form_ids = {"plain_f32": ms.FORM_PLAIN_F32, "mixed_f16": ms.FORM_FUSED_F16,
                "mixed_f16_mfma": ms.FORM_FUSED_F16_MFMA, "hankel_f16": ms.FORM_HANKEL_F16,
                "hankel_bf16x3": ms.FORM_HANKEL_BF16X3}
    wanted = [f for f in args.forms.split(",") if f]
    if "mixed_f16" not in wanted:
        wanted.append("mixed_f16")
    if args.no_plain and "plain_f32" in wanted:
        wanted.remove("plain_f32")
    forms = [(f, form_ids[f]) for f in FORM_NAMES if f in wanted]
    plan_info = None
    fold = None
    pcie = None
    hankel_flops = None
    for name, form in forms:
        plan = ms.Plan(bw, dms, f0, fftpoint=args.fftpoint, count=args.count, form=form)
        n = plan.count * plan.M
        ncalls = L // n
        eng = ms.Engine(plan, stream=stream)
        plan_info = dict(M=plan.M, K_max=plan.K_max, count=plan.count, calls_per_step=ncalls)

        def step():
            for c in range(ncalls):
                eng.filter_u16(d_in[c * n:(c + 1) * n])

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        w0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            step()
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms_step = ev0.elapsed_time(ev1) / args.steps
        value = len(dms) * (ncalls * n) / (ms_step / 1e3) / 1e6
        # per-kernel durations (HIP events around each launch, engine stream)
        eng.set_timing(True)
        step()
        timing = eng.get_timing()
        eng.set_timing(False)
        results[name] = dict(value=value, ms_per_step=ms_step, wall_s=wall, stages={
            k: dict(ms_total=v[0], launches=v[1], ms_per_launch=(v[0] / v[1] if v[1] else 0.0))
            for k, v in timing.items() if v[1]})
        log(f"{name}: {ms_step:.2f} ms/step -> {value:.1f} {UNIT}; stages "
            + ", ".join(f"{k}={v['ms_per_launch']:.3f}ms" for k, v in results[name]["stages"].items()))
        if name == "mixed_f16":
            pcie = host_input_rate(torch, eng, u, n, ncalls, len(dms), args.steps)
            log(f"{name}, input from page-locked host memory: {pcie['ms_per_step']:.2f} ms/step -> "
                f"{pcie['value']:.1f} {UNIT}; H2D copy {pcie['h2d_ms_per_call']:.3f} ms per call")
            fold = fold_bench(ms, torch, eng, plan, stream, 0.008 * bw, args.fold_bins)
            log(f"phase fold ({args.fold_bins} bins): {fold['ms_per_launch']:.3f} ms per launch, "
                f"{fold['achieved']:.0f} GB/s")
        if name.startswith("hankel") and hankel_flops is None:
            hankel_flops = hankel_flop_count(plan)
        if not args.no_check:
            blocks = check_blocks(eng, plan, d_in, n)
            if fp64_ref is None:
                t1 = time.time()
                fp64_ref = fp64_same_plan(eng, plan, host_reader(u), blocks["blocks"])
                log(f"FP64 same-plan reference for {len(dms)} trials, blocks {blocks['blocks']}: "
                    f"{time.time() - t1:.1f}s")
            fp64[name] = compare_fp64(blocks, fp64_ref, dms, GATES[name])
            log(f"{name}: rel L2 vs FP64 over {len(dms)} trials: max {fp64[name]['max']:.3e} "
                f"(DM {fp64[name]['worst_dm']}), median {fp64[name]['median']:.3e}")
            if name == "plain_f32":
                plain_ref = blocks
            elif plain_ref is not None:
                checks[name] = compare_blocks(blocks, plain_ref, dms)
            if name == "mixed_f16":
                mixed_rows = blocks
        eng.close()
        del eng
        torch.cuda.synchronize()

    osm = None
    if fp64_ref is not None and mixed_rows is not None:
        t1 = time.time()
        osm = osm_informational(ms, bw, f0, dms, host_reader(u), mixed_rows, fp64_ref, plan_info["M"])
        log(f"against FP64 single-segment OSM ({time.time() - t1:.1f}s): {json.dumps(osm)}")
    counters, counters_note = None, "skipped (--no-counters)"
    if not args.no_counters:
        t1 = time.time()
        counters, counters_note = counter_passes(args, ",".join(f for f in wanted if f != "plain_f32"))
        log(f"counter passes ({time.time() - t1:.0f}s): {counters_note or 'ok'}")
    head_traffic = hbm_traffic(counters, "k_inv_fused_nb<") if counters else None
    mixed = results["mixed_f16"]
    st = mixed["stages"]["fft_inv"]
    units = len(dms) * plan_info["count"]
    alg_bytes = units * plan_info["M"] * (8 + 8)  # gather 2 FP16 bins + write 1 complex64 per DM-sample
    achieved = alg_bytes / (st["ms_per_launch"] / 1e3) / 1e9
    kname = f"k_inv_fused_nb<{(2 * plan_info['M']).bit_length() - 1},fp16>"
    roofline = dict(bound="hbm", kernel=kname, achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4),
                    traffic=head_traffic["total"] if head_traffic else None,
                    traffic_read_write=[head_traffic["read"], head_traffic["written"]] if head_traffic else None,
                    alg_bytes_per_launch=alg_bytes, ms_per_launch=st["ms_per_launch"])
    kernels = kernel_rooflines(results, plan_info, len(dms), fold, hankel_flops, counters)
    copy_gbs = copy_bandwidth(torch, dev)
    fftcmp = fft_vs_library(ms, torch, dev)
    log(f"batched inverse FFT vs torch.fft (ROCm's library): {json.dumps(fftcmp)}")
    log(f"device copy (1 GiB read + 1 GiB write): {copy_gbs:.0f} GB/s")
    kernels[kname] = {k: roofline[k] for k in ("achieved", "unit", "peak", "frac", "ms_per_launch")}
    if head_traffic:
        kernels[kname]["traffic_bytes"] = head_traffic["total"]
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(bw, f0, dms, args.fftpoint // 2, u, args.cpu_log2_samples)
        log(f"cpu baseline: {cpu}")
    others = None if args.no_configs else other_configs(ms, torch, dev)
    if others:
        log(f"other workloads: {json.dumps(others)}")
    line = {
        "metric": METRIC,
        "value": round(mixed["value"], 2),
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16-storage/f32-compute",
        "storage": "f16 spectra ring and filter table, f32 arithmetic, complex64 output",
        "data": "synthetic (seeded pulsar at DM 57.3 + noise, uint16 offset binary)",
        "config": {"workload": "multi-DM batched: 2^%d complex samples, 64 MHz, f0 1368 MHz, %d DM trials in "
                               "[0,100], FFT 2^14 (M 8192, K 1..%d), count %d"
                               % (args.log2_samples, len(dms), plan_info["K_max"], plan_info["count"]),
                   "input_samples": L, "num_dms": len(dms), "fftpoint": args.fftpoint,
                   "calls_per_step": plan_info["calls_per_step"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": {k: {"value": round(v["value"], 2), "ms_per_step": round(v["ms_per_step"], 3),
                      "stages_ms_per_launch": {s: round(x["ms_per_launch"], 4) for s, x in v["stages"].items()}}
                  for k, v in results.items()},
        "ratio_mixed_vs_plain_f32": (round(mixed["value"] / results["plain_f32"]["value"], 3)
                                     if "plain_f32" in results else None),
        "kernels": kernels,
        "counters": {"source": "rocprofv3 --pmc passes over a 6-call slice of each form in this run",
                     "note": counters_note,
                     "calibration": ({k: head_traffic[k] for k in ("fetch_factor", "write_factor")}
                                     if head_traffic else None)},
        "hbm_copy_gbs": round(copy_gbs, 1),
        "fft_vs_rocm_library": fftcmp,
        "host_input_pcie": pcie,
        "rel_l2_vs_fp64": fp64 or None,
        "rel_l2_vs_fp64_osm": osm,
        "rel_l2_vs_plain_f32": {k: v for k, v in checks.items()},
        "other_configs": others,
    }
    print(json.dumps(line), flush=True)
    if "mixed_f16" in fp64 and not fp64["mixed_f16"]["max"] < GATES["mixed_f16"]:
        log(f"headline form above the {GATES['mixed_f16']:g} gate against FP64")
        sys.exit(1)


def random_u16_words(torch, dev, L, seed):
    """L packed offset-binary uint16 pairs (codes 1..65535) on the device."""
    g = torch.Generator(device=dev).manual_seed(seed)
    lo = torch.randint(1, 65536, (L,), generator=g, device=dev, dtype=torch.int32)
    hi = torch.randint(1, 65536, (L,), generator=g, device=dev, dtype=torch.int32)
    return lo | (hi << 16)


def time_engine(ms, torch, plan, d_in, L, steps=2, warmup=1, gate=None):
    """DM-trial*Msamples/s of one engine over the first whole calls of d_in;
    with a gate, the first two blocks of the last call against the FP64
    execution of the same plan (returned as rel_l2_vs_fp64, with the blocks'
    first sample and rows for further comparisons)."""
    stream = torch.cuda.current_stream()
    eng = ms.Engine(plan, stream=stream)
    n = plan.count * plan.M
    calls = L // n

    def step():
        for c in range(calls):
            eng.filter_u16(d_in[c * n:(c + 1) * n])

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / steps
    res = dict(value=round(plan.num_dms * calls * n / (ms_step / 1e3) / 1e6, 1), ms_per_step=round(ms_step, 3),
               samples=calls * n, fftpoint=2 * plan.M, count=plan.count)
    rows = None
    if gate is not None:
        nb = min(2, plan.count)
        b0 = (calls - 1) * plan.count
        rows = eng.output_c64()[:, :nb].reshape(plan.num_dms, -1).astype(np.complex128)
        ref = fp64_same_plan(eng, plan, device_reader(d_in), (b0, b0 + nb))
        err = max(rel(rows[i], ref[i]) for i in range(plan.num_dms))
        res["rel_l2_vs_fp64"] = dict(max=err, blocks=[b0, b0 + nb], gate=gate, passed=bool(err < gate))
        res["_check"] = (b0 * plan.M, rows, ref)
    eng.close()
    return res


def osm_rows(ms, bw, dm, f0, d_in, forms):
    """rel L2 of each single-trial form's check rows (and of the FP64 same-plan
    MS-OSM rows) against FP64 single-segment OSM over the same samples;
    removes the rows from the results."""
    s0 = min(r["_check"][0] for r in forms.values())
    n = max(r["_check"][0] + r["_check"][1].shape[1] for r in forms.values()) - s0
    o, Mo = osm_fp64(ms, bw, dm, f0, device_reader(d_in), s0, n)
    for r in forms.values():
        a, rows, ref = r.pop("_check")
        seg = o[a - s0:a - s0 + rows.shape[1]]
        r["rel_l2_vs_fp64_osm"] = dict(form=rel(rows[0], seg), msosm_f64=rel(ref[0], seg), osm_fftpoint=2 * Mo)


def other_configs(ms, torch, dev, only=None):
    """Short runs of the other workloads (inputs resident in HBM, uniform
    random codes): a single trial at 16 MHz (DM 10, FFT 1024) and at 128 MHz
    (DM 100 at its plan, FFT 2^16 with K 152, and at FFT 2^14 with K 608), and
    24 independent streams of 8-64 MHz, 2^20-2^25 samples, 1-64 trials,
    autotuned FFT sizes, forms and launch shapes, up to 2^22 samples per call,
    each phase-folded, replayed as one graph, and the reference's science
    plan (FFT 2^21, one block per call).  only = "single16", "single128",
    "streams24" or "science" runs that one alone."""
    out = {}
    if only in (None, "single16"):
        out.update(config_single16(ms, torch, dev))
    if only in (None, "single128"):
        out.update(config_single128(ms, torch, dev))
    if only in (None, "streams24"):
        out.update(config_streams24(ms, torch, dev))
    if only in (None, "science"):
        out.update(config_science(ms, torch, dev))
    return out


def config_science(ms, torch, dev):
    """The plan the reference's science driver uses (msosm_psrdada.cpp:94-95:
    fftpoint 2097152, count 1) on a single 64 MHz trial (DM 100, f0 1368 MHz):
    2^26 samples, 64 calls of one block of 2^20 output samples."""
    L = 1 << 26
    d_in = random_u16_words(torch, dev, L, 0x5EED0005)
    c = {}
    for name, form, gate in (("mixed_f16", ms.FORM_FUSED_F16, 1e-3), ("fused_f32", ms.FORM_FUSED_F32, 1e-5),
                             ("plain_f32", ms.FORM_PLAIN_F32, 1e-5)):
        c[name] = time_engine(ms, torch, ms.Plan(64e6, 100.0, 1368e6, fftpoint=1 << 21, count=1, form=form), d_in,
                              L, steps=2, gate=gate)
        c[name].pop("_check", None)
    del d_in
    torch.cuda.empty_cache()
    return {"single trial, 64 MHz, DM 100, 2^26 samples, FFT 2^21, 1 block per call (the reference's science "
            "plan)": c}


def config_single16(ms, torch, dev):
    out = {}
    L = 1 << 22
    d_in = random_u16_words(torch, dev, L, 0x5EED0001)
    c1 = {}
    for name, form, gate in (("mixed_f16", ms.FORM_FUSED_F16, 1e-3), ("fused_f32", ms.FORM_FUSED_F32, 1e-5),
                             ("plain_f32", ms.FORM_PLAIN_F32, 1e-5)):
        c1[name] = time_engine(ms, torch, ms.Plan(16e6, 10.0, 1392e6, fftpoint=1024, count=1024, form=form), d_in,
                               L, steps=3, gate=gate)
    osm_rows(ms, 16e6, 10.0, 1392e6, d_in, c1)
    out["single trial, 16 MHz, DM 10, 2^22 samples, FFT 1024, 1024 blocks per call"] = c1
    return out


def config_single128(ms, torch, dev):
    out = {}
    L = 1 << 28
    d_in = random_u16_words(torch, dev, L, 0x5EED0002)
    c2 = {"plain_f32, FFT 2^16 (K 152)": time_engine(ms, torch, ms.Plan(128e6, 100.0, 1336e6, fftpoint=65536,
                                                                          count=256, form=ms.FORM_PLAIN_F32),
                                                     d_in, L, steps=2, gate=1e-5)}
    for name, form, gate in (("mixed_f16", ms.FORM_FUSED_F16, 1e-3), ("fused_f32", ms.FORM_FUSED_F32, 1e-5)):
        c2[name + ", FFT 2^16 (K 152)"] = time_engine(ms, torch, ms.Plan(128e6, 100.0, 1336e6, fftpoint=65536,
                                                                         count=256, form=form), d_in, L, steps=2,
                                                      gate=gate)
        c2[name + ", FFT 2^14 (K 608)"] = time_engine(ms, torch, ms.Plan(128e6, 100.0, 1336e6, fftpoint=16384,
                                                                         count=1024, form=form), d_in, L, steps=2,
                                                      gate=gate)
    osm_rows(ms, 128e6, 100.0, 1336e6, d_in, c2)
    out["single trial, 128 MHz, DM 100, 2^28 samples, 2^23 samples per call"] = c2
    del d_in
    torch.cuda.empty_cache()
    return out


def config_streams24(ms, torch, dev):
    out = {}
    rng = np.random.Generator(np.random.Philox(0x5EED0004))
    engines, inputs, lens, periods, shapes, work = [], [], [], [], [], 0
    for s in range(24):
        bw = float(rng.choice([8e6, 16e6, 32e6, 64e6]))
        f0 = 1400e6 - bw / 2
        Ls = 1 << int(rng.integers(20, 26))
        ndm = int(rng.integers(1, 65))
        dms = np.sort(rng.uniform(0, 300, ndm)).astype(np.float32)
        period = float(np.exp(rng.uniform(np.log(1e-3), np.log(1.0))))
        tuned = ms.autotune(bw, dms, f0, min_fftpoint=1024, max_fftpoint=65536, target_samples=min(Ls, 1 << 22),
                            calls=1, form=ms.FORM_FUSED_F16)
        fp, form = tuned["fftpoint"], tuned["form"]
        M = fp // 2
        # up to 2^22 samples per call: small transforms get many blocks per
        # launch instead of many launches
        count = max(1, min(Ls, 1 << 22) // M)
        plan = ms.Plan(bw, dms, f0, fftpoint=fp, count=count, form=form)
        n = plan.count * plan.M
        used = (Ls // n) * n
        engines.append(ms.Engine(plan))
        engines[-1].set_tuning(**tuned["tuning"])
        shapes.append(dict(bw_mhz=bw / 1e6, log2_samples=int(np.log2(Ls)), trials=ndm, fftpoint=fp, form=form,
                           count=count))
        inputs.append(random_u16_words(torch, dev, used, 1000 + s))
        lens.append(used)
        periods.append(period * bw)
        work += ndm * used
    torch.cuda.synchronize()
    batch = ms.Batch(engines, inputs, lens, periods, nbins=512)
    batch.launch()
    batch.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.launch()
    batch.synchronize()
    dt = (time.perf_counter() - t0) / reps
    batch.close()
    # each stream alone, as a graph of its own: where the pass's time goes
    per = []
    for k in range(len(engines)):
        b1 = ms.Batch([engines[k]], [inputs[k]], [lens[k]], [periods[k]], nbins=512)
        b1.launch()
        b1.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            b1.launch()
        b1.synchronize()
        ms1 = (time.perf_counter() - t1) / reps * 1e3
        b1.close()
        per.append(dict(shapes[k], ms=round(ms1, 3),
                        value=round(engines[k].plan.num_dms * lens[k] / (ms1 / 1e3) / 1e6, 1)))
    per.sort(key=lambda r: -r["ms"])
    out["24 streams, 8-64 MHz, up to 2^22 samples per call, one graph with phase folds"] = {
        "value": round(work / dt / 1e6, 1), "ms_per_pass": round(dt * 1e3, 3), "dm_trial_samples": work,
        "graph_nodes": batch.graph_nodes, "streams_alone_ms_sum": round(sum(r["ms"] for r in per), 3),
        "per_stream_alone": per}
    for e in engines:
        e.close()
    return out


# counter passes: one block of counters per rocprofv3 run (FETCH_SIZE and
# WRITE_SIZE each fill most of the TCC block; six SQ counters)
COUNTER_SETS = (("FETCH_SIZE",), ("WRITE_SIZE",),
                ("SQ_INSTS_MFMA", "SQ_VALU_MFMA_BUSY_CYCLES", "SQ_BUSY_CYCLES", "SQ_INSTS_VALU_MFMA_MOPS_F16",
                 "SQ_INSTS_VALU_MFMA_MOPS_BF16", "SQ_WAVES"))


def _kernel_key(name):
    """Short kernel name: template arguments kept, parameter list dropped."""
    n = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\(.*", "", n).replace("msosm::", "").strip()


def counter_passes(args, forms, timeout_s=150):
    """Hardware counters per dispatch for the kernels of `forms` and the phase
    fold: one rocprofv3 --pmc pass per counter set over a child process that
    runs --pmc-workload with this run's workload sizes.  Returns (counters,
    None) with counters = {kernel: {counter: mean per dispatch}} plus the
    calibration copy's FETCH_SIZE and WRITE_SIZE, or (None, reason)."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None, "rocprofv3 not on PATH"
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp", prefix="msosm_pmc_") as td:
        for i, cs in enumerate(COUNTER_SETS):
            d = os.path.join(td, f"p{i}")
            cmd = [exe, "--pmc", *cs, "--output-format", "csv", "-d", d, "-o", "pmc", "--", sys.executable, "-u",
                   os.path.abspath(__file__), "--pmc-workload", "--pmc-forms", forms, "--num-dms", str(args.num_dms),
                   "--count", str(args.count), "--fftpoint", str(args.fftpoint), "--fold-bins", str(args.fold_bins)]
            log(f"counter pass {i + 1}/{len(COUNTER_SETS)}: {' '.join(cs)}")
            with open(os.path.join(td, f"p{i}.log"), "w") as lf:
                p = subprocess.Popen(cmd, cwd="/tmp", env=dict(os.environ, TMPDIR="/tmp"), stdout=lf,
                                     stderr=subprocess.STDOUT, start_new_session=True)
                try:
                    rc = p.wait(timeout=timeout_s)
                except subprocess.TimeoutExpired:
                    os.killpg(p.pid, signal.SIGKILL)
                    p.wait()
                    return None, f"counter pass {' '.join(cs)} exceeded {timeout_s} s"
            text = open(os.path.join(td, f"p{i}.log")).read()
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            # the profiler may end with its own fault after the program has
            # printed its last line and the files are whole: accept that
            if not files or '"pmc_workload"' not in text:
                return None, f"counter pass {' '.join(cs)} failed (rc {rc}): {text[-300:]}"
            rows = list(csv.DictReader(open(files[0])))
            copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
            big = max((int(r["Grid_Size"]) for r in copies), default=0)
            for r in rows:
                k = "calibration copy" if ("copyBuffer" in r["Kernel_Name"] and int(r["Grid_Size"]) == big) \
                    else _kernel_key(r["Kernel_Name"])
                out.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in out.items()}, None


def hbm_traffic(counters, prefix, copy_bytes=1 << 30):
    """HBM bytes per dispatch of the kernels whose short name starts with
    `prefix`: FETCH_SIZE and WRITE_SIZE (KiB) divided by what the same
    counters read for the 1 GiB calibration copy (read and write 1 GiB)."""
    cal = counters.get("calibration copy", {})
    if not cal.get("FETCH_SIZE") or not cal.get("WRITE_SIZE"):
        return None
    fetch = [v["FETCH_SIZE"] for k, v in counters.items() if k.startswith(prefix) and "FETCH_SIZE" in v]
    write = [v["WRITE_SIZE"] for k, v in counters.items() if k.startswith(prefix) and "WRITE_SIZE" in v]
    if not fetch or not write:
        return None
    kib = copy_bytes / 1024
    rf, wf = cal["FETCH_SIZE"] / kib, cal["WRITE_SIZE"] / kib
    read, written = fetch[0] * 1024 / rf, write[0] * 1024 / wf
    return dict(read=int(read), written=int(written), total=int(read + written), fetch_factor=round(rf, 4),
                write_factor=round(wf, 4))


def mfma_counters(counters, key, num_se=32, simds_per_se=32):
    """Matrix-core counters per dispatch of kernel `key`: MFMA instructions,
    counted flops (MOPS counters x 512), and the MFMA busy fraction =
    SQ_VALU_MFMA_BUSY_CYCLES / (SIMD count x per-engine busy cycles), the
    latter SQ_BUSY_CYCLES summed over the 32 shader engines / 32."""
    c = counters.get(key)
    if not c or "SQ_INSTS_MFMA" not in c:
        return None
    flops = 512 * (c.get("SQ_INSTS_VALU_MFMA_MOPS_F16", 0) + c.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0))
    busy = c["SQ_VALU_MFMA_BUSY_CYCLES"] / (num_se * simds_per_se * c["SQ_BUSY_CYCLES"] / num_se)
    return dict(mfma_insts=int(c["SQ_INSTS_MFMA"]), mfma_flops_counted=int(flops), mfma_busy_frac=round(busy, 4))


def pmc_workload(ms, torch, dev, bw, f0, dms, args, calls=6, folds=4):
    """Dispatches for counter passes (FETCH_SIZE, WRITE_SIZE, SQ_*): for each
    form of --pmc-forms, `calls` calls on random offset-binary input and
    `folds` phase folds of the last output (one launch per form: every trial,
    count*M samples, --fold-bins bins); then two copies of 1 GiB between
    device buffers -- a kernel whose HBM bytes are known (read 1 GiB, write
    1 GiB), giving the counters' scale on this device."""
    form_ids = {"plain_f32": ms.FORM_PLAIN_F32, "mixed_f16": ms.FORM_FUSED_F16,
                "mixed_f16_mfma": ms.FORM_FUSED_F16_MFMA, "hankel_f16": ms.FORM_HANKEL_F16,
                "hankel_bf16x3": ms.FORM_HANKEL_BF16X3}
    g = torch.Generator(device=dev).manual_seed(7)
    for name in [f for f in args.pmc_forms.split(",") if f]:
        plan = ms.Plan(bw, dms, f0, fftpoint=args.fftpoint, count=args.count, form=form_ids[name])
        n = plan.count * plan.M
        lo = torch.randint(1, 65536, (calls * n,), generator=g, device=dev, dtype=torch.int32)
        hi = torch.randint(1, 65536, (calls * n,), generator=g, device=dev, dtype=torch.int32)
        d_in = lo | (hi << 16)
        del lo, hi
        stream = torch.cuda.current_stream()
        eng = ms.Engine(plan, stream=stream)
        print(f"pmc workload: {name}, {calls} calls", flush=True)
        for c in range(calls):
            eng.filter_u16(d_in[c * n:(c + 1) * n])
        prof = torch.zeros((plan.num_dms, args.fold_bins), dtype=torch.float32, device=dev)
        hits = torch.zeros(args.fold_bins, dtype=torch.float32, device=dev)
        for i in range(folds):
            ms.fold_phase(eng.output_ptr(), n, 0.008 * bw, args.fold_bins, prof, hits, ndm=plan.num_dms, stride=n,
                          n0=i * n, stream=stream)
        torch.cuda.synchronize()
        eng.close()
        del d_in
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"pmc_workload": {"forms": args.pmc_forms, "calls": calls, "folds": folds,
                                       "units_per_call": len(dms) * plan.count, "copy_bytes": src.numel() * 4}}),
          flush=True)


def copy_bandwidth(torch, dev, reps=5):
    """Achievable HBM bandwidth on this device: a 1 GiB device-to-device copy,
    (read + write bytes) / time, best of reps."""
    src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        best = max(best, 2 * src.numel() * 4 / (e0.elapsed_time(e1) / 1e3) / 1e9)
    del src, dst
    return best


def fft_vs_library(ms, torch, dev, shapes=((14, 16384), (16, 1024)), reps=10):
    """The build's batched C2C inverse (msosm_k_fft_c2c) beside ROCm's FFT
    library as torch.fft.ifft exposes it (a labelled comparator, not the
    product path), same complex64 input: time per batch and the two outputs'
    relative difference.  16384 x 2^14 is one call of the headline workload's
    inverse transforms (256 trials x 64 blocks)."""
    res = {}
    for log2n, batch in shapes:
        n = 1 << log2n
        g = torch.Generator(device=dev).manual_seed(log2n)
        x = torch.randn(batch, n, dtype=torch.complex64, device=dev, generator=g)
        out = torch.empty_like(x)

        def timed(fn):
            for _ in range(2):
                fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            return e0.elapsed_time(e1) / reps

        ours = timed(lambda: ms.fft_c2c(x, out, n, batch, direction=+1))
        holder = {}
        lib = timed(lambda: holder.__setitem__("y", torch.fft.ifft(x, dim=1)))
        diff = (torch.linalg.norm(out / n - holder["y"]) / torch.linalg.norm(holder["y"])).item()
        gb = batch * n * 16 / 1e9
        res[f"2^{log2n} x {batch}"] = dict(ours_ms=round(ours, 4), torch_fft_ms=round(lib, 4),
                                           ours_gbs=round(gb / ours * 1e3, 1), torch_fft_gbs=round(gb / lib * 1e3, 1),
                                           rel_diff=float(f"{diff:.3g}"))
        del x, out, holder
    torch.cuda.empty_cache()
    return res


def host_input_rate(torch, eng, u, n, ncalls, ndm, steps):
    """The same steps with the input handed over in page-locked host memory:
    each call copies its count*M packed samples (4 B each) over PCIe before
    its kernels run.  Wall-clock time, as the host waits for every copy."""
    h_in = torch.from_numpy(u.view(np.int32)).pin_memory()

    def step():
        for c in range(ncalls):
            eng.filter_u16(h_in[c * n:(c + 1) * n])

    step()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t) * 1e3 / steps
    eng.set_timing(True)
    step()
    timing = eng.get_timing()
    eng.set_timing(False)
    c = timing.get("h2d", (0.0, 0))
    del h_in
    return dict(value=round(ndm * ncalls * n / (ms_step / 1e3) / 1e6, 2), ms_per_step=round(ms_step, 3),
                h2d_ms_per_call=round(c[0] / c[1], 4) if c[1] else None,
                input="page-locked host memory, 4 B per sample copied host to device each call")


def fold_bench(ms, torch, eng, plan, stream, period_samples, nbins, launches=20):
    """Phase fold of the engine's last output (all trials, count*M samples
    each) into nbins bins, timed with HIP events on the engine stream;
    algorithmic bytes = 8 B (complex64) read per DM-sample."""
    n = plan.count * plan.M
    prof = torch.zeros((plan.num_dms, nbins), dtype=torch.float32, device="cuda:0")
    hits = torch.zeros(nbins, dtype=torch.float32, device="cuda:0")
    ptr = eng.output_ptr()
    ms.fold_phase(ptr, n, period_samples, nbins, prof, hits, ndm=plan.num_dms, stride=n, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(launches):
        ms.fold_phase(ptr, n, period_samples, nbins, prof, hits, ndm=plan.num_dms, stride=n, n0=i * n, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / launches
    alg = plan.num_dms * n * 8
    return dict(ms_per_launch=t, alg_bytes_per_launch=alg, achieved=alg / (t / 1e3) / 1e9)


def hankel_flop_count(plan):
    """Useful and issued flops of one GEMM launch.  Useful: one complex MAC
    (8 real flops) per (trial, bin, block).  Issued: per (bin, 16-trial tile,
    16-block tile) the 32-deep MFMA steps that cover the tile's delay band,
    2 MFMAs (real, imaginary) of 16x16x32 each; the dense recast would issue
    8 K_max per (trial, bin, block)."""
    N, count, ndm = plan.fftpoint, plan.count, plan.num_dms
    d = np.stack([plan.delay_points(i) for i in range(ndm)])  # [ndm][N]
    steps = 0
    for t0 in range(0, ndm, 16):
        blk = d[t0:t0 + 16]
        steps += int(((blk.max(axis=0) - blk.min(axis=0)) // 16 + 1).sum())
    gtiles = (count + 15) // 16
    issued_per_term = steps * gtiles * 2 * (16 * 16 * 32 * 2)
    useful = 8 * ndm * N * count
    return dict(useful=useful, issued_per_term=issued_per_term, dense=8 * plan.K_max * ndm * N * count)


def kernel_rooflines(results, plan_info, ndm, fold, hankel_flops, counters=None):
    """Per-kernel rooflines from the HIP-event stage times; where the counter
    passes ran, each entry also carries its counted HBM bytes per launch
    (traffic_bytes) and, for matrix-core kernels, the MFMA counters."""
    M, count = plan_info["M"], plan_info["count"]
    N = 2 * M
    out = {}

    def put(name, alg, ms_launch, unit="GB/s", peak=HBM_PEAK_GBS, ckey=None, **extra):
        if not ms_launch:
            return
        a = alg / (ms_launch / 1e3) / (1e9 if unit == "GB/s" else 1e12)
        out[name] = dict(achieved=round(a, 1), unit=unit, peak=peak, frac=round(a / peak, 4),
                         ms_per_launch=round(ms_launch, 4), **extra)
        if counters and ckey:
            tr = hbm_traffic(counters, ckey)
            if tr:
                out[name]["traffic_bytes"] = tr["total"]
            mf = mfma_counters(counters, ckey)
            if mf and mf["mfma_insts"]:
                out[name].update(mf)
                out[name]["mfma_flops_counted_tflops"] = round(mf["mfma_flops_counted"] / (ms_launch / 1e3) / 1e12, 1)

    st = results["mixed_f16"]["stages"]
    # forward FFT: count windows of N points, uint16 in (4 B) and FP16 out (4 B)
    put("k_fwd_u16<14,fp16>", count * N * 8, st["fft_fwd"]["ms_per_launch"], ckey="k_fwd_u16<14, true, false>")
    if "plain_f32" in results:
        sp = results["plain_f32"]["stages"]
        put("k_fft_c2c<14,inverse> (plain)", ndm * count * N * 16, sp["fft_inv"]["ms_per_launch"])
        # per (trial, bin): read H, write the ring slot (8 B each); the spectrum is shared by all trials
        put("k_multiply_delay (plain)", ndm * count * N * 16, sp["multiply"]["ms_per_launch"])
    if "mixed_f16_mfma" in results:
        # same algorithmic bytes as the vector-ALU fused inverse
        put("k_inv_mfma<14> (matrix-core DFTs)", ndm * count * M * 16,
            results["mixed_f16_mfma"]["stages"]["fft_inv"]["ms_per_launch"], ckey="k_inv_mfma")
    for name, mode, terms, tmpl in (("hankel_f16", "fp16", 1, 0), ("hankel_bf16x3", "bf16x3", 3, 2)):
        if name not in results or hankel_flops is None:
            continue
        sh = results[name]["stages"]
        t = sh["multiply"]["ms_per_launch"]
        issued = hankel_flops["issued_per_term"] * terms
        put(f"k_hankel<{mode}> (useful flops)", hankel_flops["useful"], t, unit="TFLOP/s",
            peak=MFMA_F16_PEAK_TFLOPS, ckey=f"k_hankel<{tmpl}>", issued_tflops=round(issued / (t / 1e3) / 1e12, 1),
            issued_over_useful=round(issued / hankel_flops["useful"], 2),
            dense_over_useful=round(hankel_flops["dense"] / hankel_flops["useful"], 1))
        ybytes = 4 if mode == "fp16" else 8
        put(f"k_inv_dense<14,{mode}>", ndm * count * (N * ybytes + M * 8), sh["fft_inv"]["ms_per_launch"],
            ckey=f"k_inv_dense_nb<14, {'true' if mode == 'fp16' else 'false'}>")
    if fold:
        put("k_fold_phase", fold["alg_bytes_per_launch"], fold["ms_per_launch"], ckey="k_fold_phase<")
    return out


GATES = {"plain_f32": 1e-5, "mixed_f16": 1e-3, "mixed_f16_mfma": 1e-3, "hankel_f16": 1e-3, "hankel_bf16x3": 1e-3}


def check_blocks(eng, plan, d_in, n, nb=2):
    """Re-run from a reset engine to the first call whose blocks lie past the
    longest delay line (K_max blocks) and keep that call's first nb blocks
    of every trial (the same blocks for every form)."""
    eng.reset()
    c_chk = (plan.K_max + plan.count - 1) // plan.count
    for c in range(c_chk + 1):
        eng.filter_u16(d_in[c * n:(c + 1) * n])
    out = eng.output_c64()
    return dict(blocks=[c_chk * plan.count, c_chk * plan.count + nb],
                rows=out[:, :nb].reshape(plan.num_dms, -1).astype(np.complex128))


def host_reader(u):
    """Sample-range reader over packed uint16 pairs in host memory."""
    return lambda a, b: u[2 * a:2 * b]


def device_reader(d_in):
    """Sample-range reader over packed uint16 pairs in a device tensor."""
    return lambda a, b: d_in[a:b].cpu().numpy().view(np.uint16)


def unpacked(read, lo, hi):
    """Complex128 samples [lo, hi) of the stream (zero before sample 0; code
    0 of a component is "no data", 0.0)."""
    x = np.zeros(hi - lo, np.complex128)
    s = max(0, lo)
    w = read(s, hi)
    v = w.astype(np.float64) - 32768.0
    v[w == 0] = 0.0
    x[s - lo:] = v[0::2] + 1j * v[1::2]
    return x


def fp64_same_plan(eng, plan, read, blocks):
    """FP64 execution of the same plan (numpy, complex128) for output blocks
    [b0, b1) of every trial: the engine's own per-bin delay table and filter
    table H, the forward spectra of the overlapped windows of the packed
    input, the delayed gather, the inverse transform and the discard, at the
    engine's scale (the last M points of IDFT(X_{g-d} H / N) / N)."""
    b0, b1 = blocks
    M, N, ndm = plan.M, 2 * plan.M, plan.num_dms
    d = np.stack([plan.delay_points(i) for i in range(ndm)]).astype(np.int64)
    H = np.stack([eng.dedisp_params(i) for i in range(ndm)]).astype(np.complex128)
    g0 = max(0, b0 - int(d.max()))
    # window of block g: samples [(g-1)M, (g+1)M); the last block b1 - 1 ends at b1 M
    x = unpacked(read, (g0 - 1) * M, b1 * M)
    win = np.lib.stride_tricks.sliding_window_view(x, N)[::M][:b1 - g0]
    X = np.fft.fft(win, axis=1)
    n = np.arange(N)
    out = np.empty((ndm, (b1 - b0) * M), np.complex128)
    for t in range(ndm):
        for k, g in enumerate(range(b0, b1)):
            src = g - d[t] - g0                     # negative: a block before the stream start
            Z = np.where(src >= 0, X[np.maximum(src, 0), n], 0) * H[t] / N
            out[t, k * M:(k + 1) * M] = np.fft.ifft(Z)[M:]
    return out


def osm_fp64(ms, bw, dm, f0, read, s0, n):
    """FP64 single-segment overlap-save output samples [s0, s0+n) of one
    trial (transform of 2 Mo points, Mo = 2 nextpow2(Nd); the filter table of
    an OSM engine with the chirp on an exact FP64 grid), numpy complex128."""
    plan = ms.Plan(bw, dm, f0, fftpoint=0, count=1, method=ms.METHOD_OSM, chirp=ms.CHIRP_EXACT_F64,
                   form=ms.FORM_PLAIN_F32)
    eng = ms.Engine(plan)
    H = eng.dedisp_params(0).astype(np.complex128)
    eng.close()
    Mo = plan.M
    out = np.empty(n, np.complex128)
    g = s0 // Mo
    while g * Mo < s0 + n:
        X = np.fft.fft(unpacked(read, (g - 1) * Mo, (g + 1) * Mo))
        y = np.fft.ifft(X * H / (2 * Mo))[Mo:]
        a, b = max(s0, g * Mo), min(s0 + n, (g + 1) * Mo)
        out[a - s0:b - s0] = y[a - g * Mo:b - g * Mo]
        g += 1
    return out, Mo


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def osm_informational(ms, bw, f0, dms, read, rows, ref, M, trials=(14.5, 57.3, 100.0)):
    """Distance of the headline output and of the FP64 same-plan MS-OSM output
    from the FP64 single-segment OSM output (exact chirp grid), for the
    trials nearest the given DMs, on the check blocks.  This is the MS-OSM
    algorithm's approximation (segmented chirp) plus the reference's float
    chirp grid, not an arithmetic error."""
    b0, b1 = rows["blocks"]
    out = {}
    for dm in trials:
        i = int(np.argmin(np.abs(dms - dm)))
        o, Mo = osm_fp64(ms, bw, float(dms[i]), f0, read, b0 * M, (b1 - b0) * M)
        out[f"dm{float(dms[i]):.2f}"] = dict(osm_fftpoint=2 * Mo, mixed_f16=rel(rows["rows"][i], o),
                                             msosm_f64=rel(ref[i], o))
    return out


def compare_fp64(blocks, ref, dms, gate):
    """Per-trial relative L2 of a form's check blocks against the FP64
    execution of the same plan, over every trial."""
    got = blocks["rows"]
    err = np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)
    w = int(np.argmax(err))
    return dict(max=float(err.max()), median=float(np.median(err)), worst_dm=round(float(dms[w]), 3),
                trials=int(len(err)), blocks=blocks["blocks"], gate=gate, passed=bool(err.max() < gate))


def compare_blocks(blocks, ref, dms):
    """Relative L2 of a form's check blocks against the plain FP32 form's
    (the comparison the metric names), over every trial."""
    err = np.linalg.norm(blocks["rows"] - ref["rows"], axis=1) / np.linalg.norm(ref["rows"], axis=1)
    return dict(blocks=ref["blocks"], max=float(err.max()), median=float(np.median(err)))


def cpu_baseline(bw, f0, dms, M, u, log2_samples, reps=5, log2_single=20):
    """The oracle's FP32 MS-OSM restatement (OpenMP over DM x block) on the
    host cores: median of `reps` runs over the first 2^log2_samples samples
    for all DM trials with every core the process may use, after a warm-up;
    and one single-thread run over the first 2^log2_single samples."""
    import oracle
    ncores = len(os.sched_getaffinity(0))
    Lc = 1 << log2_samples
    delays = np.stack([oracle.delay_points(bw, float(dm), f0, M)[0] for dm in dms])
    H = np.stack([oracle.dedisp_params(bw, float(dm), f0, M) for dm in dms]).astype(np.complex64)
    us = np.ascontiguousarray(u[:2 * Lc])
    oracle.cpu_msosm_f32(us[:2 * 4 * M], M, delays[:2], H[:2], nthreads=ncores)  # warm-up
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        oracle.cpu_msosm_f32(us, M, delays, H, nthreads=ncores)
        times.append(time.perf_counter() - t)
    dt = float(np.median(times))
    L1 = 1 << log2_single
    t = time.perf_counter()
    oracle.cpu_msosm_f32(np.ascontiguousarray(u[:2 * L1]), M, delays, H, nthreads=1)
    dt1 = time.perf_counter() - t
    return dict(value=round(len(dms) * Lc / dt / 1e6, 3), unit=UNIT, cores=ncores, kind="port",
                sample="first 2^%d samples x %d DM trials, FP32 MS-OSM (oracle restatement), median of %d runs "
                       "(%.1f s each)" % (log2_samples, len(dms), reps, dt),
                single_thread=dict(value=round(len(dms) * L1 / dt1 / 1e6, 3), cores=1,
                                   sample="first 2^%d samples x %d DM trials, %.1f s" % (log2_single, len(dms), dt1)),
                single_trial_16mhz=cpu_single_trial(ncores))


def cpu_single_trial(ncores, reps=3):
    """The single-trial CPU workload whole: 2^22 samples of a simulated
    16 MHz pulsar (f0 1392 MHz, DM 10, P 5 ms, 5% duty, per-sample S/N 0.1),
    FP32 MS-OSM at FFT 1024 (K 16) with the oracle restatement, then the
    oracle's sequential fold (the reference's float state machine) of the
    output intensity into 256 bins; best of `reps` with every core for the
    filter (the fold is sequential)."""
    import oracle
    from msosm_gpu_amd import simulate
    bw, f0, dm, period, M, L = 16e6, 1392e6, 10.0, 0.005, 512, 1 << 22
    u = simulate.pulsar_u16(L, bw, f0, dm, period, duty=0.05, snr=0.1, seed=0x5EED0001)
    d = oracle.delay_points(bw, dm, f0, M)[0][None]
    H = oracle.dedisp_params(bw, dm, f0, M)[None].astype(np.complex64)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        _, y = oracle.cpu_msosm_f32(u, M, d, H, want_output=True, nthreads=ncores)
        t1 = time.perf_counter()
        f = oracle.Fold(period, bw, L, 256)
        f.discard_samples(oracle.nd(bw, dm, f0))
        f.fold_data((y[0].real.astype(np.float32) ** 2 + y[0].imag.astype(np.float32) ** 2))
        f.bins()
        t2 = time.perf_counter()
        if best is None or t2 - t0 < best[0] + best[1]:
            best = (t1 - t0, t2 - t1)
    return dict(value=round(L / sum(best) / 1e6, 3), unit="Msamples/s", cores=ncores, kind="port",
                filter_s=round(best[0], 3), fold_s=round(best[1], 3),
                sample="2^22 samples, 16 MHz, DM 10, FFT 1024 (K 16), FP32 MS-OSM + fold into 256 bins, best of %d"
                       % reps)


if __name__ == "__main__":
    main()
