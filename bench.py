#!/usr/bin/env python
"""Throughput of the MS-OSM coherent-dedispersion engine on one MI355X.

Workload (multi-DM batched precision configuration): 2^26 complex samples of
synthetic 64 MHz baseband (f0 = 1368 MHz, a pulsar at DM 57.3 with P = 8 ms),
256 DM trials evenly spaced on [0, 100] pc cm^-3, FFT size 2^14 (M = 8192,
K = 1..152 segments), count = 64 blocks per call.  One step = one pass of
the whole input through all 256 trials (128 calls), input resident in HBM,
output written per call to the engine's complex64 output buffer.

The headline value is the FP16-storage fused form; the plain FP32 form (the
reference's product-ring structure) is timed beside it for the ratio.  Both
are checked against the FP64 same-plan oracle on a block range.

Metric: DM-trial * Msamples / s = num_dms * input_samples / time / 1e6.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DM-trial·Msamples/s on 1 MI355X, mixed vs FP32 plain at ≤1e-3 rel err"
UNIT = "DM-trial·Msamples/s"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2-samples", type=int, default=26)
    ap.add_argument("--num-dms", type=int, default=256)
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--fftpoint", type=int, default=16384)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP32 comparison")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 spot check")
    ap.add_argument("--cpu-log2-samples", type=int, default=25)
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine shards over DM trials with no exchange beyond an input "
              "broadcast, and this build runs on single-GPU boxes only; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import msosm_gpu_amd as ms
    from msosm_gpu_amd import simulate

    dev = torch.device("cuda:0")
    ms.lib()
    L = 1 << args.log2_samples
    bw, f0 = 64e6, 1368e6
    dms = np.linspace(0.0, 100.0, args.num_dms).astype(np.float32)
    t0 = time.time()
    u = simulate.pulsar_u16(L, bw, f0, 57.3, 0.008, duty=0.05, snr=0.1, seed=0x5EED0003)
    log(f"input: {L} samples generated in {time.time() - t0:.1f}s")
    d_in = torch.from_numpy(u.view(np.int32)).to(dev)
    stream = torch.cuda.current_stream()
    results = {}
    checks = {}
    forms = [("mixed_f16", ms.FORM_FUSED_F16)]
    if not args.no_plain:
        forms.insert(0, ("plain_f32", ms.FORM_PLAIN_F32))
    plan_info = None
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
        if not args.no_check:
            checks[name] = spot_check(ms, eng, plan, d_in, u, bw, f0, dms, n)
            log(f"{name}: max rel L2 vs FP64 on spot check = {checks[name]['max_rel_l2']:.3e}")
        eng.close()
        del eng
        torch.cuda.synchronize()

    mixed = results["mixed_f16"]
    st = mixed["stages"]["fft_inv"]
    units = len(dms) * plan_info["count"]
    alg_bytes = units * plan_info["M"] * (8 + 8)  # gather 2 FP16 bins + write 1 complex64 per DM-sample
    achieved = alg_bytes / (st["ms_per_launch"] / 1e3) / 1e9
    roofline = dict(bound="hbm", kernel="k_inv_fused<14,fp16>", achieved=round(achieved, 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    alg_bytes_per_launch=alg_bytes, ms_per_launch=st["ms_per_launch"])
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(bw, f0, dms, args.fftpoint // 2, u, args.cpu_log2_samples)
        log(f"cpu baseline: {cpu}")
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
        "dtype": "f32",
        "storage": "f16 spectra ring and filter table, complex64 output",
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
        "rel_l2_vs_fp64": {k: v for k, v in checks.items()},
    }
    print(json.dumps(line), flush=True)


def spot_check(ms, eng, plan, d_in, u, bw, f0, dms, n):
    """Re-run from a reset engine up to call c and compare its first blocks
    for a few DMs against the FP64 same-plan oracle."""
    import oracle
    eng.reset()
    c_chk = 2
    for c in range(c_chk + 1):
        eng.filter_u16(d_in[c * n:(c + 1) * n])
    out = eng.output_c64()
    M = plan.M
    nb = 4
    b0 = c_chk * plan.count
    x = oracle.unpack_u16(u[:2 * (b0 + nb) * M]).astype(np.complex128)
    errs = {}
    for i in (0, 37, 146, len(dms) - 1):
        d, K, _ = oracle.delay_points(bw, float(dms[i]), f0, M)
        H = oracle.dedisp_params(bw, float(dms[i]), f0, M)
        ref = oracle.msosm_run(x, M, d, H, b0=b0, nb=nb)
        got = out[i, :nb].reshape(-1)
        errs[f"dm{float(dms[i]):.2f}"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    return dict(blocks=[b0, b0 + nb], per_dm=errs, max_rel_l2=max(errs.values()))


def cpu_baseline(bw, f0, dms, M, u, log2_samples):
    """The oracle's FP32 MS-OSM restatement (OpenMP over DM x block) on the
    host cores, on the first 2^log2_samples samples for all DM trials."""
    import oracle
    ncores = len(os.sched_getaffinity(0))
    Lc = 1 << log2_samples
    delays = np.stack([oracle.delay_points(bw, float(dm), f0, M)[0] for dm in dms])
    H = np.stack([oracle.dedisp_params(bw, float(dm), f0, M) for dm in dms]).astype(np.complex64)
    us = np.ascontiguousarray(u[:2 * Lc])
    oracle.cpu_msosm_f32(us[:2 * 4 * M], M, delays[:2], H[:2], nthreads=ncores)  # warm-up
    t = time.perf_counter()
    oracle.cpu_msosm_f32(us, M, delays, H, nthreads=ncores)
    dt = time.perf_counter() - t
    return dict(value=round(len(dms) * Lc / dt / 1e6, 3), unit=UNIT, cores=ncores, kind="port",
                sample="first 2^%d samples x %d DM trials, FP32 MS-OSM (oracle restatement), %.1f s"
                       % (log2_samples, len(dms), dt))


if __name__ == "__main__":
    main()
