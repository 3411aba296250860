"""The shared-ring forms (fused FP32 and FP16 storage) at FFT sizes 2^15 ..
2^21, run as four-step transforms, against the FP64 same-plan oracle and
against the reference-structured form of the same plan; plus the FFT plan
API (msosm_fft_plan_*) against the per-call entry point."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle
from test_gpu_engine import make_input, oracle_out, rel_l2, run_engine

pytestmark = pytest.mark.gpu

# (log2 FFT size, bandwidth, DMs, f0, blocks per call, calls)
CASES = [
    (15, 64e6, [30.0], 1368e6, 3, 4),
    (16, 64e6, [0.0, 20.0, 57.3], 1368e6, 4, 3),
    (17, 32e6, [150.0], 1384e6, 2, 4),
    (18, 64e6, [200.0], 1368e6, 2, 3),
    (21, 128e6, [100.0], 1336e6, 1, 6),
]


@pytest.mark.parametrize("case", CASES, ids=[f"2^{c[0]}" for c in CASES])
@pytest.mark.parametrize("form", [ms.FORM_FUSED_F32, ms.FORM_FUSED_F16])
def test_fused_large_sizes_vs_fp64(device, case, form):
    log2n, bw, dms, f0, count, calls = case
    M = 1 << (log2n - 1)
    p = ms.Plan(bw, dms, f0, fftpoint=2 * M, count=count, form=form)
    assert p.M == M and not p.clamped
    u = make_input(calls * count * M, 900 + log2n)
    print(f"fused 2^{log2n} form={form} K={p.K_max} trials={len(dms)}", flush=True)
    y = run_engine(p, u, device=device)
    tol = 1e-5 if form == ms.FORM_FUSED_F32 else 1e-3
    for i, dm in enumerate(dms):
        err = rel_l2(y[i], oracle_out(bw, dm, f0, M, u))
        print(f"  DM {dm}: rel L2 vs FP64 {err:.2e}", flush=True)
        assert err < tol, (dm, err)


def test_config2_plan_full_delay_line(device):
    """128 MHz, DM 100 at its stated plan (FFT 2^16, K 152) in the FP16 form,
    over 160 blocks so the delay line is full, against the FP64 oracle on the
    last blocks and against the reference-structured FP32 form on all."""
    import torch
    bw, dm, f0 = 128e6, 100.0, 1336e6
    M, count, calls = 32768, 16, 10
    u = make_input(calls * count * M, 1234)
    outs = {}
    for form in (ms.FORM_FUSED_F16, ms.FORM_PLAIN_F32):
        p = ms.Plan(bw, dm, f0, fftpoint=2 * M, count=count, form=form)
        assert p.K_max == 152
        outs[form] = run_engine(p, u, device=device)[0]
    a, b = outs[ms.FORM_FUSED_F16], outs[ms.FORM_PLAIN_F32]
    assert rel_l2(a, b) < 1e-3
    L = calls * count * M
    b0 = L // M - 2
    s0 = b0 - 153
    x = oracle.unpack_u16(u[2 * s0 * M:]).astype(np.complex128)
    d, K, _ = oracle.delay_points(bw, dm, f0, M)
    ref = oracle.msosm_run(x, M, d, oracle.dedisp_params(bw, dm, f0, M), b0=b0 - s0, nb=2)
    err = rel_l2(a[b0 * M:], ref)
    print(f"FFT 2^16, K 152: FP16 four-step vs FP64 on the last two blocks {err:.2e}; vs plain FP32 "
          f"{rel_l2(a, b):.2e}")
    assert err < 1e-3
    torch.cuda.synchronize()


def test_large_size_state_and_host_input(device):
    # host-side input staging and export/import at 2^17 (ring in position order)
    bw, dms, f0 = 32e6, [40.0, 150.0], 1384e6
    p = ms.Plan(bw, dms, f0, fftpoint=1 << 17, count=2, form=ms.FORM_FUSED_F16)
    u = make_input(5 * p.count * p.M, 77)
    a = run_engine(p, u, device_input=True, device=device)
    b = run_engine(p, u, device_input=False, device=device)
    np.testing.assert_array_equal(a, b)
    import torch
    dev = torch.from_numpy(u.view(np.int32).copy()).to(device)
    n = p.count * p.M
    e1, e2 = ms.Engine(p), ms.Engine(p)
    for c in range(2):
        e1.filter_u16(dev[c * n:(c + 1) * n])
    e2.state_import(e1.state_export())
    for c in range(2, 5):
        e1.filter_u16(dev[c * n:(c + 1) * n])
        e2.filter_u16(dev[c * n:(c + 1) * n])
        np.testing.assert_array_equal(e1.output_c64(), e2.output_c64())
    e1.close()
    e2.close()


@pytest.mark.parametrize("log2n,batch", [(12, 3), (16, 4), (20, 2)])
def test_fft_plan_matches_entry_point(device, log2n, batch):
    import torch
    n = 1 << log2n
    rng = np.random.default_rng(log2n)
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    xd = torch.from_numpy(x).to(device)
    a = torch.empty_like(xd)
    b = torch.empty_like(xd)
    plan = ms.FftPlan(n, batch)
    for direction in (-1, +1):
        plan(xd, a, direction=direction)
        ms.fft_c2c(xd, b, n, batch, direction=direction)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    plan.close()


@pytest.mark.parametrize("kind", ["full_scale_noise", "pulse_train_32", "constant"])
def test_fp16_row_scratch_full_scale(device, kind):
    """The FP16-storage form above 2^14 keeps the inverse's row scratch in
    FP16 scaled by 2^-9: inputs at the uint16 limits (maximum power spread
    over all bins, a pulse train whose spectrum adds coherently within rows,
    a constant whose whole power sits in one bin) stay finite and within the
    mixed-precision gate against FP64, at FFT 2^16 (4 rows of 2^14)."""
    M, count, calls = 1 << 15, 3, 3
    L = calls * count * M
    rng = np.random.default_rng(12)
    if kind == "full_scale_noise":
        u = rng.choice(np.array([1, 65535], np.uint16), 2 * L)
    elif kind == "pulse_train_32":
        u = np.full(2 * L, 32768, np.uint16)
        u[0::64] = 65535
        u[1::64] = 65535
    else:
        u = np.full(2 * L, 65535, np.uint16)
    bw, dm, f0 = 64e6, 20.0, 1368e6
    p = ms.Plan(bw, dm, f0, fftpoint=2 * M, count=count, form=ms.FORM_FUSED_F16)
    y = run_engine(p, u, device=device)[0]
    assert np.all(np.isfinite(y))
    err = rel_l2(y, oracle_out(bw, dm, f0, M, u))
    print(f"fp16 row scratch {kind}: rel L2 vs FP64 {err:.2e}", flush=True)
    assert err < 1e-3


@pytest.mark.parametrize("form,gate", [(ms.FORM_FUSED_F32, 1e-5), (ms.FORM_FUSED_F16, 1e-3)])
@pytest.mark.parametrize("log2n,count,calls,dm", [(16, 80, 3, 57.3), (15, 300, 2, 30.0)],
                         ids=["2^16x80", "2^15x300"])
def test_launches_with_more_rows_than_cus(device, form, gate, log2n, count, calls, dm):
    """Calls whose forward row launch has more 2^14-point rows than the GPU
    has CUs (2^16 x 80 blocks: 320 rows, runs of 2; 2^15 x 300 blocks: 600
    rows, runs of 3): the row pass then runs rows in runs per workgroup.  The
    last call's last two blocks against the FP64 same-plan oracle, past the
    delay line."""
    bw, f0 = 64e6, 1368e6
    M = 1 << (log2n - 1)
    p = ms.Plan(bw, dm, f0, fftpoint=2 * M, count=count, form=form)
    assert p.M == M and not p.clamped
    K = p.K_max
    u = make_input(calls * count * M, 4321 + log2n)
    print(f"rows > CUs: 2^{log2n} count={count} K={K} form={form}", flush=True)
    y = run_engine(p, u, device=device)[0]
    b0 = calls * count - 2
    s0 = max(0, b0 - K - 1)
    x = oracle.unpack_u16(u[2 * s0 * M:]).astype(np.complex128)
    d, _, _ = oracle.delay_points(bw, dm, f0, M)
    ref = oracle.msosm_run(x, M, d, oracle.dedisp_params(bw, dm, f0, M), b0=b0 - s0, nb=2)
    err = rel_l2(y[b0 * M:], ref)
    print(f"  rel L2 vs FP64 {err:.2e} (gate {gate:g})", flush=True)
    assert err < gate
