"""GPU parity of the dedispersion engine (all execution forms) against the
FP64 same-plan oracle.  Gates: FP32 forms <= 1e-5 relative L2, the
FP16-storage and matrix-core forms <= 1e-3 relative L2 (per DM, whole
stream including the zero-filled transient)."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
MIXED_TOL = 1e-3
FORMS = [ms.FORM_PLAIN_F32, ms.FORM_FUSED_F32, ms.FORM_FUSED_F16, ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16X3]
TOL = {ms.FORM_PLAIN_F32: FP32_TOL, ms.FORM_FUSED_F32: FP32_TOL, ms.FORM_FUSED_F16: MIXED_TOL,
       ms.FORM_HANKEL_F16: MIXED_TOL, ms.FORM_HANKEL_BF16X3: MIXED_TOL}


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def make_input(L, seed, sigma=900.0, zeros=32):
    rng = np.random.Generator(np.random.Philox(seed))
    u = np.clip(np.round(32768 + sigma * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    if zeros:
        u[rng.integers(0, 2 * L, zeros)] = 0
    return u


def run_engine(plan, u, device_input=True, device=None):
    """Feed u (2L uint16) call by call; returns output [ndm, L] complex64."""
    import torch
    eng = ms.Engine(plan)
    n = plan.count * plan.M
    L = u.size // 2
    assert L % n == 0
    outs = []
    words = u.view(np.uint32)
    dev = torch.from_numpy(words.view(np.int32).copy()).to(device) if device_input else None
    for c in range(L // n):
        if device_input:
            eng.filter_u16(dev[c * n:(c + 1) * n])
        else:
            eng.filter_u16(words[c * n:(c + 1) * n])
        outs.append(eng.output_c64())
    eng.close()
    y = np.concatenate(outs, axis=1)  # [ndm, calls*count, M]
    return y.reshape(plan.num_dms, -1)


def oracle_out(bw, dm, f0, M, u, mode=oracle.CHIRP_REF_F32W, method=ms.METHOD_MSOSM):
    x = oracle.unpack_u16(u).astype(np.complex128)
    if method == ms.METHOD_MSOSM:
        d, K, _ = oracle.delay_points(bw, dm, f0, M)
    else:
        d = np.zeros(2 * M, np.int32)
    H = oracle.dedisp_params(bw, dm, f0, M, mode)
    return oracle.msosm_run(x, M, d, H)


@pytest.mark.parametrize("form", FORMS)
def test_golden_pair_config1_plan(device, golden, form):
    arrays, meta = golden
    c = meta["msosm"]
    u = arrays["msosm_in_u16"]
    p = ms.Plan(c["bw"], c["dm"], c["f0"], fftpoint=2 * c["M"], count=8, form=form)
    print(f"golden pair form={form}", flush=True)
    y = run_engine(p, u, device=device)[0]
    assert rel_l2(y, arrays["msosm_out_ref"]) < TOL[form]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("mode", [ms.CHIRP_REF_F32W, ms.CHIRP_EXACT_F64])
def test_filter_tables_match_oracle(device, form, mode):
    bw, dm, f0 = 16e6, 10.0, 1392e6
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=2, form=form, chirp=mode)
    eng = ms.Engine(p)
    H = eng.dedisp_params(0)
    eng.close()
    assert rel_l2(H, oracle.dedisp_params(bw, dm, f0, 512, mode)) < 1e-6


@pytest.mark.parametrize("form", FORMS)
def test_config1_single_dm_vs_fp64(device, form):
    # 16 MHz, DM 10, f0 1392 MHz, N = 1024 (M = 512, K = 16), 2^18 samples
    bw, dm, f0 = 16e6, 10.0, 1392e6
    u = make_input(1 << 18, 0x5EED0001)
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=32, form=form)
    assert (p.M, p.K_max, p.R) == (512, 16, 64)
    print(f"config1 form={form}", flush=True)
    y = run_engine(p, u, device=device)[0]
    ref = oracle_out(bw, dm, f0, 512, u)
    err = rel_l2(y, ref)
    print(f"  rel L2 = {err:.3e}")
    assert err < TOL[form]


@pytest.mark.parametrize("form", FORMS)
def test_exact_chirp_mode(device, form):
    bw, dm, f0 = 32e6, 60.0, 1384e6
    u = make_input(1 << 17, 7)
    p = ms.Plan(bw, dm, f0, fftpoint=4096, count=4, form=form, chirp=ms.CHIRP_EXACT_F64)
    y = run_engine(p, u, device=device)[0]
    ref = oracle_out(bw, dm, f0, 2048, u, mode=oracle.CHIRP_EXACT_F64)
    assert rel_l2(y, ref) < TOL[form]


@pytest.mark.parametrize("form", FORMS)
def test_multi_dm_vs_fp64(device, form):
    # 8 DM trials on one 16384-point plan, K from 1 (DM 0) to 44
    bw, f0 = 64e6, 1368e6
    dms = [0.0, 0.4, 3.0, 7.5, 12.0, 18.25, 24.0, 29.0]
    p = ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=form)
    u = make_input(16 * 8192, 99)
    print(f"multi-dm form={form} K_max={p.K_max}", flush=True)
    y = run_engine(p, u, device=device)
    for i, dm in enumerate(dms):
        ref = oracle_out(bw, dm, f0, 8192, u)
        err = rel_l2(y[i], ref)
        assert err < TOL[form], (i, dm, err)


@pytest.mark.parametrize("count,calls", [(1, 8), (3, 5)])
@pytest.mark.parametrize("form", FORMS + [ms.FORM_FUSED_F16_MFMA])
def test_short_delay_line_few_blocks(device, form, count, calls):
    # K from 1 to 3 at FFT 16384 with 1 or 3 blocks per call: a ring of only
    # K_max - 1 + count slots that every call wraps, and tiles of 16 blocks
    # (Hankel) or of 16 / 32 blocks per workgroup that a call fills in part
    bw, f0 = 64e6, 1368e6
    dms = [0.0, 1.0, 1.5]
    p = ms.Plan(bw, dms, f0, fftpoint=16384, count=count, form=form)
    assert p.K_max == 3
    u = make_input(calls * count * 8192, 31 + count)
    print(f"short delay line form={form} count={count} calls={calls}", flush=True)
    y = run_engine(p, u, device=device)
    for i, dm in enumerate(dms):
        err = rel_l2(y[i], oracle_out(bw, dm, f0, 8192, u))
        assert err < TOL.get(form, MIXED_TOL), (dm, err)


def test_block_split_invariance(device):
    # 1 call of 16 blocks == 16 calls of 1 block == 4 calls of 4 (bit for bit)
    bw, dm, f0 = 16e6, 10.0, 1392e6
    u = make_input(64 * 512, 3)
    for form in FORMS:
        outs = []
        for count in (16, 1, 4):
            p = ms.Plan(bw, dm, f0, fftpoint=1024, count=count, form=form)
            outs.append(run_engine(p, u, device=device))
        np.testing.assert_array_equal(outs[0], outs[1])
        np.testing.assert_array_equal(outs[0], outs[2])


@pytest.mark.parametrize("form", [ms.FORM_FUSED_F32, ms.FORM_FUSED_F16])
def test_blocks_per_workgroup_choice(device, form):
    """With 512 trials the fused inverse runs 32 blocks per workgroup (here
    20 blocks per call: one partial chunk); 8 trials alone run 1 per
    workgroup (too few workgroups otherwise).  Each trial's output must not
    depend on that choice (bit for bit), and must match the FP64 oracle."""
    bw, f0 = 64e6, 1368e6
    dms = np.linspace(0.0, 50.0, 512).astype(np.float32)
    u = make_input(2 * 20 * 2048, 41)
    big = ms.Plan(bw, dms, f0, fftpoint=4096, count=20, form=form)
    small = ms.Plan(bw, dms[:8], f0, fftpoint=4096, count=20, form=form)
    assert big.M == 2048 and big.K_max > 100
    a = run_engine(big, u, device=device)
    b = run_engine(small, u, device=device)
    np.testing.assert_array_equal(a[:8], b)
    for i in (0, 255, 511):
        err = rel_l2(a[i], oracle_out(bw, float(dms[i]), f0, 2048, u))
        print(f"form {form} DM {dms[i]:.2f}: rel L2 {err:.2e}", flush=True)
        assert err < TOL[form]


def test_host_and_device_input_agree(device):
    bw, dm, f0 = 16e6, 10.0, 1392e6
    u = make_input(32 * 512, 4)
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=8, form=ms.FORM_FUSED_F16)
    a = run_engine(p, u, device_input=True, device=device)
    b = run_engine(p, u, device_input=False, device=device)
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("form", [ms.FORM_PLAIN_F32, ms.FORM_FUSED_F16])
def test_host_input_refilled_buffer_pipelined(device, form):
    # One host buffer refilled before every call and no output read in
    # between: the copies alternate staging halves and overlap the previous
    # call's kernels.  The last output depends on the previous K+1 calls'
    # input (K = 16 with count 1), so a copy that overwrote a half still being
    # read, or a call that read stale input, changes it.
    bw, dm, f0 = 16e6, 10.0, 1392e6
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=1, form=form)
    assert p.K_max == 16
    ncalls = 40
    u = make_input(ncalls * 512, 31).view(np.uint32)
    ref = run_engine(p, u.view(np.uint16), device=device)[0].reshape(ncalls, 512)
    eng = ms.Engine(p)
    buf = np.empty(512, np.uint32)
    for c in range(ncalls):
        buf[:] = u[c * 512:(c + 1) * 512]
        eng.filter_u16(buf)
        if c == 25:
            np.testing.assert_array_equal(eng.output_c64()[0, 0], ref[c])
    last = eng.output_c64()[0, 0]
    eng.close()
    np.testing.assert_array_equal(last, ref[-1])


def test_forms_agree(device):
    bw, f0 = 8e6, 1396e6
    dms = [50.0, 300.0]
    u = make_input(1 << 16, 12)
    outs = {f: run_engine(ms.Plan(bw, dms, f0, fftpoint=1024, count=8, form=f), u, device=device) for f in FORMS}
    for i in range(2):
        assert rel_l2(outs[ms.FORM_FUSED_F32][i], outs[ms.FORM_PLAIN_F32][i]) < 2e-6
        assert rel_l2(outs[ms.FORM_FUSED_F16][i], outs[ms.FORM_PLAIN_F32][i]) < MIXED_TOL


@pytest.mark.parametrize("form", [ms.FORM_PLAIN_F32, ms.FORM_FUSED_F32, ms.FORM_FUSED_F16])
def test_osm_two_pass_fft(device, form):
    # OSM plan of the 16 MHz / DM 10 case: M = 16384, FFT 32768 (two-pass for
    # the plain form, four-step with a one-slot delay line for the fused ones)
    bw, dm, f0 = 16e6, 10.0, 1392e6
    p = ms.Plan(bw, dm, f0, fftpoint=0, count=2, method=ms.METHOD_OSM, form=form)
    assert p.M == 16384 and p.K_max == 1
    u = make_input(4 * 16384, 21)
    y = run_engine(p, u, device=device)[0]
    ref = oracle_out(bw, dm, f0, 16384, u, method=ms.METHOD_OSM)
    err = rel_l2(y, ref)
    print(f"OSM FFT 32768 form {form}: rel L2 {err:.2e}", flush=True)
    assert err < TOL[form]


@pytest.mark.parametrize("form", [ms.FORM_FUSED_F32, ms.FORM_FUSED_F16, ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16X3])
def test_osm_method_shared_ring_forms(device, form):
    # single-segment overlap-save (K = 1, no delay line) for three trials with
    # a common M = 4096 (FFT 8192), through the shared-ring forms; 4 blocks
    # per call against a ring of 4 slots also covers the Hankel staging's slot
    # modulo (a 16-block tile reaches back several times around the ring)
    bw, f0 = 16e6, 1392e6
    dms = [0.5, 1.0, 2.0]
    p = ms.Plan(bw, dms, f0, fftpoint=0, count=4, method=ms.METHOD_OSM, form=form)
    assert p.M == 4096 and p.K_max == 1
    u = make_input(3 * 4 * 4096, 22)
    y = run_engine(p, u, device=device)
    for i, dm in enumerate(dms):
        # each trial's own OSM filter at the common M
        x = oracle.unpack_u16(u).astype(np.complex128)
        H = oracle.dedisp_params(bw, dm, f0, 4096)
        ref = oracle.msosm_run(x, 4096, np.zeros(8192, np.int32), H)
        err = rel_l2(y[i], ref)
        print(f"OSM form {form} DM {dm}: rel L2 {err:.2e}", flush=True)
        assert err < TOL[form]


def test_config2_fft_size_plain(device):
    # N = 2^16 (the 128 MHz production plan's FFT size) at DM 5: two-pass path
    bw, dm, f0 = 128e6, 5.0, 1336e6
    p = ms.Plan(bw, dm, f0, fftpoint=65536, count=2)
    u = make_input(8 * 32768, 22)
    y = run_engine(p, u, device=device)[0]
    ref = oracle_out(bw, dm, f0, 32768, u)
    assert rel_l2(y, ref) < FP32_TOL


def test_output_u16_matches_pack(device):
    bw, dm, f0 = 16e6, 10.0, 1392e6
    u = make_input(8 * 512, 5)
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=8)
    eng = ms.Engine(p)
    eng.filter_u16(u.view(np.uint32))
    c = eng.output_c64()
    q = eng.output_u16()
    eng.close()
    np.testing.assert_array_equal(q.reshape(-1), oracle.pack_u16(c.reshape(-1)))


def test_reset_restarts_stream(device):
    bw, dm, f0 = 16e6, 10.0, 1392e6
    u = make_input(16 * 512, 6)
    for form in FORMS:
        p = ms.Plan(bw, dm, f0, fftpoint=1024, count=16, form=form)
        eng = ms.Engine(p)
        eng.filter_u16(u.view(np.uint32))
        a = eng.output_c64()
        eng.filter_u16(u.view(np.uint32))
        eng.reset()
        eng.filter_u16(u.view(np.uint32))
        b = eng.output_c64()
        eng.close()
        np.testing.assert_array_equal(a, b)


def test_timing_reports_stages(device):
    p = ms.Plan(16e6, 10.0, 1392e6, fftpoint=1024, count=8, form=ms.FORM_PLAIN_F32)
    eng = ms.Engine(p)
    eng.set_timing(True)
    u = make_input(8 * 512, 8)
    for _ in range(3):
        eng.filter_u16(u.view(np.uint32))
    t = eng.get_timing()
    eng.close()
    assert t["fft_fwd"][1] == 3 and t["fft_inv"][1] == 3 and t["multiply"][1] == 3
    assert t["h2d"][1] == 3  # host input: one copy per call, timed on the copy stream
    assert all(v[0] >= 0 for v in t.values())
