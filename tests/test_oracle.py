"""CPU tests that pin the oracle (oracle/msosm_oracle.c) before anything is
checked against it: golden fixtures from the independent numpy restatement
(tests/golden/gen_golden.py), closed forms and brute force."""
import math

import numpy as np
import pytest

import oracle

KDM = 4.15e15


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- planner --
def test_plan_scalars_match_golden(golden):
    _, meta = golden
    for r in meta["plans"]:
        nd = oracle.nd(r["bw"], r["dm"], r["f0"])
        assert nd == r["Nd"], r
        mmin = oracle.msosm_min_order(nd) if r["method"] == "msosm" else oracle.osm_min_order(nd)
        assert mmin == r["M_min"], r
        M, clamped = oracle.clamp_order(r["fftpoint"], mmin)
        assert M == r["M"] and int(clamped) == r["clamped"], r
        if r["method"] == "msosm":
            assert max(oracle.delaycount(nd, M), 1) == r["K"], r


def test_nd_closed_form():
    # Nd ~ kdm*DM*(1/f0^2 - 1/(f0+bw)^2)*fs within float rounding of the inputs
    for bw, dm, f0 in [(16e6, 10, 1392e6), (128e6, 100, 1336e6), (64e6, 57.3, 1368e6), (8e6, 300, 1396e6)]:
        exact = KDM * dm * (1 / f0 ** 2 - 1 / (f0 + bw) ** 2) * bw
        assert abs(oracle.nd(bw, dm, f0) - exact) <= 1e-6 * exact + 1


def test_min_order_closed_form():
    for nd in [1, 2, 3, 4, 5, 15, 16, 17, 7743, 4976580, 1240246]:
        M = oracle.msosm_min_order(nd)
        assert M >= 2 * math.sqrt(nd) and M / 2 < 2 * math.sqrt(nd)
        assert M & (M - 1) == 0
    assert oracle.msosm_min_order(0) == 0
    assert oracle.osm_min_order(7743) == 16384 and oracle.osm_min_order(0) == 2


def test_delay_points_match_golden(golden):
    arrays, meta = golden
    for i, c in enumerate(meta["segw"]):
        d, K, holes = oracle.delay_points(c["bw"], c["dm"], c["f0"], c["M"])
        assert holes == 0
        assert K == c["K"]
        np.testing.assert_array_equal(d, arrays[f"delay_{i}"].astype(np.int32))
        w, neg = oracle.segment_w(c["bw"], c["dm"], c["f0"], c["M"])
        assert neg == c["negcount"]
        np.testing.assert_array_equal(w, arrays[f"segw_{i}"])


@pytest.mark.parametrize("bw,dm,f0,M", [(16e6, 10.0, 1392e6, 512), (64e6, 57.3, 1368e6, 8192),
                                        (8e6, 300.0, 1396e6, 512), (128e6, 100.0, 1336e6, 32768),
                                        (2e6, 3.3, 300e6, 128)])
def test_delay_invariants_and_brute_force(bw, dm, f0, M):
    d, K, holes = oracle.delay_points(bw, dm, f0, M)
    N = 2 * M
    assert holes == 0
    assert d.min() == 0 and d.max() == K - 1
    assert len(np.unique(d)) == K
    # frequency order: bins [M, 2M) are the lower half band, then [0, M)
    order = np.concatenate([np.arange(M, N), np.arange(0, M)])
    assert np.all(np.diff(d[order]) >= 0), "delay must not decrease with frequency"
    # brute force: d = floor(tau(f) * fs / M), tau relative to the band bottom
    k = np.arange(N)
    foff = np.where(k < M, k, k - N) * (bw / N) + bw / 2    # offset from f0
    tau = KDM * dm * (1 / f0 ** 2 - 1 / (f0 + foff) ** 2)
    brute = np.floor(tau * bw / M).astype(np.int64)
    brute = np.clip(brute, 0, K - 1)
    mism = np.nonzero(brute != d)[0]
    # disagreements only next to a segment boundary (float rounding there)
    bpos = set(np.nonzero(np.diff(d[order]))[0])
    pos = {int(np.nonzero(order == m)[0][0]) for m in mism}
    for p in pos:
        assert any(abs(p - b) <= 1 for b in bpos), (p, len(mism))
    # ... and only on a bin whose exact segment floor(tau*fs/M) differs from
    # a frequency neighbour's (the bin next to where tau*fs/M crosses an
    # integer), off by exactly one segment
    bo = brute[order]
    for p in pos:
        nb = [bo[q] for q in (p - 1, p + 1) if 0 <= q < N]
        assert any(v != bo[p] for v in nb), p
    assert np.all(np.abs(brute[mism] - d[mism]) == 1)


def test_dm_zero_is_single_segment():
    d, K, holes = oracle.delay_points(64e6, 0.0, 1368e6, 256)
    assert K == 1 and holes == 0 and not d.any()


# ------------------------------------------------------------------ chirp --
def test_chirp_matches_golden_and_is_unit_modulus(golden):
    arrays, meta = golden
    for i, (bw, dm, f0, M) in enumerate(meta["h_cases"]):
        c = oracle.chirp(bw, dm, f0, M, oracle.CHIRP_REF_F32W)
        np.testing.assert_array_equal(c.astype(np.complex64), arrays[f"chirp_ref_{i}"])
        assert np.allclose(np.abs(c), 1, atol=1e-6)
        ce = oracle.chirp(bw, dm, f0, M, oracle.CHIRP_EXACT_F64)
        assert np.allclose(np.abs(ce), 1, atol=1e-12)


def test_chirp_exact_phase_closed_form():
    bw, dm, f0, M = 8e6, 300.0, 1396e6, 256
    c = oracle.chirp(bw, dm, f0, M, oracle.CHIRP_EXACT_F64)
    i = np.arange(M)
    f = bw * i / M
    f0f = float(np.float32(f0))
    ph = 2 * np.pi * KDM * float(np.float32(dm)) * f * f / (f0f * f0f * (f0f + f))
    idx = np.where(i < M // 2, M // 2 + i, i - M // 2)
    expect = np.zeros(M, complex)
    expect[idx] = np.exp(-1j * ph)
    assert np.max(np.abs(c - expect)) < 1e-9


def test_dedisp_params_golden_and_closed_form(golden):
    arrays, meta = golden
    for i, (bw, dm, f0, M) in enumerate(meta["h_cases"]):
        for mode, key in ((oracle.CHIRP_REF_F32W, "ref"), (oracle.CHIRP_EXACT_F64, "exact")):
            H = oracle.dedisp_params(bw, dm, f0, M, mode)
            assert rel_l2(H, arrays[f"H_{key}_{i}"]) < 1e-12
            c = oracle.chirp(bw, dm, f0, M, mode)
            # even bins of the 2M-point DFT of the zero-padded M-point IDFT
            assert np.max(np.abs(H[0::2] - M * c)) < 1e-9 * M


def test_dedisp_params_textbook_dft():
    bw, dm, f0, M = 1e6, 1.0, 400e6, 16
    c = oracle.chirp(bw, dm, f0, M)
    n = np.arange(M)
    h = np.array([np.sum(c * np.exp(2j * np.pi * n * t / M)) for t in range(M)])   # unnormalised IDFT
    k = np.arange(2 * M)
    H = np.array([np.sum(h * np.exp(-2j * np.pi * kk * n / (2 * M))) for kk in k])
    assert rel_l2(oracle.dedisp_params(bw, dm, f0, M), H) < 1e-12


# -------------------------------------------------------------------- fft --
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert rel_l2(oracle.fft(x, -1), W @ x) < 1e-13
    assert rel_l2(oracle.fft(x, +1), W.conj() @ x) < 1e-13


def test_fft_parseval_and_numpy():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(4096) + 1j * rng.standard_normal(4096)
    X = oracle.fft(x)
    assert abs(np.sum(np.abs(X) ** 2) / 4096 - np.sum(np.abs(x) ** 2)) < 1e-9 * np.sum(np.abs(x) ** 2)
    assert rel_l2(X, np.fft.fft(x)) < 1e-13


# ---------------------------------------------------------- pack / unpack --
def test_unpack_pack_exhaustive():
    u = np.arange(65536, dtype=np.uint16)
    pairs = np.stack([u, u[::-1]], axis=1).reshape(-1)
    c = oracle.unpack_u16(pairs)
    expect_re = np.where(u != 0, u.astype(np.float32) - 32768, 0)
    np.testing.assert_array_equal(c.real, expect_re)
    back = oracle.pack_u16(c).reshape(-1, 2)
    exp_back = np.where(u == 32768, 0, u)           # value 0 collides with "no data"
    np.testing.assert_array_equal(back[:, 0], exp_back)
    # truncation toward zero of v + 32768 and saturation
    v = np.array([0.0, 1.9, -1.9, 40000.0, -40000.0, np.nan, 32767.9, -32767.5], np.float32)
    got = oracle.pack_u16(v.astype(np.complex64))[0::2]
    np.testing.assert_array_equal(got, [0, 32769, 32766, 65535, 0, 0, 65535, 0])


def test_intensity_bit_exact():
    rng = np.random.default_rng(3)
    a = (rng.standard_normal(1000) + 1j * rng.standard_normal(1000)).astype(np.complex64) * 100
    b = (rng.standard_normal(1000) + 1j * rng.standard_normal(1000)).astype(np.complex64) * 100
    s = a.real * a.real
    s = s + a.imag * a.imag
    s = s + b.real * b.real
    s = s + b.imag * b.imag
    np.testing.assert_array_equal(oracle.intensity(a, b), np.sqrt(s))


# ---------------------------------------------------------------- filters --
def _unpack64(u):
    return oracle.unpack_u16(u).astype(np.complex128)


def test_msosm_matches_golden(golden):
    arrays, meta = golden
    c = meta["msosm"]
    x = _unpack64(arrays["msosm_in_u16"])
    d, K, _ = oracle.delay_points(c["bw"], c["dm"], c["f0"], c["M"])
    for mode, key in ((oracle.CHIRP_REF_F32W, "ref"), (oracle.CHIRP_EXACT_F64, "exact")):
        H = oracle.dedisp_params(c["bw"], c["dm"], c["f0"], c["M"], mode)
        y = oracle.msosm_run(x, c["M"], d, H)
        assert rel_l2(y, arrays[f"msosm_out_{key}"]) < 1e-6


def test_osm_matches_golden_and_direct_convolution(golden):
    arrays, meta = golden
    c = meta["osm"]
    M = c["M"]
    x = _unpack64(arrays["osm_in_u16"])
    H = oracle.dedisp_params(c["bw"], c["dm"], c["f0"], M)
    y = oracle.msosm_run(x, M, np.zeros(2 * M, np.int32), H)
    assert rel_l2(y, arrays["osm_out"]) < 1e-6
    # K=1: y = 1/2 * (x * h), h = M-point IDFT of the chirp (length M <= M+1)
    h = np.fft.ifft(oracle.chirp(c["bw"], c["dm"], c["f0"], M))
    direct = 0.5 * np.convolve(x[:8 * M], h)[:8 * M]
    assert rel_l2(y[:8 * M], direct) < 1e-12


def test_msosm_linearity_and_block_shift():
    bw, dm, f0, M = 2e6, 3.3, 300e6, 128
    d, K, _ = oracle.delay_points(bw, dm, f0, M)
    H = oracle.dedisp_params(bw, dm, f0, M)
    rng = np.random.default_rng(11)
    L = 96 * M
    x1 = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    x2 = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    y1, y2 = oracle.msosm_run(x1, M, d, H), oracle.msosm_run(x2, M, d, H)
    y12 = oracle.msosm_run(2.5 * x1 - 1j * x2, M, d, H)
    assert rel_l2(y12, 2.5 * y1 - 1j * y2) < 1e-12
    xs = np.concatenate([np.zeros(3 * M), x1[:-3 * M]])
    ys = oracle.msosm_run(xs, M, d, H)
    assert rel_l2(ys[3 * M:], y1[:-3 * M]) < 1e-12
    # sub-range of output blocks equals the full run
    part = oracle.msosm_run(x1, M, d, H, b0=50, nb=20)
    assert rel_l2(part, y1[50 * M:70 * M]) < 1e-13


def test_msosm_multi_trial_equals_single_trial_runs():
    # shared forward spectra: every trial's rows equal its own single-trial
    # run bit for bit, for windows that start at the stream, inside the delay
    # line of the largest trial, and past it
    bw, f0, M = 2e6, 300e6, 128
    dms = [0.0, 1.1, 3.3, 6.0]
    rng = np.random.default_rng(12)
    L = 80 * M
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    d = np.stack([oracle.delay_points(bw, dm, f0, M)[0] for dm in dms])
    H = np.stack([oracle.dedisp_params(bw, dm, f0, M) for dm in dms])
    assert d.max() > 20
    for b0, nb in ((0, 5), (7, 3), (60, 20)):
        y = oracle.msosm_run_multi(x, M, d, H, b0=b0, nb=nb)
        for t in range(len(dms)):
            np.testing.assert_array_equal(y[t], oracle.msosm_run(x, M, d[t], H[t], b0=b0, nb=nb))


def test_msosm_impulse_lands_after_segment_delays():
    # an impulse's response energy sits in blocks t0 .. t0 + K (delays 0..K-1)
    bw, dm, f0, M = 2e6, 3.3, 300e6, 128
    d, K, _ = oracle.delay_points(bw, dm, f0, M)
    H = oracle.dedisp_params(bw, dm, f0, M)
    L = (K + 40) * M
    x = np.zeros(L, complex)
    t0 = 10 * M + 5
    x[t0] = 1
    y = oracle.msosm_run(x, M, d, H)
    e = np.abs(y) ** 2
    assert e[:10 * M].sum() < 1e-20
    assert e[(10 + K + 2) * M:].sum() < 1e-6 * e.sum()
    # the scale is 1/2: output energy ~ 1/4 of input energy times sum|h|^2 = 1
    assert abs(e.sum() - 0.25) < 0.05


def test_cpu_baseline_f32_matches_fp64():
    bw, f0, M = 2e6, 300e6, 128
    dms = [0.0, 1.1, 3.3]
    d = np.stack([oracle.delay_points(bw, dm, f0, M)[0] for dm in dms])
    H = np.stack([oracle.dedisp_params(bw, dm, f0, M) for dm in dms])
    rng = np.random.default_rng(5)
    L = 64 * M
    u = np.clip(np.round(32768 + 800 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    cs, y = oracle.cpu_msosm_f32(u, M, d, H.astype(np.complex64), want_output=True, nthreads=2)
    x = _unpack64(u)
    tot = 0.0
    for i in range(len(dms)):
        ref = oracle.msosm_run(x, M, d[i], H[i])
        assert rel_l2(y[i], ref) < 1e-5
        tot += np.sum(np.abs(ref) ** 2)
    assert abs(cs - tot) < 1e-4 * tot


# ------------------------------------------------------------------- fold --
def test_fold_matches_golden(golden):
    arrays, meta = golden
    for j, c in enumerate(meta["fold"]):
        f = oracle.Fold(c["period"], c["fs"], c["size"], c["time_bin"])
        f.discard_samples(c["nd"])
        seq = arrays[f"fold_in_{j}"]
        traces = []
        for k in range(c["calls"]):
            f.fold_data(seq[k * c["size"]:(k + 1) * c["size"]])
            traces.append(f.trace())
        f.bins()
        assert f.period_samples == c["P"]
        np.testing.assert_array_equal(f.raw(), arrays[f"fold_raw_{j}"])
        np.testing.assert_array_equal(f.folded(), arrays[f"fold_bins_{j}"])
        np.testing.assert_array_equal(np.concatenate(traces), arrays[f"fold_trace_{j}"])


def test_fold_skip_schedule_closed_form():
    # P*fs = 1000.25: drift 0.25/period; a sample is dropped when the float
    # sum exceeds 1.0 -- first after the 5th period (0.25*4 == 1.0 is not > 1), then
    # every 4 periods.  Impulses on every period start all land in bin 0.
    P, nper = 1000, 40
    starts, pos, diff = [], 0, np.float32(0)
    for k in range(nper):
        starts.append(pos)
        pos += P
        diff = np.float32(diff + np.float32(0.25))
        if diff > 1:
            diff = np.float32(diff - 1)
            pos += 1
    assert starts[5] - starts[4] == P + 1 and starts[4] - starts[3] == P
    size = pos
    x = np.zeros(size, np.float32)
    x[starts] = 1
    f = oracle.Fold(1.0, 1000.25, size, -1)
    f.discard_samples(0)
    f.fold_data(x)
    raw = f.raw()
    assert raw[0] == nper and raw[1:].sum() == 0


def test_fold_call_split_invariance():
    rng = np.random.default_rng(2)
    x = rng.random(60000).astype(np.float32)
    a = oracle.Fold(1.0, 1234.5, 60000, 100)
    a.discard_samples(3000)
    a.fold_data(x)
    b = oracle.Fold(1.0, 1234.5, 6000, 100)
    b.discard_samples(3000)
    for k in range(10):
        b.fold_data(x[k * 6000:(k + 1) * 6000])
    np.testing.assert_array_equal(a.raw(), b.raw())


@pytest.mark.parametrize("nd,skipped", [(1, 1), (999, 1), (1000, 1), (1001, 2), (2500, 3), (3000, 3), (3001, 4)])
def test_fold_discard_count_closed_form(nd, skipped):
    # fold_gpu.cpp:29-33: discard_samples(Nd) skips the first ceil(Nd / P)
    # whole periods.  P*fs = 1000 exactly (no drift); period k holds the
    # value k+1 everywhere, so every raw bin sums the periods kept.
    P, nper = 1000, 10
    x = np.repeat(np.arange(1, nper + 1, dtype=np.float32), P)
    f = oracle.Fold(1.0, 1000.0, x.size, -1)
    f.discard_samples(nd)
    f.fold_data(x)
    want = float(sum(range(skipped + 1, nper + 1)))
    raw = f.raw()
    assert raw.shape == (P,) and np.all(raw == want), (raw[:3], want)


@pytest.mark.parametrize("time_bin,per_bin", [(100, 10), (250, 4), (300, 3), (999, 1)])
def test_fold_bins_average_closed_form(time_bin, per_bin):
    # fold_gpu.cpp:175-194: bins_per_time_bin = P / time_bin (integer
    # division); output bin i is the mean of raw bins [i*b, (i+1)*b), so only
    # the first time_bin*b raw bins are used.  A ramp x[n] = n mod P folded
    # over nper periods gives raw[j] = nper*j and bin i = nper*(b*i + (b-1)/2).
    P, nper = 1000, 6
    x = np.tile(np.arange(P, dtype=np.float32), nper)
    f = oracle.Fold(1.0, 1000.0, x.size, time_bin)
    f.discard_samples(0)
    f.fold_data(x)
    f.bins()
    np.testing.assert_array_equal(f.raw(), nper * np.arange(P, dtype=np.float32))
    i = np.arange(time_bin)
    want = (nper * (per_bin * i + (per_bin - 1) / 2.0)).astype(np.float32)
    np.testing.assert_array_equal(f.folded(), want)
