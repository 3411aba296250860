"""The C++ drop-in facade (include/msosm/*.h), driven by a C++ program the
way the reference's drivers drive its classes, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "facade_check")


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_facade_program(device, tmp_path):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_check is not built (run __graft_entry__.build())")
    rng = np.random.Generator(np.random.Philox(77))
    L = 1 << 15
    u = np.clip(np.round(32768 + 900 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    inp = tmp_path / "in.u16"
    u.tofile(inp)
    pre = str(tmp_path / "out")
    r = subprocess.run([BIN, str(inp), str(L), pre], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    meta = open(pre + ".meta").read()
    assert "batch M 512 M_min 256 Nd 7743 K 16 count 8" in meta
    bw, f0, M = 16e6, 1392e6, 512
    x = oracle.unpack_u16(u).astype(np.complex128)

    def ref(dm, M=M, osm=False):
        d = np.zeros(2 * M, np.int32) if osm else oracle.delay_points(bw, dm, f0, M)[0]
        return oracle.msosm_run(x, M, d, oracle.dedisp_params(bw, dm, f0, M))

    y = np.fromfile(pre + ".batch.c64", np.complex64)
    assert rel_l2(y, ref(10.0)) < 1e-5
    conc = np.fromfile(pre + ".conc.c64", np.complex64).reshape(-1, 3, 8 * M)
    for i, dm in enumerate((5.0, 10.0, 20.0)):
        assert rel_l2(conc[:, i].reshape(-1), ref(dm)) < 1e-5
    cu = np.fromfile(pre + ".conc.u16", np.uint16)
    np.testing.assert_array_equal(cu, oracle.pack_u16(conc.reshape(-1)))
    loop = np.fromfile(pre + ".loop.c64", np.complex64).reshape(-1, 3, 8 * M)
    np.testing.assert_array_equal(loop, conc)
    osm_line = [ln for ln in meta.splitlines() if ln.startswith("osm")][0].split()
    Mo, co = int(osm_line[2]), int(osm_line[4])
    yo = np.fromfile(pre + ".osm.c64", np.complex64)
    n = yo.size
    assert n == (L // (co * Mo)) * co * Mo
    assert rel_l2(yo, ref(2.0, M=Mo, osm=True)[:n]) < 1e-5
    # fold chain: intensity of (y, y) on the GPU, then the fold, bit for bit
    inten = np.fromfile(pre + ".inten.f32", np.float32)
    np.testing.assert_array_equal(inten, oracle.intensity(y, y))
    f = oracle.Fold(0.0005, bw, 8 * M, 100)
    f.discard_samples(7743)
    for c in range(L // (8 * M)):
        f.fold_data(inten[c * 8 * M:(c + 1) * 8 * M])
    f.bins()
    np.testing.assert_array_equal(np.fromfile(pre + ".fold.f32", np.float32), f.folded())
    txt = open(pre + ".fold.txt").read().split()
    assert len(txt) == 100
    # DM stream engine: per-trial hand-out of the same outputs as the
    # concurrent engine, bit for bit, in both formats
    ds = np.fromfile(pre + ".dmstream.c64", np.complex64).reshape(-1, 3, 8 * M)
    np.testing.assert_array_equal(ds, conc)
    np.testing.assert_array_equal(np.fromfile(pre + ".dmstream.u16", np.uint16), cu)
    # independent streams with their own FFT sizes (M 512 and 1024)
    assert "bstream M 512 1024 fftpoint 1024 2048" in meta
    b0 = np.fromfile(pre + ".bstream0.c64", np.complex64)
    b1 = np.fromfile(pre + ".bstream1.c64", np.complex64)
    calls = b1.size // (4 * 1024)
    assert calls >= 4 and b0.size == calls * 4 * 512
    assert rel_l2(b0, ref(10.0)[:b0.size]) < 1e-5
    assert rel_l2(b1, ref(20.0, M=1024)[:b1.size]) < 1e-5
    # the drivers' simulated signal through the facade class: the library's
    # generator, four periods of 80,000 samples
    assert "sim Np 80000 Nd 7743 range 1 size 320000" in meta
    sig, u16, _ = ms.refsim_generate(bw, 10.0, f0, 0.005, repeat=4)
    np.testing.assert_array_equal(np.fromfile(pre + ".sim.u16", np.uint16), u16)
    np.testing.assert_array_equal(np.fromfile(pre + ".sim.c64", np.float32), sig.view(np.float32))


def test_facade_reference_default_plan(device, tmp_path):
    """MSOSM_GPU_BATCH at the recording program's plan (fftpoint 2^21, count
    1, msosm_psrdada.cpp:94-95): 128 MHz, DM 100 (K 5), eight calls, every
    output block against the FP64 same-plan oracle.  The facade runs the
    fused FP32 form, now four-step at this size."""
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_check is not built (run __graft_entry__.build())")
    rng = np.random.Generator(np.random.Philox(78))
    M = 1 << 20
    L = 8 * M
    u = np.clip(np.round(32768 + 900 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    inp = tmp_path / "in.u16"
    u.tofile(inp)
    pre = str(tmp_path / "out")
    r = subprocess.run([BIN, str(inp), str(L), pre, "default_plan"], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    meta = open(pre + ".meta").read()
    assert "big M 1048576 Nd 4976580 K 5 count 1" in meta, meta
    y = np.fromfile(pre + ".big.c64", np.complex64)
    assert y.size == L
    bw, dm, f0 = 128e6, 100.0, 1336e6
    x = oracle.unpack_u16(u).astype(np.complex128)
    ref = oracle.msosm_run(x, M, oracle.delay_points(bw, dm, f0, M)[0], oracle.dedisp_params(bw, dm, f0, M))
    err = rel_l2(y, ref)
    print(f"facade at FFT 2^21, count 1: rel L2 vs FP64 {err:.2e}")
    assert err < 1e-5


def test_facade_writer_threads(device, tmp_path):
    """get_output_thread with two writer threads (the reference driver's
    multiple_dm_concurrent_msosm.cpp:117-146 pattern): the outputs go through
    the output ring asynchronously; every call's output reaches its writer
    exactly once, in call order, equal to the synchronous get_output."""
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/facade_check is not built (run __graft_entry__.build())")
    rng = np.random.Generator(np.random.Philox(79))
    L = 1 << 17
    u = np.clip(np.round(32768 + 900 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    inp = tmp_path / "in.u16"
    u.tofile(inp)
    pre_t, pre_s = str(tmp_path / "thr"), str(tmp_path / "sync")
    for pre, mode in ((pre_t, ["writer_threads"]), (pre_s, [])):
        r = subprocess.run([BIN, str(inp), str(L), pre] + mode, capture_output=True, text=True, timeout=240)
        print(r.stdout, r.stderr)
        assert r.returncode == 0, r.stderr
    calls = L // (8 * 512)
    assert f"writer_threads calls {calls}" in open(pre_t + ".meta").read()
    per = 3 * 8 * 512 * 2                       # uint16 values per call (3 trials)
    sync = np.fromfile(pre_s + ".conc.u16", np.uint16).reshape(calls, per)
    for tid in (0, 1):
        got = np.fromfile(pre_t + f".thread{tid}.u16", np.uint16)
        want = sync[tid::2].reshape(-1)
        assert got.size == want.size, (tid, got.size, want.size)
        np.testing.assert_array_equal(got, want)
