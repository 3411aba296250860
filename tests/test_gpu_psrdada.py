"""Dual-polarisation PSRDADA recording -> coherent dedispersion of each
polarisation -> intensity -> fold, end to end through the msosm_psrdada
program (C++ on the drop-in classes), against the FP64 oracle: MS-OSM per
polarisation, the reference's float intensity and the reference's
sequential fold state machine, on the same synthetic recording (a dispersed
pulsar in both polarisations, with a leading stretch of "no data" frames
that the reader must skip)."""
import os
import subprocess

import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle
from msosm_gpu_amd import psrdada, simulate

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "msosm_gpu_amd", "bin", "msosm_psrdada")


@pytest.mark.timeout(300)
def test_recording_end_to_end_vs_oracle(device, tmp_path):
    if not os.path.exists(APP):
        pytest.fail("msosm_psrdada is not built (run __graft_entry__.build())")
    bw_mhz, fc_mhz, dm, period = 16.0, 1400.0, 10.0, 0.005
    bw, f0 = bw_mhz * 1e6, (fc_mhz - bw_mhz / 2) * 1e6
    fftpoint, count, time_bin = 1024, 256, 256
    M = fftpoint // 2
    n = count * M
    reads = 24
    L = reads * n
    pols = [simulate.pulsar_u16(L, bw, f0, dm, period, duty=0.05, snr=0.3, seed=s) for s in (11, 12)]
    lead = np.zeros(2 * n, np.uint16)  # one read of "no data" before the recording starts
    rec = tmp_path / "pulsar.dada"
    psrdada.write_dada(rec, np.concatenate([lead, pols[0]]), np.concatenate([lead, pols[1]]), bw_mhz, fc_mhz)
    out = tmp_path / "folded.txt"
    print(f"running msosm_psrdada on {L} samples per polarisation", flush=True)
    r = subprocess.run([APP, "--file", str(rec), "--dm", str(dm), "--period", str(period), "--fftpoint",
                        str(fftpoint), "--count", str(count), "--time-bin", str(time_bin), "--out", str(out)],
                       capture_output=True, text=True, timeout=240)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"blocks used {reads} of {reads + 1}" in r.stdout
    got = np.loadtxt(out, dtype=np.float64)
    assert got.shape == (time_bin,)

    d, K, _ = oracle.delay_points(bw, dm, f0, M)
    H = oracle.dedisp_params(bw, dm, f0, M)
    y = [oracle.msosm_run(oracle.unpack_u16(p).astype(np.complex128), M, d, H).astype(np.complex64) for p in pols]
    inten = oracle.intensity(y[0], y[1])
    fold = oracle.Fold(period, bw, n, time_bin)
    fold.discard_samples(oracle.nd(bw, dm, f0))
    for c in range(reads):
        fold.fold_data(inten[c * n:(c + 1) * n])
    fold.bins()
    ref = fold.folded().astype(np.float64)
    err = np.max(np.abs(got - ref) / np.abs(ref))
    print(f"folded profile vs FP64 oracle: max relative difference {err:.2e}")
    assert err < 2e-5
    # the pulse stands out of the folded profile
    med = np.median(ref)
    mad = np.median(np.abs(ref - med)) * 1.4826
    assert (got.max() - med) / mad > 10
