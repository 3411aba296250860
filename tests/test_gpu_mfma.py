"""GPU parity of the mixed form whose inverse FFT runs its DFT stages on the
matrix cores (MSOSM_FORM_FUSED_F16_MFMA, FFT size 16384) against the FP64
same-plan oracle, gate 1e-3 relative L2 per trial; block-split and
host/device-input invariance bit for bit; closeness to the vector-ALU mixed
form."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle
from test_gpu_engine import make_input, oracle_out, rel_l2, run_engine

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-3
BW, F0 = 64e6, 1368e6
DMS = [0.0, 0.4, 3.0, 7.5, 12.0, 18.25, 24.0, 29.0]


@pytest.mark.parametrize("count", [4, 5, 13])
def test_mfma_multi_dm_vs_fp64(device, count):
    # 8 trials, K from 1 (DM 0) to 44; counts that leave a workgroup with
    # fewer than 8 blocks
    p = ms.Plan(BW, DMS, F0, fftpoint=16384, count=count, form=ms.FORM_FUSED_F16_MFMA)
    u = make_input(count * 4 * 8192, 99 + count)
    print(f"mfma multi-dm count={count} K_max={p.K_max}", flush=True)
    y = run_engine(p, u, device=device)
    errs = []
    for i, dm in enumerate(DMS):
        errs.append(rel_l2(y[i], oracle_out(BW, dm, F0, 8192, u)))
    print("  rel L2 per trial: " + " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) < MIXED_TOL


def test_mfma_exact_chirp_mode(device):
    dm = 57.3
    p = ms.Plan(BW, dm, F0, fftpoint=16384, count=8, form=ms.FORM_FUSED_F16_MFMA, chirp=ms.CHIRP_EXACT_F64)
    u = make_input(24 * 8192, 17)
    y = run_engine(p, u, device=device)[0]
    err = rel_l2(y, oracle_out(BW, dm, F0, 8192, u, mode=oracle.CHIRP_EXACT_F64))
    print(f"mfma exact chirp: {err:.2e}", flush=True)
    assert err < MIXED_TOL


def test_mfma_close_to_vector_form(device):
    u = make_input(16 * 8192, 5)
    a = run_engine(ms.Plan(BW, DMS, F0, fftpoint=16384, count=8, form=ms.FORM_FUSED_F16_MFMA), u, device=device)
    b = run_engine(ms.Plan(BW, DMS, F0, fftpoint=16384, count=8, form=ms.FORM_FUSED_F16), u, device=device)
    for i in range(len(DMS)):
        assert rel_l2(a[i], b[i]) < MIXED_TOL


def test_mfma_block_split_and_input_invariance(device):
    # 1 call of 16 blocks == 16 calls of 1 block == 4 calls of 4, and host
    # input == device input (bit for bit)
    u = make_input(32 * 8192, 3)
    outs = [run_engine(ms.Plan(BW, DMS[:3], F0, fftpoint=16384, count=c, form=ms.FORM_FUSED_F16_MFMA), u,
                       device=device) for c in (16, 1, 4)]
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])
    h = run_engine(ms.Plan(BW, DMS[:3], F0, fftpoint=16384, count=16, form=ms.FORM_FUSED_F16_MFMA), u,
                   device_input=False, device=device)
    np.testing.assert_array_equal(outs[0], h)


def test_mfma_reset_restarts_stream(device):
    u = make_input(8 * 8192, 6)
    p = ms.Plan(BW, DMS[:2], F0, fftpoint=16384, count=8, form=ms.FORM_FUSED_F16_MFMA)
    eng = ms.Engine(p)
    eng.filter_u16(u.view(np.uint32))
    a = eng.output_c64()
    eng.filter_u16(u.view(np.uint32))
    eng.reset()
    eng.filter_u16(u.view(np.uint32))
    b = eng.output_c64()
    eng.close()
    np.testing.assert_array_equal(a, b)


def test_mfma_needs_fft_size_16384(device):
    with pytest.raises(ms.MsosmError):
        ms.Engine(ms.Plan(BW, DMS[:2], F0, fftpoint=8192, count=4, form=ms.FORM_FUSED_F16_MFMA))


@pytest.mark.parametrize("kind", ["full_scale_noise", "pulse_train_32", "dc"])
def test_mfma_full_scale_inputs(device, kind):
    # Inputs at the uint16 limits: maximum power spread over all bins, a
    # period-32 pulse train (its spectrum is one pass-0 column: every
    # product of that column adds coherently) and a constant.  The FP16
    # exchanges keep their headroom: outputs finite and within the gate.
    L = 12 * 8192
    rng = np.random.default_rng(11)
    if kind == "full_scale_noise":
        u = rng.choice(np.array([1, 65535], np.uint16), 2 * L)
    elif kind == "pulse_train_32":
        u = np.full(2 * L, 32768, np.uint16)
        u[0::64] = 65535
        u[1::64] = 65535
    else:
        u = np.full(2 * L, 65535, np.uint16)
    dm = 57.3
    p = ms.Plan(BW, dm, F0, fftpoint=16384, count=4, form=ms.FORM_FUSED_F16_MFMA)
    y = run_engine(p, u, device=device)[0]
    assert np.all(np.isfinite(y))
    err = rel_l2(y, oracle_out(BW, dm, F0, 8192, u))
    print(f"mfma {kind}: rel L2 {err:.2e}", flush=True)
    assert err < MIXED_TOL


def test_mfma_blocks_per_workgroup_choice(device):
    # 512 trials run 16 blocks per workgroup (20 per call: a full chunk and a
    # partial one), 8 trials alone run 1: each trial's output must not depend
    # on that choice (bit for bit)
    dms = np.linspace(0.0, 50.0, 512).astype(np.float32)
    u = make_input(2 * 20 * 8192, 43)
    big = ms.Plan(BW, dms, F0, fftpoint=16384, count=20, form=ms.FORM_FUSED_F16_MFMA)
    small = ms.Plan(BW, dms[:8], F0, fftpoint=16384, count=20, form=ms.FORM_FUSED_F16_MFMA)
    a = run_engine(big, u, device=device)
    b = run_engine(small, u, device=device)
    np.testing.assert_array_equal(a[:8], b)
    err = rel_l2(a[511], oracle_out(BW, float(dms[511]), F0, 8192, u))
    print(f"mfma 512 trials, DM 50: rel L2 {err:.2e}", flush=True)
    assert err < MIXED_TOL
