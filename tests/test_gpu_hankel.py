"""The delay-line accumulate as per-bin complex GEMMs on the matrix cores
(FORM_HANKEL_*): equality with the one-hot gather form on the same operands
up to the rounding of the stored products, the error of each operand form
against the FP64 same-plan oracle, many DM tiles with their delay bands, and
the refusal of plans whose delay span does not fit."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle
from test_gpu_engine import make_input, oracle_out, rel_l2, run_engine

pytestmark = pytest.mark.gpu

MULTI = (64e6, [0.0, 0.4, 3.0, 7.5, 12.0, 18.25, 24.0, 29.0], 1368e6)


def test_operand_forms_vs_fp64(device):
    """FP16 operands within the mixed gate; BF16 hi+lo about 2^-16 relative;
    single BF16 operands (2^-8 relative rounding) outside the 1e-3 budget,
    which is why the split form exists."""
    bw, dms, f0 = MULTI
    u = make_input(16 * 8192, 99)
    refs = [oracle_out(bw, dm, f0, 8192, u) for dm in dms]
    errs = {}
    for form in (ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16, ms.FORM_HANKEL_BF16X3):
        print(f"form {form}", flush=True)
        y = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=form), u, device=device)
        errs[form] = max(rel_l2(y[i], refs[i]) for i in range(len(dms)))
    print("max rel L2 vs FP64:", errs)
    assert errs[ms.FORM_HANKEL_F16] < 1e-3
    assert errs[ms.FORM_HANKEL_BF16X3] < 5e-5
    assert 1e-3 < errs[ms.FORM_HANKEL_BF16] < 2e-2
    assert errs[ms.FORM_HANKEL_BF16X3] < errs[ms.FORM_HANKEL_BF16] / 30


def test_equals_gather_form(device):
    """Same FP16 operands as the fused gather form: the only difference is
    the FP16 rounding of the stored products (<= 2^-11 relative each)."""
    bw, dms, f0 = MULTI
    u = make_input(16 * 8192, 5)
    a = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_HANKEL_F16), u, device=device)
    b = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_FUSED_F16), u, device=device)
    for i in range(len(dms)):
        assert rel_l2(a[i], b[i]) < 4e-4, (i, rel_l2(a[i], b[i]))
    c = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_HANKEL_BF16X3), u, device=device)
    d = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=4, form=ms.FORM_FUSED_F32), u, device=device)
    for i in range(len(dms)):
        assert rel_l2(c[i], d[i]) < 5e-5, (i, rel_l2(c[i], d[i]))


def test_many_trials_delay_bands(device):
    """256 trials on [0, 100] at 64 MHz (K up to 152): 16 DM tiles whose
    delay bands span one or two 16-wide k steps, trials past the last full
    tile, two calls so the Hankel blocks reach into the previous call."""
    bw, f0 = 64e6, 1368e6
    dms = list(np.linspace(0, 100, 250))
    u = make_input(2 * 16 * 8192, 77)
    p16 = ms.Plan(bw, dms, f0, fftpoint=16384, count=16, form=ms.FORM_HANKEL_F16)
    assert p16.K_max == 152
    a = run_engine(p16, u, device=device)
    b = run_engine(ms.Plan(bw, dms, f0, fftpoint=16384, count=16, form=ms.FORM_FUSED_F16), u, device=device)
    worst = max(rel_l2(a[i], b[i]) for i in range(len(dms)))
    print(f"max rel L2 Hankel FP16 vs gather FP16 over {len(dms)} trials: {worst:.2e}")
    assert worst < 4e-4
    for i in (0, 97, 249):
        assert rel_l2(a[i], oracle_out(bw, dms[i], f0, 8192, u)) < 1e-3


def test_refuses_delay_span_beyond_lds(device):
    p = ms.Plan(64e6, [300.0], 1368e6, fftpoint=8192, count=2, form=ms.FORM_HANKEL_F16)
    assert p.K_max == 909
    with pytest.raises(ms.MsosmError):
        ms.Engine(p)
