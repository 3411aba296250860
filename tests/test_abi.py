"""The C ABI library loads and exports every function include/msosm_c.h
declares (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

import msosm_gpu_amd as ms

HEADER = os.path.join(ms.INCLUDE_DIR, "msosm_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(msosm_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_declared_symbols():
    L = ms.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the python binding covers the whole header
    assert set(names) <= set(ms.SIGNATURES), set(names) - set(ms.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", ms.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (msosm_\w+)", out))
    assert set(names) <= exported


def test_version_string():
    assert "gfx950" in ms.version()


def test_library_contains_gfx950_code():
    # the offload bundle inside the shared object names its target
    data = open(ms.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_no_gpu_call_needed_for_plans():
    # planning works without a GPU; engine creation reports a clean error when none is present
    p = ms.Plan(16e6, 10.0, 1392e6, fftpoint=1024, count=4)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        h = ctypes.c_void_p()
        rc = ms.lib().msosm_engine_create(p.h, None, ctypes.byref(h))
        assert rc != 0 and h.value is None


def test_engine_rejects_a_ring_beyond_32bit_offsets():
    # the shared-ring forms address the spectra ring with 32-bit byte offsets:
    # 151 blocks of history + 40000 per call x 16384 points x 8 B > 4 GiB
    # (checked before any device is touched)
    p = ms.Plan(64e6, [100.0], 1368e6, fftpoint=16384, count=40000, form=ms.FORM_FUSED_F32)
    with pytest.raises(ms.MsosmError, match="below 4 GiB"):
        ms.Engine(p)


def test_noalloc_fft_refuses_oversized_scratch_on_the_host():
    # checked before any device call: batch * n beyond the fixed 2^22-sample
    # scratch, bad sizes and bad directions are argument errors
    L = ms.lib()
    p = ctypes.c_void_p(0x1000)
    assert L.msosm_k_fft_c2c_noalloc(p, p, 1 << 21, 3, 1 << 21, 1 << 21, -1, None) == ms.ERR_ARG
    assert b"exceeds the fixed scratch" in L.msosm_last_error()
    assert L.msosm_k_fft_c2c_noalloc(p, p, 1 << 23, 1, 1 << 23, 1 << 23, 1, None) == ms.ERR_ARG
    assert L.msosm_k_fft_c2c_noalloc(p, p, 3000, 1, 3000, 3000, 1, None) == ms.ERR_ARG
    assert L.msosm_k_fft_c2c_noalloc(p, p, 1024, 1, 1024, 1024, 0, None) == ms.ERR_ARG


def test_single_bf16_hankel_form_is_refused():
    # single BF16 operands miss the 1e-3 mixed-precision gate: the engine
    # refuses the form before touching a device
    p = ms.Plan(64e6, [10.0], 1368e6, fftpoint=16384, count=2, form=ms.FORM_HANKEL_BF16)
    with pytest.raises(ms.MsosmError, match="1e-3"):
        ms.Engine(p)
