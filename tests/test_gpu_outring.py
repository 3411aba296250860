"""Output streaming ring between an engine and a consumer thread: every
call's output reaches the consumer exactly once, in order, equal to the
engine's synchronous output; the producer blocks (with a time limit) while
every slot is taken."""
import threading
import time

import numpy as np
import pytest

import msosm_gpu_amd as ms
from test_gpu_engine import make_input

pytestmark = pytest.mark.gpu

BW, F0, DMS = 16e6, 1392e6, [5.0, 10.0, 20.0]


def _setup(device, calls, seed):
    import torch
    p = ms.Plan(BW, DMS, F0, fftpoint=1024, count=8, form=ms.FORM_FUSED_F16)
    n = p.count * p.M
    u = make_input(calls * n, seed)
    words = u.view(np.uint32)
    dev = torch.from_numpy(words.view(np.int32).copy()).to(device)
    return p, n, words, dev


@pytest.mark.parametrize("fmt", [ms.OutRing.UINT16, ms.OutRing.COMPLEX64])
def test_stream_in_order_and_complete(device, fmt):
    calls = 24
    p, n, words, dev = _setup(device, calls, 41)
    ref_eng = ms.Engine(p)
    refs = []
    for c in range(calls):
        ref_eng.filter_u16(words[c * n:(c + 1) * n])
        refs.append((ref_eng.output_u16() if fmt == ms.OutRing.UINT16 else ref_eng.output_c64()).tobytes())
    ref_eng.close()
    eng = ms.Engine(p)
    ring = ms.OutRing(len(refs[0]), slots=3)
    got, errors = [], []

    def consumer():
        try:
            for _ in range(calls):
                got.append(ring.pop(timeout_ms=30000))
                time.sleep(0.001)  # a slower writer than the producer
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    t = threading.Thread(target=consumer)
    t.start()
    seqs = []
    for c in range(calls):
        eng.filter_u16(dev[c * n:(c + 1) * n])
        seqs.append(ring.push(eng, fmt, timeout_ms=30000))
    t.join(timeout=120)
    assert not t.is_alive() and not errors, errors
    assert seqs == list(range(calls))
    assert [s for s, _ in got] == list(range(calls))
    for c, (_, data) in enumerate(got):
        assert data == refs[c], c
    ring.close()
    eng.close()


def test_push_waits_for_a_free_slot(device):
    p, n, words, dev = _setup(device, 4, 42)
    eng = ms.Engine(p)
    ring = ms.OutRing(p.num_dms * p.count * p.M * 4, slots=2)
    for c in range(2):
        eng.filter_u16(dev[c * n:(c + 1) * n])
        ring.push(eng)
    eng.filter_u16(dev[2 * n:3 * n])
    t0 = time.time()
    with pytest.raises(ms.MsosmError):
        ring.push(eng, timeout_ms=100)
    assert time.time() - t0 >= 0.09
    seq, _ = ring.pop(timeout_ms=1000)
    assert seq == 0
    assert ring.push(eng, timeout_ms=1000) == 2
    assert [ring.pop(1000)[0], ring.pop(1000)[0]] == [1, 2]
    with pytest.raises(ms.MsosmError):
        ring.pop(timeout_ms=50)  # nothing queued
    ring.close()
    eng.close()
