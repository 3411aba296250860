"""Engine stream state (checkpoint/resume through the C ABI), launch-shape
tuning and the batch graph's handling of engine cursors.

- Export after call k and import into a fresh engine of the same plan: the
  fresh engine continues bit for bit like the original (every form, host and
  device images); images of another plan are refused.
- Blocks per workgroup of the inverse and tiles per Hankel group change only
  the launch shape: outputs are bit-identical to the default rule's.
- Capturing a batch graph leaves an engine's eager stream where it was, and
  a launch moves it to the end of the replayed pass.
"""
import numpy as np
import pytest

import msosm_gpu_amd as ms
from test_gpu_engine import make_input, rel_l2

pytestmark = pytest.mark.gpu

BW, F0 = 64e6, 1368e6
DMS = [0.0, 3.0, 12.5, 29.0]


def _dev(u, device):
    import torch
    return torch.from_numpy(u.view(np.int32).copy()).to(device)


def _calls(eng, dev, first, last):
    n = eng.plan.count * eng.plan.M
    outs = []
    for c in range(first, last):
        eng.filter_u16(dev[c * n:(c + 1) * n])
        outs.append(eng.output_c64())
    return outs


@pytest.mark.parametrize("form", [ms.FORM_PLAIN_F32, ms.FORM_FUSED_F32, ms.FORM_FUSED_F16, ms.FORM_FUSED_F16_MFMA,
                                  ms.FORM_HANKEL_F16])
@pytest.mark.parametrize("where", ["host", "device"])
def test_export_import_continues_bit_for_bit(device, form, where):
    import torch
    p = ms.Plan(BW, DMS, F0, fftpoint=16384, count=3, form=form)
    calls = 7
    u = make_input(calls * p.count * p.M, 400 + form)
    dev = _dev(u, device)
    a = ms.Engine(p)
    _calls(a, dev, 0, 3)
    print(f"state form={form} where={where}: {a.state_size()} bytes", flush=True)
    if where == "host":
        img = a.state_export()
    else:
        img = torch.empty(a.state_size(), dtype=torch.uint8, device=device)
        a.state_export(img)
    rest_a = _calls(a, dev, 3, calls)
    b = ms.Engine(p)
    _calls(b, dev, 0, 1)        # some other history first: the import replaces it
    b.state_import(img)
    rest_b = _calls(b, dev, 3, calls)
    for x, y in zip(rest_a, rest_b):
        np.testing.assert_array_equal(x, y)
    # an image of another plan is refused: different trials with the same
    # buffer sizes, and a different ring length
    q = ms.Engine(ms.Plan(BW, DMS[:2] + [12.0, 29.0], F0, fftpoint=16384, count=3, form=form))
    assert q.state_size() == a.state_size()
    with pytest.raises(ms.MsosmError, match="different plan"):
        q.state_import(img)
    q.close()
    q = ms.Engine(ms.Plan(BW, DMS[:3] + [30.0], F0, fftpoint=16384, count=3, form=form))
    with pytest.raises(ms.MsosmError, match="different plan|state is"):
        q.state_import(img)
    for e in (a, b, q):
        e.close()


def test_import_into_a_reset_stream_equals_a_fresh_engine(device):
    # the image of a fresh engine is all history zero and cursor 0: importing
    # it is the same as a reset
    p = ms.Plan(BW, DMS, F0, fftpoint=8192, count=4, form=ms.FORM_FUSED_F16)
    u = make_input(3 * p.count * p.M, 9)
    dev = _dev(u, device)
    fresh = ms.Engine(p)
    img = fresh.state_export()
    e = ms.Engine(p)
    _calls(e, dev, 0, 2)
    e.state_import(img)
    x = _calls(e, dev, 0, 3)
    y = _calls(fresh, dev, 0, 3)
    for a, b in zip(x, y):
        np.testing.assert_array_equal(a, b)
    e.close()
    fresh.close()


@pytest.mark.parametrize("form", [ms.FORM_FUSED_F32, ms.FORM_FUSED_F16, ms.FORM_FUSED_F16_MFMA])
def test_inverse_blocks_per_workgroup_change_only_the_launch(device, form):
    dms = list(np.linspace(0, 40, 9))
    p = ms.Plan(BW, dms, F0, fftpoint=16384, count=12, form=form)
    u = make_input(2 * p.count * p.M, 31)
    dev = _dev(u, device)
    ref = None
    for nbpw in (0, 1, 3, 8, 32):
        e = ms.Engine(p)
        e.set_tuning(inv_blocks_per_wg=nbpw)
        assert e.get_tuning()["inv_blocks_per_wg"] == nbpw
        print(f"form={form} blocks per workgroup {nbpw}", flush=True)
        out = _calls(e, dev, 0, 2)
        e.close()
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("form", [ms.FORM_HANKEL_F16, ms.FORM_HANKEL_BF16X3])
def test_hankel_group_size_changes_only_the_launch(device, form):
    dms = list(np.linspace(0, 100, 70))          # 5 tiles of 16 trials
    p = ms.Plan(BW, dms, F0, fftpoint=16384, count=8, form=form)
    u = make_input(3 * p.count * p.M, 57)
    dev = _dev(u, device)
    ref = None
    seen = set()
    for tpg in (0, 1, 2, 5):
        e = ms.Engine(p)
        e.set_tuning(hankel_tiles_per_group=tpg)
        eff = e.get_tuning()["hankel_tiles_per_group"]
        seen.add(eff)
        print(f"form={form} tiles per group requested {tpg}, in effect {eff}", flush=True)
        out = _calls(e, dev, 0, 3)
        e.close()
        if ref is None:
            ref = out
        for a, b in zip(out, ref):
            assert rel_l2(a, b) == 0.0
    assert len(seen) >= 3


def test_batch_capture_keeps_the_eager_stream_in_place(device):
    import torch
    p = ms.Plan(16e6, [10.0, 25.0], 1392e6, fftpoint=2048, count=4, form=ms.FORM_FUSED_F16)
    n = p.count * p.M
    u = make_input(8 * n, 77)
    dev = _dev(u, device)
    ref = ms.Engine(p)
    expect = _calls(ref, dev, 0, 6)
    e = ms.Engine(p)
    got = _calls(e, dev, 0, 2)
    # capture a 3-call pass over other input; no launch yet
    other = _dev(make_input(3 * n, 78), device)
    b = ms.Batch([e], [other], [3 * n])
    got += _calls(e, dev, 2, 4)        # eager calls between capture and launch
    for x, y in zip(got, expect[:4]):
        np.testing.assert_array_equal(x, y)
    # after a launch the engine has run the 3-call pass: its next eager call is
    # call 3 of that stream
    b.launch()
    b.synchronize()
    solo = ms.Engine(p)
    want = _calls(solo, other, 0, 3)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(e.output_c64(), want[-1])
    b.close()
    for x in (ref, e, solo):
        x.close()


def test_batch_refuses_engines_with_timing_on(device):
    import torch
    p = ms.Plan(16e6, [10.0], 1392e6, fftpoint=2048, count=4, form=ms.FORM_FUSED_F16)
    e = ms.Engine(p)
    e.set_timing(True)
    x = torch.zeros(4 * p.M * p.count, dtype=torch.int32, device=device)
    with pytest.raises(ms.MsosmError, match="timing"):
        ms.Batch([e], [x], [p.count * p.M])
    e.set_timing(False)
    e.close()
