"""Host-side planner of the product (libmsosm.so C ABI, no GPU needed)
against the golden plan table and the oracle."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle


def test_plan_table_matches_golden(golden):
    _, meta = golden
    for r in meta["plans"]:
        method = ms.METHOD_MSOSM if r["method"] == "msosm" else ms.METHOD_OSM
        if r["method"] == "osm" and r["M"] > (1 << 27):
            continue
        p = ms.Plan(r["bw"], r["dm"], r["f0"], fftpoint=r["fftpoint"], count=r["count"], method=method)
        info = p.dm_info(0)
        assert info["Nd"] == r["Nd"], r
        assert info["M_min"] == r["M_min"], r
        assert p.M == r["M"], r
        assert info["K"] == r["K"], r
        assert info["delay_block_size"] == r["R"], r
        assert p.R == r["R"] and int(p.clamped) == r["clamped"], r


def test_delay_points_match_golden_and_oracle(golden):
    arrays, meta = golden
    for i, c in enumerate(meta["segw"]):
        p = ms.Plan(c["bw"], c["dm"], c["f0"], fftpoint=2 * c["M"], count=1)
        assert p.M == c["M"]
        d = p.delay_points(0)
        np.testing.assert_array_equal(d, arrays[f"delay_{i}"].astype(np.int32))
        od, K, _ = oracle.delay_points(c["bw"], c["dm"], c["f0"], c["M"])
        np.testing.assert_array_equal(d, od)


def test_multi_dm_common_order():
    dms = np.linspace(0, 100, 256).astype(np.float32)
    p = ms.Plan(64e6, dms, 1368e6, fftpoint=16384, count=8)
    assert p.num_dms == 256 and p.M == 8192
    Ks = [p.dm_info(i)["K"] for i in range(256)]
    assert Ks[0] == 1 and max(Ks) == 152 and p.K_max == 152
    assert p.R == 160  # roundup(152 + 8 - 1, 8)
    for i in (0, 1, 57, 146, 255):
        od, K, _ = oracle.delay_points(64e6, float(dms[i]), 1368e6, 8192)
        np.testing.assert_array_equal(p.delay_points(i), od)
        assert K == Ks[i]
    # common M = max of per-DM minima when fftpoint = 0
    q = ms.Plan(64e6, dms, 1368e6, fftpoint=0, count=1)
    assert q.M == max(oracle.msosm_min_order(oracle.nd(64e6, float(d), 1368e6)) for d in dms)


def test_input_size_sets_count():
    p = ms.Plan(64e6, [10.0, 20.0], 1368e6, fftpoint=16384, count=1, input_size=1 << 20)
    assert p.count == (1 << 20) // 8192


def test_clamp(capfd):
    p = ms.Plan(16e6, 10.0, 1392e6, fftpoint=64, count=2)
    assert p.M == 256 and p.clamped
    assert "minimum FFT point 512" in capfd.readouterr().out


@pytest.mark.parametrize("kwargs,code", [
    (dict(bw=16e6, dm=-1.0, f0=1392e6), ms.MsosmError),
    (dict(bw=0.0, dm=1.0, f0=1392e6), ms.MsosmError),
    (dict(bw=16e6, dm=10.0, f0=1392e6, fftpoint=3000), ms.MsosmError),   # not a power of two
    (dict(bw=16e6, dm=0.0, f0=1392e6), ms.MsosmError),                   # M = 0
    (dict(bw=16e6, dm=10.0, f0=1392e6, count=0), ms.MsosmError),
    (dict(bw=16e6, dm=10.0, f0=1392e6, method=7), ms.MsosmError),
    (dict(bw=16e6, dm=10.0, f0=1392e6, form=9), ms.MsosmError),
])
def test_plan_errors(kwargs, code):
    with pytest.raises(code):
        ms.Plan(**kwargs)


def test_error_message_is_reported():
    with pytest.raises(ms.MsosmError) as e:
        ms.Plan(16e6, -5.0, 1392e6)
    assert "dm" in str(e.value)
