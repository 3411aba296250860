"""GPU fold against the sequential oracle fold: folded_data_raw and the
time bins must be bit for bit equal (same float additions, same order)."""
import numpy as np
import pytest

import msosm_gpu_amd as ms
import oracle

pytestmark = pytest.mark.gpu


def _load_intensity(fold, arr, device):
    fold.load_intensity(np.ascontiguousarray(arr, np.float32))  # host copy, synchronous


def test_fold_matches_oracle_bit_exact(device, golden):
    arrays, meta = golden
    for j, c in enumerate(meta["fold"]):
        f = ms.Fold(c["period"], c["fs"], c["size"], c["time_bin"])
        f.discard_samples(c["nd"])
        o = oracle.Fold(c["period"], c["fs"], c["size"], c["time_bin"])
        o.discard_samples(c["nd"])
        seq = arrays[f"fold_in_{j}"]
        print(f"fold case {j} P={c['P']} calls={c['calls']}", flush=True)
        for k in range(c["calls"]):
            chunk = seq[k * c["size"]:(k + 1) * c["size"]]
            _load_intensity(f, chunk, device)
            f.fold_data()
            o.fold_data(chunk)
        f.bins()
        o.bins()
        folded, raw, _ = f.get()
        np.testing.assert_array_equal(raw, o.raw())
        np.testing.assert_array_equal(raw, arrays[f"fold_raw_{j}"])
        np.testing.assert_array_equal(folded, o.folded())
        assert f.ready == o.ready
        f.close()


def test_intensity_then_fold_end_to_end(device):
    # two polarisations through two engines, intensity and fold on the GPU,
    # against the same chain on the CPU (msosm_psrdada.cpp:134-141)
    import torch
    bw, dm, f0 = 16e6, 10.0, 1392e6
    count, M = 16, 512
    rng = np.random.Generator(np.random.Philox(42))
    L = 8 * count * M
    u1 = np.clip(np.round(32768 + 700 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    u2 = np.clip(np.round(32768 + 700 * rng.standard_normal(2 * L)), 1, 65535).astype(np.uint16)
    p = ms.Plan(bw, dm, f0, fftpoint=1024, count=count)
    e1, e2 = ms.Engine(p), ms.Engine(p)
    fold = ms.Fold(0.005, bw, count * M, 256)
    fold.discard_samples(p.dm_info(0)["Nd"])
    ofold = oracle.Fold(0.005, bw, count * M, 256)
    ofold.discard_samples(p.dm_info(0)["Nd"])
    n = count * M
    for c in range(L // n):
        e1.filter_u16(u1.view(np.uint32)[c * n:(c + 1) * n])
        e2.filter_u16(u2.view(np.uint32)[c * n:(c + 1) * n])
        e1.synchronize()
        e2.synchronize()
        fold.intensity(e1.output_ptr(), e2.output_ptr())
        fold.fold_data()
        # CPU chain on the engines' own outputs
        a, b = e1.output_c64()[0, :, :].reshape(-1), e2.output_c64()[0, :, :].reshape(-1)
        inten = oracle.intensity(a, b)
        ofold.fold_data(inten)
        _, _, gi = fold.get()
        np.testing.assert_array_equal(gi, inten)
    fold.bins()
    ofold.bins()
    folded, raw, _ = fold.get()
    np.testing.assert_array_equal(raw, ofold.raw())
    np.testing.assert_array_equal(folded, ofold.folded())
    e1.close()
    e2.close()
    fold.close()


@pytest.mark.parametrize("P", [1234.567, 987654.321])
@pytest.mark.parametrize("n,stride,first", [(100000, 100352, 0), (99999, 100352, 0), (99999, 100001, 0),
                                            (99998, 100352, 1)])
def test_phase_fold_lds_histogram(device, P, n, stride, first):
    # FP64 phase per sample, LDS histogram per workgroup: bin assignment
    # (hit counts) exact, sums to float tolerance, for 3 trials at once; the
    # long period puts whole waves into one bin (wave-sum path), the short
    # one splits most waves over several bins.  Rows 16-byte aligned (two
    # samples per lane and load, also with an odd length), an odd stride and
    # rows starting at an odd sample (one sample per lane and load).
    import torch
    rng = np.random.default_rng(9)
    ndm, nbins, n0 = 3, 64, 5000
    a = (rng.standard_normal((ndm, stride)) + 1j * rng.standard_normal((ndm, stride))).astype(np.complex64)
    b = (rng.standard_normal((ndm, stride)) + 1j * rng.standard_normal((ndm, stride))).astype(np.complex64)
    ad = torch.from_numpy(a).to(device).reshape(-1)[first:]
    bd = torch.from_numpy(b).to(device).reshape(-1)[first:]
    a = np.stack([a.reshape(-1)[first + d * stride:first + d * stride + n] for d in range(ndm)])
    b = np.stack([b.reshape(-1)[first + d * stride:first + d * stride + n] for d in range(ndm)])
    for power in (True, False):
        prof = torch.zeros((ndm, nbins), dtype=torch.float32, device=device)
        hits = torch.zeros(nbins, dtype=torch.float32, device=device)
        ms.fold_phase(ad, n, P, nbins, prof, hits, b=bd, ndm=ndm, stride=stride, n0=n0, power=power)
        torch.cuda.synchronize()
        ph = (np.arange(n, dtype=np.float64) + n0) * (1.0 / P)
        ph -= np.floor(ph)
        bins = np.minimum((ph * nbins).astype(np.int64), nbins - 1)
        np.testing.assert_array_equal(hits.cpu().numpy(), np.bincount(bins, minlength=nbins).astype(np.float32))
        v = np.abs(a[:, :n].astype(np.complex128)) ** 2 + np.abs(b[:, :n].astype(np.complex128)) ** 2
        if not power:
            v = np.sqrt(v)
        exp = np.stack([np.bincount(bins, weights=v[d], minlength=nbins) for d in range(ndm)])
        assert np.max(np.abs(prof.cpu().numpy() - exp) / np.maximum(exp, 1e-30)) < 1e-5


def test_phase_fold_max_bins(device):
    # 8192 bins (the most one launch takes): the histogram and hit counts
    # plus the per-wave staging areas exceed 64 KiB of LDS per workgroup
    import torch
    rng = np.random.default_rng(10)
    n, nbins, P = 300000, 8192, 250000.5
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    ad = torch.from_numpy(a).to(device)
    prof = torch.zeros((1, nbins), dtype=torch.float32, device=device)
    hits = torch.zeros(nbins, dtype=torch.float32, device=device)
    ms.fold_phase(ad, n, P, nbins, prof, hits)
    torch.cuda.synchronize()
    ph = np.arange(n, dtype=np.float64) * (1.0 / P)
    ph -= np.floor(ph)
    bins = np.minimum((ph * nbins).astype(np.int64), nbins - 1)
    np.testing.assert_array_equal(hits.cpu().numpy(), np.bincount(bins, minlength=nbins).astype(np.float32))
    exp = np.bincount(bins, weights=np.abs(a.astype(np.complex128)) ** 2, minlength=nbins)
    assert np.max(np.abs(prof.cpu().numpy()[0] - exp) / np.maximum(exp, 1e-30)) < 1e-5
