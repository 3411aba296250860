"""The dual-polarisation PSRDADA ingest, host side: the header parser and
read bookkeeping of the C++ reader (include/msosm/psrdada.h, through the
msosm_psrdada program's --header mode) and the Python writer/reader round
trip.  No GPU calls."""
import os
import subprocess

import numpy as np
import pytest

from msosm_gpu_amd import psrdada

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "msosm_gpu_amd", "bin", "msosm_psrdada")


def _pols(L, seed):
    rng = np.random.default_rng(seed)
    return [rng.integers(1, 65536, 2 * L).astype(np.uint16) for _ in range(2)]


def test_write_read_round_trip(tmp_path):
    p1, p2 = _pols(3 * 2048, 1)
    f = tmp_path / "rec.dada"
    psrdada.write_dada(f, p1, p2, 16.0, 1400.0, extra={"SOURCE": "J0000+0000"})
    assert os.path.getsize(f) == 4096 + 2 * p1.nbytes
    h, q1, q2 = psrdada.read_dada(f)
    np.testing.assert_array_equal(q1, p1)
    np.testing.assert_array_equal(q2, p2)
    assert h["BW"] == "16.0" and h["SOURCE"] == "J0000+0000"
    # frames alternate: the second 8 KiB of data is pol2's first frame
    raw = np.fromfile(f, np.uint16, offset=4096)
    np.testing.assert_array_equal(raw[2 * 2048:4 * 2048], p2[:2 * 2048])


@pytest.mark.skipif(not os.path.exists(APP), reason="msosm_psrdada not built")
def test_cpp_header_parser(tmp_path):
    p1, p2 = _pols(2048, 2)
    f = tmp_path / "obs.dada"
    # keys in any order, padded values, unknown keys, a band inverted width
    psrdada.write_dada(f, p1, p2, -64.0, 1368.5, extra={"TELESCOPE": "FAST", "NPOL": "2"})
    r = subprocess.run([APP, "--file", str(f), "--header"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    keys = {"BW", "FREQ", "NBIT", "NCHAN", "NDIM", "NPOL", "filesize"}
    vals = dict(line.split(" ", 1) for line in r.stdout.splitlines() if line.split(" ", 1)[0] in keys)
    assert float(vals["BW"]) == -64.0 and float(vals["FREQ"]) == 1368.5
    assert (vals["NBIT"], vals["NCHAN"], vals["NDIM"], vals["NPOL"]) == ("16", "1", "2", "2")
    assert int(vals["filesize"]) == 4096 + 2 * p1.nbytes


@pytest.mark.skipif(not os.path.exists(APP), reason="msosm_psrdada not built")
def test_cpp_reader_rejects_short_file(tmp_path):
    f = tmp_path / "short.dada"
    f.write_bytes(b"BW 16\n")
    r = subprocess.run([APP, "--file", str(f), "--header"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "shorter" in r.stdout
