"""ctypes view of the CPU oracle (oracle/msosm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker (or the timed CPU
baseline), never as a product code path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmsosm_oracle.so")

CHIRP_REF_F32W = 0
CHIRP_EXACT_F64 = 1

_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    ul, i32p, dp, fp = ctypes.c_ulong, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double), \
        ctypes.POINTER(ctypes.c_float)
    vp, sz, f, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_float, ctypes.c_int
    L.ora_nd.restype = ul
    L.ora_nd.argtypes = [f, f, f]
    L.ora_msosm_min_order.restype = ul
    L.ora_msosm_min_order.argtypes = [ul]
    L.ora_osm_min_order.restype = ul
    L.ora_osm_min_order.argtypes = [ul]
    L.ora_clamp_order.restype = ul
    L.ora_clamp_order.argtypes = [ul, ul, ctypes.POINTER(i)]
    L.ora_delaycount.restype = i
    L.ora_delaycount.argtypes = [ul, ul]
    L.ora_delay_points.restype = i
    L.ora_delay_points.argtypes = [f, f, f, ul, vp, ctypes.POINTER(i)]
    L.ora_segment_w.restype = i
    L.ora_segment_w.argtypes = [f, f, f, ul, vp, ctypes.POINTER(i)]
    L.ora_chirp.restype = None
    L.ora_chirp.argtypes = [f, f, f, ul, i, vp]
    L.ora_dedisp_params.restype = None
    L.ora_dedisp_params.argtypes = [f, f, f, ul, i, vp]
    L.ora_fft.restype = None
    L.ora_fft.argtypes = [vp, sz, i]
    L.ora_unpack_u16.argtypes = [vp, vp, sz]
    L.ora_pack_u16.argtypes = [vp, vp, sz]
    L.ora_intensity.argtypes = [vp, vp, vp, sz]
    L.ora_msosm_run.restype = i
    L.ora_msosm_run.argtypes = [vp, sz, ul, vp, vp, vp]
    L.ora_msosm_run_blocks.restype = i
    L.ora_msosm_run_blocks.argtypes = [vp, sz, ul, vp, vp, sz, sz, vp]
    L.ora_msosm_run_blocks_multi.restype = i
    L.ora_msosm_run_blocks_multi.argtypes = [vp, sz, ul, i, vp, vp, sz, sz, vp]
    L.ora_cpu_msosm_f32.restype = ctypes.c_double
    L.ora_cpu_msosm_f32.argtypes = [vp, sz, ul, i, vp, vp, vp, i]
    L.ora_fold_create.restype = vp
    L.ora_fold_create.argtypes = [f, f, ul, ul]
    L.ora_fold_destroy.argtypes = [vp]
    L.ora_fold_discard_samples.argtypes = [vp, ul]
    L.ora_fold_fold_data.argtypes = [vp, vp]
    L.ora_fold_bins.argtypes = [vp]
    L.ora_fold_period_samples.restype = ul
    L.ora_fold_period_samples.argtypes = [vp]
    L.ora_fold_time_bin.restype = ul
    L.ora_fold_time_bin.argtypes = [vp]
    L.ora_fold_raw.restype = fp
    L.ora_fold_raw.argtypes = [vp]
    L.ora_fold_folded.restype = fp
    L.ora_fold_folded.argtypes = [vp]
    L.ora_fold_trace.argtypes = [vp, vp]
    L.ora_fold_ready.restype = i
    L.ora_fold_ready.argtypes = [vp]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---- planner ---------------------------------------------------------------
def nd(bw, dm, f0):
    return int(lib().ora_nd(bw, dm, f0))


def msosm_min_order(nd_):
    return int(lib().ora_msosm_min_order(nd_))


def osm_min_order(nd_):
    return int(lib().ora_osm_min_order(nd_))


def clamp_order(fftpoint, m_min):
    c = ctypes.c_int(0)
    M = lib().ora_clamp_order(fftpoint, m_min, ctypes.byref(c))
    return int(M), bool(c.value)


def delaycount(nd_, M):
    return int(lib().ora_delaycount(nd_, M))


def delay_points(bw, dm, f0, M):
    out = np.zeros(2 * M, np.int32)
    holes = ctypes.c_int(0)
    K = lib().ora_delay_points(bw, dm, f0, M, _p(out), ctypes.byref(holes))
    return out, int(K), int(holes.value)


def segment_w(bw, dm, f0, M):
    K = delaycount(nd(bw, dm, f0), M)
    w = np.zeros(max(K, 1) + 1, np.float32)
    neg = ctypes.c_int(0)
    K = lib().ora_segment_w(bw, dm, f0, M, _p(w), ctypes.byref(neg))
    return w[:K + 1], int(neg.value)


def chirp(bw, dm, f0, M, mode=CHIRP_REF_F32W):
    c = np.zeros(2 * M, np.float64)
    lib().ora_chirp(bw, dm, f0, M, mode, _p(c))
    return c.view(np.complex128)


def dedisp_params(bw, dm, f0, M, mode=CHIRP_REF_F32W):
    H = np.zeros(4 * M, np.float64)
    lib().ora_dedisp_params(bw, dm, f0, M, mode, _p(H))
    return H.view(np.complex128)


def fft(x, sign=-1):
    a = np.ascontiguousarray(x, np.complex128).copy()
    lib().ora_fft(_p(a), a.size, sign)
    return a


# ---- element kernels ---------------------------------------------------------
def unpack_u16(u16pairs):
    u = np.ascontiguousarray(u16pairs, np.uint16).reshape(-1)
    out = np.zeros(u.size, np.float32)
    lib().ora_unpack_u16(_p(u), _p(out), u.size // 2)
    return out.view(np.complex64)


def pack_u16(c64):
    c = np.ascontiguousarray(c64, np.complex64)
    out = np.zeros(2 * c.size, np.uint16)
    lib().ora_pack_u16(_p(c), _p(out), c.size)
    return out


def intensity(a, b):
    a = np.ascontiguousarray(a, np.complex64)
    b = np.ascontiguousarray(b, np.complex64)
    out = np.zeros(a.size, np.float32)
    lib().ora_intensity(_p(a), _p(b), _p(out), a.size)
    return out


# ---- filters -----------------------------------------------------------------
def msosm_run(x, M, delay, H, b0=0, nb=None):
    x = np.ascontiguousarray(x, np.complex128)
    L = x.size
    if nb is None:
        nb = L // M - b0
    d = np.ascontiguousarray(delay, np.int32)
    Hc = np.ascontiguousarray(H, np.complex128)
    y = np.zeros(nb * M, np.complex128)
    rc = lib().ora_msosm_run_blocks(_p(x), L, M, _p(d), _p(Hc), b0, nb, _p(y))
    if rc != 0:
        raise ValueError("ora_msosm_run_blocks: bad sizes")
    return y


def msosm_run_multi(x, M, delays, H, b0=0, nb=None):
    """Output blocks b0 .. b0+nb-1 for every trial (rows of delays [ndm][2M]
    and H [ndm][2M]) of one stream x; returns [ndm][nb*M] complex128."""
    x = np.ascontiguousarray(x, np.complex128)
    L = x.size
    if nb is None:
        nb = L // M - b0
    d = np.ascontiguousarray(delays, np.int32)
    Hc = np.ascontiguousarray(H, np.complex128)
    ndm = d.shape[0]
    assert d.shape == (ndm, 2 * M) and Hc.shape == (ndm, 2 * M)
    y = np.zeros((ndm, nb * M), np.complex128)
    rc = lib().ora_msosm_run_blocks_multi(_p(x), L, M, ndm, _p(d), _p(Hc), b0, nb, _p(y))
    if rc != 0:
        raise ValueError("ora_msosm_run_blocks_multi: bad sizes")
    return y


def cpu_msosm_f32(u16, M, delays, H32, want_output=False, nthreads=0):
    """FP32 CPU baseline over all DMs; returns (checksum, output or None)."""
    u = np.ascontiguousarray(u16, np.uint16).reshape(-1)
    L = u.size // 2
    d = np.ascontiguousarray(delays, np.int32)
    ndm = d.shape[0]
    H = np.ascontiguousarray(H32, np.complex64)
    y = np.zeros((ndm, L), np.complex64) if want_output else None
    cs = lib().ora_cpu_msosm_f32(_p(u), L, M, ndm, _p(d), _p(H), _p(y) if y is not None else None, nthreads)
    return cs, y


class Fold:
    """Sequential fold with the reference's float state (fold_gpu.cpp)."""

    def __init__(self, period, fs, size, time_bin=-1):
        tb = ctypes.c_ulong(time_bin).value
        self.h = lib().ora_fold_create(period, fs, size, tb)
        self.size = size

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_fold_destroy(self.h)
            self.h = None

    def discard_samples(self, nd_):
        lib().ora_fold_discard_samples(self.h, nd_)

    def fold_data(self, intensity_):
        a = np.ascontiguousarray(intensity_, np.float32)
        assert a.size == self.size
        lib().ora_fold_fold_data(self.h, _p(a))

    def bins(self):
        lib().ora_fold_bins(self.h)

    @property
    def period_samples(self):
        return int(lib().ora_fold_period_samples(self.h))

    @property
    def time_bin(self):
        return int(lib().ora_fold_time_bin(self.h))

    @property
    def ready(self):
        return bool(lib().ora_fold_ready(self.h))

    def raw(self):
        P = self.period_samples
        return np.ctypeslib.as_array(lib().ora_fold_raw(self.h), (P,)).copy()

    def folded(self):
        T = self.time_bin
        return np.ctypeslib.as_array(lib().ora_fold_folded(self.h), (T,)).copy()

    def trace(self):
        out = np.zeros(self.size, np.int64)
        lib().ora_fold_trace(self.h, _p(out))
        return out
