"""Python binding of libmsosm.so, the MI355X-native MS-OSM coherent
dedispersion engine (C ABI: include/msosm_c.h).

The library is built in-tree (msosm_gpu_amd/lib/libmsosm.so) by
msosm_gpu_amd.build.  There is no fallback: every call goes to the HIP
library, and a missing library raises.

Device buffers may be passed as torch CUDA tensors (data_ptr) or as raw
integer device addresses; host buffers as numpy arrays.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmsosm.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

OK = 0
ERR_ARG, ERR_HIP, ERR_STATE, ERR_UNSUPPORTED, ERR_ALLOC = 1, 2, 3, 4, 5
METHOD_MSOSM, METHOD_OSM = 0, 1
CHIRP_REF_F32W, CHIRP_EXACT_F64 = 0, 1
FORM_PLAIN_F32, FORM_FUSED_F32, FORM_FUSED_F16 = 0, 1, 2
FORM_HANKEL_F16, FORM_HANKEL_BF16, FORM_HANKEL_BF16X3 = 3, 4, 5
FORM_FUSED_F16_MFMA = 6
STAGES = ("unpack", "fft_fwd", "multiply", "fft_inv", "discard", "copy", "h2d")

_lib = None


class MsosmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"msosm error {code}: {msg}")
        self.code = code


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("num_dms", ctypes.c_int),
        ("bw", ctypes.POINTER(ctypes.c_float)),
        ("dm", ctypes.POINTER(ctypes.c_float)),
        ("f0", ctypes.POINTER(ctypes.c_float)),
        ("fftpoint", ctypes.c_ulong),
        ("count", ctypes.c_int),
        ("input_size", ctypes.c_ulong),
        ("method", ctypes.c_int),
        ("chirp", ctypes.c_int),
        ("form", ctypes.c_int),
    ]


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("M", ctypes.c_ulong),
        ("M_min", ctypes.c_ulong),
        ("fftpoint", ctypes.c_ulong),
        ("count", ctypes.c_int),
        ("num_dms", ctypes.c_int),
        ("K_max", ctypes.c_int),
        ("R", ctypes.c_int),
        ("clamped", ctypes.c_int),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("inv_blocks_per_wg", ctypes.c_int),
        ("hankel_tiles_per_group", ctypes.c_int),
        ("scratch_chunk_kib", ctypes.c_int),
    ]


class AutotuneResult(ctypes.Structure):
    _fields_ = [
        ("fftpoint", ctypes.c_ulong),
        ("form", ctypes.c_int),
        ("tuning", Tuning),
        ("ns_per_dm_sample", ctypes.c_double),
        ("candidates", ctypes.c_int),
    ]


# Exported C ABI: name -> (restype, argtypes)
_vp, _i, _ul, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_ulong, ctypes.c_size_t
_f, _ll = ctypes.c_float, ctypes.c_longlong
SIGNATURES = {
    "msosm_plan_create": (_i, [ctypes.POINTER(PlanDesc), ctypes.POINTER(_vp)]),
    "msosm_plan_destroy": (None, [_vp]),
    "msosm_dm_model": (_i, [_f, _f, _f, _i, ctypes.POINTER(_ul), ctypes.POINTER(_ul)]),
    "msosm_plan_get_info": (_i, [_vp, ctypes.POINTER(PlanInfo)]),
    "msosm_plan_dm_info": (_i, [_vp, _i, ctypes.POINTER(_ul), ctypes.POINTER(_ul), ctypes.POINTER(_i),
                                ctypes.POINTER(_i)]),
    "msosm_plan_delay_points": (_i, [_vp, _i, _vp]),
    "msosm_engine_create": (_i, [_vp, _vp, ctypes.POINTER(_vp)]),
    "msosm_engine_destroy": (_i, [_vp]),
    "msosm_engine_filter_u16": (_i, [_vp, _vp, _i]),
    "msosm_engine_output_c64": (_i, [_vp, _vp, _i]),
    "msosm_engine_output_u16": (_i, [_vp, _vp, _i]),
    "msosm_engine_output_ptr": (_i, [_vp, ctypes.POINTER(_vp)]),
    "msosm_engine_dedisp_params": (_i, [_vp, _i, _vp]),
    "msosm_engine_synchronize": (_i, [_vp]),
    "msosm_engine_reset": (_i, [_vp]),
    "msosm_engine_stream": (_i, [_vp, ctypes.POINTER(_vp)]),
    "msosm_engine_state_size": (_i, [_vp, ctypes.POINTER(_sz)]),
    "msosm_engine_state_export": (_i, [_vp, _vp, _sz, _i]),
    "msosm_engine_state_import": (_i, [_vp, _vp, _sz, _i]),
    "msosm_engine_set_tuning": (_i, [_vp, ctypes.POINTER(Tuning)]),
    "msosm_engine_get_tuning": (_i, [_vp, ctypes.POINTER(Tuning)]),
    "msosm_range_push": (_i, [ctypes.c_char_p]),
    "msosm_range_pop": (_i, []),
    "msosm_engine_set_timing": (_i, [_vp, _i]),
    "msosm_engine_get_timing": (_i, [_vp, _vp, _vp, _i]),
    "msosm_k_uint16_to_complex": (_i, [_vp, _vp, _sz, _vp]),
    "msosm_k_complex_to_uint16": (_i, [_vp, _vp, _sz, _vp]),
    "msosm_k_intensity": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "msosm_k_multiply_batch": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp]),
    "msosm_k_multiply_delay_batch": (_i, [_vp, _vp, _i, _vp, _vp, _i, _i, _i, _i, _vp]),
    "msosm_k_discard_batch": (_i, [_vp, _vp, _i, _i, _i, _vp]),
    "msosm_k_fft_c2c": (_i, [_vp, _vp, _sz, _sz, _sz, _sz, _i, _vp]),
    "msosm_k_fft_c2c_noalloc": (_i, [_vp, _vp, _sz, _sz, _sz, _sz, _i, _vp]),
    "msosm_fft_plan_create": (_i, [_sz, _sz, _sz, _sz, ctypes.POINTER(_vp)]),
    "msosm_fft_plan_exec": (_i, [_vp, _vp, _vp, _i, _vp]),
    "msosm_fft_plan_destroy": (_i, [_vp]),
    "msosm_host_alloc": (_i, [_sz, ctypes.POINTER(_vp)]),
    "msosm_refsim_info": (_i, [_f, _f, _f, _f, ctypes.POINTER(_ul), ctypes.POINTER(_ul), ctypes.POINTER(_ul),
                               ctypes.POINTER(_ul)]),
    "msosm_refsim_generate": (_i, [_f, _f, _f, _f, _ul, _i, _f, _vp, _vp, ctypes.POINTER(_f)]),
    "msosm_refsim_awgn": (_i, [_f, _ul, _vp]),
    "msosm_outring_create": (_i, [_sz, _i, ctypes.POINTER(_vp)]),
    "msosm_outring_destroy": (_i, [_vp]),
    "msosm_outring_push": (_i, [_vp, _vp, _i, _i, ctypes.POINTER(_ll)]),
    "msosm_outring_pop": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_sz), ctypes.POINTER(_ll)]),
    "msosm_outring_release": (_i, [_vp]),
    "msosm_host_free": (_i, [_vp]),
    "msosm_k_hankel_delay_gemm": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _ll, _i, _i, _vp]),
    "msosm_k_inverse_discard": (_i, [_vp, _vp, _i, _i, _sz, _f, _vp]),
    "msosm_fold_create": (_i, [_f, _f, _ul, _ul, _vp, ctypes.POINTER(_vp)]),
    "msosm_fold_destroy": (_i, [_vp]),
    "msosm_fold_discard_samples": (_i, [_vp, _ul]),
    "msosm_fold_intensity": (_i, [_vp, _vp, _vp]),
    "msosm_fold_fold_data": (_i, [_vp]),
    "msosm_fold_bins": (_i, [_vp]),
    "msosm_fold_get": (_i, [_vp, _vp, _vp, _vp]),
    "msosm_fold_info": (_i, [_vp, ctypes.POINTER(_ul), ctypes.POINTER(_ul), ctypes.POINTER(_i)]),
    "msosm_fold_intensity_ptr": (_i, [_vp, ctypes.POINTER(_vp)]),
    "msosm_fold_load_intensity": (_i, [_vp, _vp, _i]),
    "msosm_fold_phase": (_i, [_vp, _vp, _sz, _i, _sz, _ll, ctypes.c_double, _i, _i, _vp, _vp, _vp]),
    "msosm_batch_create": (_i, [ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp), ctypes.POINTER(_ul),
                                ctypes.POINTER(ctypes.c_double), _i, _vp, ctypes.POINTER(_vp)]),
    "msosm_batch_launch": (_i, [_vp]),
    "msosm_batch_synchronize": (_i, [_vp]),
    "msosm_batch_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_sz)]),
    "msosm_batch_profiles": (_i, [_vp, _i, _vp, _vp]),
    "msosm_batch_destroy": (_i, [_vp]),
    "msosm_autotune": (_i, [ctypes.POINTER(PlanDesc), _ul, _ul, _ul, _i, _vp, ctypes.POINTER(AutotuneResult)]),
    "msosm_autotune_fftpoint": (_i, [ctypes.POINTER(PlanDesc), _ul, _ul, _ul, _i, _vp, ctypes.POINTER(_ul),
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_i)]),
    "msosm_sim_pulsar": (_i, [_vp, _sz, ctypes.c_double, ctypes.c_double, _f, ctypes.c_uint, _vp]),
    "msosm_sim_disperse": (_i, [_vp, _sz, ctypes.c_double, ctypes.c_double, ctypes.c_double, _i, _vp]),
    "msosm_sim_quantize": (_i, [_vp, _sz, _f, ctypes.c_uint, _f, _vp, _vp]),
    "msosm_spectrogram_u16": (_i, [_vp, _sz, _i, _i, _vp, _vp]),
    "msosm_rows_standardize": (_i, [_vp, _i, _i, _vp]),
    "msosm_incoherent_dedisperse": (_i, [_vp, _i, _i, _vp, _i, _i, _vp, _vp]),
    "msosm_rows_harmonics": (_i, [_vp, _i, _i, _i, _vp, _vp]),
    "msosm_incoherent_fold": (_i, [_vp, _i, _i, _vp, _i, ctypes.c_double, _vp, _vp]),
    "msosm_rows_fold": (_i, [_vp, _i, _i, ctypes.c_double, _i, _f, _vp, _vp]),
    "msosm_conv1d_fwd": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "msosm_conv1d_bwd": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "msosm_pool_fwd": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "msosm_pool_bwd": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "msosm_linear_fwd": (_i, [_vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "msosm_linear_bwd": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "msosm_softmax_xent": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "msosm_adam": (_i, [_vp, _vp, _vp, _vp, _vp, _sz, _f, _f, _f, _f, _i, _vp]),
    "msosm_f32_to_f16": (_i, [_vp, _vp, _sz, _vp]),
    "msosm_last_error": (ctypes.c_char_p, []),
    "msosm_version": (ctypes.c_char_p, []),
}


def lib():
    """Load libmsosm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m msosm_gpu_amd.build` "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise MsosmError(rc, lib().msosm_last_error().decode())


def _ptr(x):
    """Device/host address of a torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"cannot take the address of {type(x)}")


def _is_device(x):
    return hasattr(x, "is_cuda") and x.is_cuda


def _stream_handle(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    raise TypeError("stream must be None, an int handle or a torch.cuda.Stream")


def _farr(v, n):
    a = np.broadcast_to(np.asarray(v, np.float32), (n,)).copy()
    return a


class Plan:
    """Planner (host only): MS-OSM / OSM plan for num_dms DM trials."""

    def __init__(self, bw, dm, f0, fftpoint=0, count=1, input_size=0, method=METHOD_MSOSM, chirp=CHIRP_REF_F32W,
                 form=FORM_PLAIN_F32):
        n = max(np.size(bw), np.size(dm), np.size(f0))
        self._bw, self._dm, self._f0 = _farr(bw, n), _farr(dm, n), _farr(f0, n)
        fp = ctypes.POINTER(ctypes.c_float)
        d = PlanDesc(n, self._bw.ctypes.data_as(fp), self._dm.ctypes.data_as(fp), self._f0.ctypes.data_as(fp),
                     fftpoint, count, input_size, method, chirp, form)
        h = ctypes.c_void_p()
        check(lib().msosm_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        info = PlanInfo()
        check(lib().msosm_plan_get_info(self.h, ctypes.byref(info)))
        self.M, self.M_min, self.fftpoint = int(info.M), int(info.M_min), int(info.fftpoint)
        self.count, self.num_dms, self.K_max = int(info.count), int(info.num_dms), int(info.K_max)
        self.R, self.clamped = int(info.R), bool(info.clamped)
        self.method, self.chirp, self.form = method, chirp, form
        self.bw, self.dm, self.f0 = self._bw, self._dm, self._f0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_plan_destroy(self.h)
            self.h = None

    def dm_info(self, i):
        nd, mm = ctypes.c_ulong(), ctypes.c_ulong()
        K, dbs = ctypes.c_int(), ctypes.c_int()
        check(lib().msosm_plan_dm_info(self.h, i, ctypes.byref(nd), ctypes.byref(mm), ctypes.byref(K),
                                       ctypes.byref(dbs)))
        return dict(Nd=int(nd.value), M_min=int(mm.value), K=int(K.value), delay_block_size=int(dbs.value))

    def delay_points(self, i):
        out = np.zeros(self.fftpoint, np.int32)
        check(lib().msosm_plan_delay_points(self.h, i, out.ctypes.data))
        return out


class Engine:
    """GPU engine: filter count*M samples per call for all DM trials."""

    def __init__(self, plan, stream=None):
        self.plan = plan
        h = ctypes.c_void_p()
        check(lib().msosm_engine_create(plan.h, _stream_handle(stream), ctypes.byref(h)))
        self.h = h
        s = ctypes.c_void_p()
        check(lib().msosm_engine_stream(self.h, ctypes.byref(s)))
        self.stream = s.value

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def out_shape(self):
        p = self.plan
        return (p.num_dms, p.count, p.M)

    def filter_u16(self, x):
        """x: count*M packed uint16 pairs: numpy (host) or torch CUDA tensor."""
        n = self.plan.count * self.plan.M
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x)
            if x.nbytes != 4 * n:
                raise ValueError(f"input must hold {n} uint16 pairs ({4 * n} bytes), got {x.nbytes}")
            check(lib().msosm_engine_filter_u16(self.h, x.ctypes.data, 0))
        else:
            if x.numel() * x.element_size() != 4 * n:
                raise ValueError(f"input must hold {n} uint16 pairs")
            check(lib().msosm_engine_filter_u16(self.h, _ptr(x), 1 if _is_device(x) else 0))

    def output_c64(self, out=None):
        if out is None:
            out = np.empty(self.out_shape, np.complex64)
            check(lib().msosm_engine_output_c64(self.h, out.ctypes.data, 0))
            return out
        check(lib().msosm_engine_output_c64(self.h, _ptr(out), 1 if _is_device(out) else 0))
        return out

    def output_u16(self, out=None):
        if out is None:
            out = np.empty(self.out_shape + (2,), np.uint16)
            check(lib().msosm_engine_output_u16(self.h, out.ctypes.data, 0))
            return out
        check(lib().msosm_engine_output_u16(self.h, _ptr(out), 1 if _is_device(out) else 0))
        return out

    def output_ptr(self):
        p = ctypes.c_void_p()
        check(lib().msosm_engine_output_ptr(self.h, ctypes.byref(p)))
        return p.value

    def dedisp_params(self, i):
        out = np.empty(self.plan.fftpoint, np.complex64)
        check(lib().msosm_engine_dedisp_params(self.h, i, out.ctypes.data))
        return out

    def synchronize(self):
        check(lib().msosm_engine_synchronize(self.h))

    def reset(self):
        check(lib().msosm_engine_reset(self.h))

    def state_size(self):
        n = ctypes.c_size_t()
        check(lib().msosm_engine_state_size(self.h, ctypes.byref(n)))
        return int(n.value)

    def state_export(self, out=None):
        """Stream state image (overlap tail, delay ring, cursor): a numpy
        uint8 array, or into a given device tensor / numpy buffer."""
        n = self.state_size()
        if out is None:
            out = np.empty(n, np.uint8)
        nbytes = out.nbytes if isinstance(out, np.ndarray) else out.numel() * out.element_size()
        check(lib().msosm_engine_state_export(self.h, _ptr(out), nbytes, 1 if _is_device(out) else 0))
        return out

    def state_import(self, image):
        nbytes = image.nbytes if isinstance(image, np.ndarray) else image.numel() * image.element_size()
        check(lib().msosm_engine_state_import(self.h, _ptr(image), nbytes, 1 if _is_device(image) else 0))

    def set_tuning(self, inv_blocks_per_wg=0, hankel_tiles_per_group=0, scratch_chunk_kib=0):
        t = Tuning(inv_blocks_per_wg, hankel_tiles_per_group, scratch_chunk_kib)
        check(lib().msosm_engine_set_tuning(self.h, ctypes.byref(t)))

    def get_tuning(self):
        t = Tuning()
        check(lib().msosm_engine_get_tuning(self.h, ctypes.byref(t)))
        return dict(inv_blocks_per_wg=int(t.inv_blocks_per_wg), hankel_tiles_per_group=int(t.hankel_tiles_per_group),
                    scratch_chunk_kib=int(t.scratch_chunk_kib))

    def set_timing(self, on=True):
        check(lib().msosm_engine_set_timing(self.h, 1 if on else 0))

    def get_timing(self):
        ms = np.zeros(len(STAGES), np.float64)
        calls = np.zeros(len(STAGES), np.int64)
        check(lib().msosm_engine_get_timing(self.h, ms.ctypes.data, calls.ctypes.data, len(STAGES)))
        return {name: (float(ms[i]), int(calls[i])) for i, name in enumerate(STAGES)}


class Fold:
    """GPU fold with the reference's Fold_GPU bin assignment."""

    def __init__(self, period, fs, size, time_bin=-1, stream=None):
        h = ctypes.c_void_p()
        check(lib().msosm_fold_create(period, fs, size, ctypes.c_ulong(time_bin).value, _stream_handle(stream),
                                      ctypes.byref(h)))
        self.h = h
        self.size = size
        P, T, r = ctypes.c_ulong(), ctypes.c_ulong(), ctypes.c_int()
        check(lib().msosm_fold_info(self.h, ctypes.byref(P), ctypes.byref(T), ctypes.byref(r)))
        self.period_samples, self.time_bin = int(P.value), int(T.value)

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_fold_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def discard_samples(self, nd):
        check(lib().msosm_fold_discard_samples(self.h, nd))

    def intensity(self, pol1, pol2=None):
        check(lib().msosm_fold_intensity(self.h, _ptr(pol1), _ptr(pol2)))

    def intensity_ptr(self):
        p = ctypes.c_void_p()
        check(lib().msosm_fold_intensity_ptr(self.h, ctypes.byref(p)))
        return p.value

    def load_intensity(self, x):
        """x: size float32 values, numpy (host) or torch CUDA tensor."""
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, np.float32)
            assert x.size == self.size
            check(lib().msosm_fold_load_intensity(self.h, x.ctypes.data, 0))
        else:
            assert x.numel() == self.size
            check(lib().msosm_fold_load_intensity(self.h, _ptr(x), 1 if _is_device(x) else 0))

    def fold_data(self):
        check(lib().msosm_fold_fold_data(self.h))

    def bins(self):
        check(lib().msosm_fold_bins(self.h))

    @property
    def ready(self):
        r = ctypes.c_int()
        check(lib().msosm_fold_info(self.h, None, None, ctypes.byref(r)))
        return bool(r.value)

    def get(self):
        folded = np.empty(self.time_bin, np.float32)
        raw = np.empty(self.period_samples, np.float32)
        inten = np.empty(self.size, np.float32)
        check(lib().msosm_fold_get(self.h, folded.ctypes.data, raw.ctypes.data, inten.ctypes.data))
        return folded, raw, inten


# ---- element kernels on device buffers (torch tensors or addresses) -----------
def uint16_to_complex(dst, src, n, stream=None):
    check(lib().msosm_k_uint16_to_complex(_ptr(dst), _ptr(src), n, _stream_handle(stream)))


def complex_to_uint16(dst, src, n, stream=None):
    check(lib().msosm_k_complex_to_uint16(_ptr(dst), _ptr(src), n, _stream_handle(stream)))


def intensity(a, b, out, n, stream=None):
    check(lib().msosm_k_intensity(_ptr(a), _ptr(b), _ptr(out), n, _stream_handle(stream)))


def multiply_batch(a, b, N, out, count, num_dms=1, stream=None):
    check(lib().msosm_k_multiply_batch(_ptr(a), _ptr(b), N, _ptr(out), count, num_dms, _stream_handle(stream)))


def multiply_delay_batch(a, b, N, ring, delay, delay_block_size, index, count, num_dms=1, stream=None):
    check(lib().msosm_k_multiply_delay_batch(_ptr(a), _ptr(b), N, _ptr(ring), _ptr(delay), delay_block_size, index,
                                             count, num_dms, _stream_handle(stream)))


def discard_batch(a, b, M, count, num_dms=1, stream=None):
    check(lib().msosm_k_discard_batch(_ptr(a), _ptr(b), M, count, num_dms, _stream_handle(stream)))


def hankel_delay_gemm(ring, H, delay, Y, form, log2n, count, num_dms, g0, rx, kmax, stream=None):
    """Per-bin complex GEMMs of the delay line on the matrix cores (device
    tensors; see msosm_k_hankel_delay_gemm)."""
    check(lib().msosm_k_hankel_delay_gemm(_ptr(ring), _ptr(H), _ptr(delay), _ptr(Y), form, log2n, count, num_dms,
                                          g0, rx, kmax, _stream_handle(stream)))


def inverse_discard(Y, out, y_f16, log2n, units, scale=1.0, stream=None):
    check(lib().msosm_k_inverse_discard(_ptr(Y), _ptr(out), 1 if y_f16 else 0, log2n, units, scale,
                                        _stream_handle(stream)))


def fft_c2c(inp, out, n, batch, idist=None, odist=None, direction=-1, stream=None):
    idist = n if idist is None else idist
    odist = n if odist is None else odist
    check(lib().msosm_k_fft_c2c(_ptr(inp), _ptr(out), n, batch, idist, odist, direction, _stream_handle(stream)))


class FftPlan:
    """Batched C2C FFT plan owning its two-pass scratch (msosm_fft_plan_*)."""

    def __init__(self, n, batch=1, idist=None, odist=None):
        h = ctypes.c_void_p()
        check(lib().msosm_fft_plan_create(n, batch, n if idist is None else idist, n if odist is None else odist,
                                          ctypes.byref(h)))
        self.h = h
        self.n, self.batch = n, batch

    def __call__(self, inp, out, direction=-1, stream=None):
        check(lib().msosm_fft_plan_exec(self.h, _ptr(inp), _ptr(out), direction, _stream_handle(stream)))

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_fft_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def refsim_info(bw, dm, f0, period):
    """(Np, Nd, range, fftpoint) of the reference drivers' simulated signal."""
    v = [ctypes.c_ulong() for _ in range(4)]
    check(lib().msosm_refsim_info(bw, dm, f0, period, *[ctypes.byref(x) for x in v]))
    return tuple(int(x.value) for x in v)


def refsim_generate(bw, dm, f0, period, repeat=1, add_noise=False, snr_db=0.0):
    """(complex64 signal [repeat*Np], uint16 pairs [2*repeat*Np], signal power dB)."""
    Np = refsim_info(bw, dm, f0, period)[0]
    sig = np.empty(repeat * Np, np.complex64)
    u16 = np.empty(2 * repeat * Np, np.uint16)
    pw = ctypes.c_float()
    check(lib().msosm_refsim_generate(bw, dm, f0, period, repeat, 1 if add_noise else 0, snr_db, sig.ctypes.data,
                                      u16.ctypes.data, ctypes.byref(pw)))
    return sig, u16, float(pw.value)


def refsim_awgn(snr_db, n):
    out = np.empty(n, np.complex64)
    check(lib().msosm_refsim_awgn(snr_db, n, out.ctypes.data))
    return out


class OutRing:
    """Ring of page-locked output slots between an engine (producer) and a
    consumer thread; see msosm_outring_* in msosm_c.h."""

    COMPLEX64, UINT16 = 0, 1

    def __init__(self, slot_bytes, slots=3):
        h = ctypes.c_void_p()
        check(lib().msosm_outring_create(slot_bytes, slots, ctypes.byref(h)))
        self.h = h

    def push(self, engine, fmt=UINT16, timeout_ms=-1):
        seq = ctypes.c_longlong()
        check(lib().msosm_outring_push(self.h, engine.h, fmt, timeout_ms, ctypes.byref(seq)))
        return seq.value

    def pop(self, timeout_ms=-1):
        """(sequence number, bytes object copy of the slot)."""
        p, n, seq = ctypes.c_void_p(), ctypes.c_size_t(), ctypes.c_longlong()
        check(lib().msosm_outring_pop(self.h, timeout_ms, ctypes.byref(p), ctypes.byref(n), ctypes.byref(seq)))
        data = ctypes.string_at(p.value, n.value)
        check(lib().msosm_outring_release(self.h))
        return seq.value, data

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_outring_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Batch:
    """Several engines (independent input streams) and their phase folds,
    captured once into a hipGraph; launch() replays one full pass."""

    def __init__(self, engines, inputs, num_samples, periods=None, nbins=512, stream=None):
        n = len(engines)
        self.engines = list(engines)
        self.inputs = list(inputs)   # keep the device buffers alive
        eh = (ctypes.c_void_p * n)(*[e.h.value for e in engines])
        ih = (ctypes.c_void_p * n)(*[_ptr(x) for x in inputs])
        ns = (ctypes.c_ulong * n)(*num_samples)
        per = (ctypes.c_double * n)(*periods) if periods is not None else None
        self.nbins = nbins
        h = ctypes.c_void_p()
        check(lib().msosm_batch_create(eh, n, ih, ns, per, nbins, _stream_handle(stream), ctypes.byref(h)))
        self.h = h
        k, nodes = ctypes.c_int(), ctypes.c_size_t()
        check(lib().msosm_batch_info(self.h, ctypes.byref(k), ctypes.byref(nodes)))
        self.graph_nodes = int(nodes.value)

    def launch(self):
        check(lib().msosm_batch_launch(self.h))

    def synchronize(self):
        check(lib().msosm_batch_synchronize(self.h))

    def profiles(self, i):
        nd = self.engines[i].plan.num_dms
        prof = np.empty((nd, self.nbins), np.float32)
        hits = np.empty(self.nbins, np.float32)
        check(lib().msosm_batch_profiles(self.h, i, prof.ctypes.data, hits.ctypes.data))
        return prof, hits

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.msosm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def autotune_fftpoint(bw, dm, f0, min_fftpoint=1024, max_fftpoint=65536, target_samples=1 << 20, calls=3,
                      method=METHOD_MSOSM, chirp=CHIRP_REF_F32W, form=FORM_FUSED_F16, stream=None):
    """Measured best FFT size for these DM trials; returns (fftpoint, ns per
    DM-trial sample, form the size runs with)."""
    n = max(np.size(bw), np.size(dm), np.size(f0))
    bwa, dma, f0a = _farr(bw, n), _farr(dm, n), _farr(f0, n)
    fp = ctypes.POINTER(ctypes.c_float)
    d = PlanDesc(n, bwa.ctypes.data_as(fp), dma.ctypes.data_as(fp), f0a.ctypes.data_as(fp), 0, 1, 0, method,
                 chirp, form)
    best, ns, bf = ctypes.c_ulong(), ctypes.c_double(), ctypes.c_int()
    check(lib().msosm_autotune_fftpoint(ctypes.byref(d), min_fftpoint, max_fftpoint, target_samples, calls,
                                        _stream_handle(stream), ctypes.byref(best), ctypes.byref(ns),
                                        ctypes.byref(bf)))
    return int(best.value), float(ns.value), int(bf.value)


def autotune(bw, dm, f0, min_fftpoint=1024, max_fftpoint=65536, target_samples=1 << 20, calls=3,
             method=METHOD_MSOSM, chirp=CHIRP_REF_F32W, form=FORM_FUSED_F16, stream=None):
    """Measured best FFT size, form and launch shapes for these DM trials:
    dict(fftpoint, form, tuning (Engine.set_tuning keywords), ns_per_dm_sample,
    candidates)."""
    n = max(np.size(bw), np.size(dm), np.size(f0))
    bwa, dma, f0a = _farr(bw, n), _farr(dm, n), _farr(f0, n)
    fp = ctypes.POINTER(ctypes.c_float)
    d = PlanDesc(n, bwa.ctypes.data_as(fp), dma.ctypes.data_as(fp), f0a.ctypes.data_as(fp), 0, 1, 0, method,
                 chirp, form)
    r = AutotuneResult()
    check(lib().msosm_autotune(ctypes.byref(d), min_fftpoint, max_fftpoint, target_samples, calls,
                               _stream_handle(stream), ctypes.byref(r)))
    t = r.tuning
    return dict(fftpoint=int(r.fftpoint), form=int(r.form),
                tuning=dict(inv_blocks_per_wg=int(t.inv_blocks_per_wg),
                            hankel_tiles_per_group=int(t.hankel_tiles_per_group),
                            scratch_chunk_kib=int(t.scratch_chunk_kib)),
                ns_per_dm_sample=float(r.ns_per_dm_sample), candidates=int(r.candidates))


def fold_phase(a, n, period_samples, nbins, prof, hits=None, b=None, ndm=1, stride=None, n0=0, power=True,
               stream=None):
    """Accumulate the phase-folded profile of ndm rows of complex64 device
    data into prof[ndm, nbins] (and sample counts into hits[nbins])."""
    stride = n if stride is None else stride
    check(lib().msosm_fold_phase(_ptr(a), _ptr(b), n, ndm, stride, n0, float(period_samples), nbins,
                                 1 if power else 0, _ptr(prof), _ptr(hits), _stream_handle(stream)))


def dm_model(bw, dm, f0, method=METHOD_MSOSM):
    """(Nd, M_min) of one DM trial."""
    nd, mm = ctypes.c_ulong(), ctypes.c_ulong()
    check(lib().msosm_dm_model(bw, dm, f0, method, ctypes.byref(nd), ctypes.byref(mm)))
    return int(nd.value), int(mm.value)


def version():
    return lib().msosm_version().decode()
