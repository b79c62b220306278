"""ctypes binding of libmdsm_amd.so (the C ABI declared in include/mdsm_abi.h).

The HIP kernels and the host logic live in the shared library; this module
only marshals numpy arrays and device pointers.  Loading fails loudly when
the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmdsm_amd.so")

(MDSM_OK, MDSM_ERR_ARG, MDSM_ERR_HIP, MDSM_ERR_NOMEM, MDSM_ERR_OVERFLOW, MDSM_ERR_STATE,
 MDSM_ERR_IO, MDSM_ERR_FORMAT) = range(8)
# F16_PAIRS: F16 rows; channels 2k, 2k+1 added in FP16 before the FP32 sum
F32, F16, U8, F16_PAIRS = 0, 1, 2, 3
DELAY_FP64_ROUND, DELAY_REFERENCE, DELAY_LEGACY = 0, 1, 2
_ITEMSIZE = {F32: 4, F16: 2, U8: 1, F16_PAIRS: 2}
_NP = {F32: np.float32, F16: np.uint16, U8: np.uint8, F16_PAIRS: np.uint16}
ALIGN = {F32: 1, F16: 2, U8: 4, F16_PAIRS: 2}  # required multiple of the input row length

CAND_DTYPE = np.dtype([("dm_idx", "<i4"), ("sample", "<i4"), ("width_log2", "<i4"), ("snr", "<f4")])


class MdsmError(RuntimeError):
    pass


class Survey(C.Structure):
    _fields_ = [("fch1", C.c_double), ("foff", C.c_double), ("tsamp", C.c_double),
                ("lowdm", C.c_double), ("dmstep", C.c_double), ("nchans", C.c_uint32),
                ("ndms", C.c_uint32), ("nsamp", C.c_uint32), ("delay_mode", C.c_int32)]

    @classmethod
    def make(cls, fch1, foff, nchans, tsamp, lowdm, dmstep, ndms, nsamp, delay_mode=DELAY_FP64_ROUND):
        return cls(fch1, foff, tsamp, lowdm, dmstep, nchans, ndms, nsamp, delay_mode)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _load():
    if not os.path.exists(LIB_PATH):
        raise MdsmError(f"{LIB_PATH} is missing: build it with `make -C mdsm_amd/csrc` "
                        "or __graft_entry__.build()")
    return C.CDLL(LIB_PATH)


_lib = _load()
_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_i = C.c_int
_f = C.c_float
_d = C.c_double
_P = C.POINTER


def _sig(name, *args):
    fn = getattr(_lib, name)
    fn.restype = C.c_int
    fn.argtypes = list(args)
    return fn


_lib.mdsm_last_error.restype = C.c_char_p
_lib.mdsm_last_error.argtypes = []
_sig("mdsm_version")
_sig("mdsm_device_count", _P(_i))
_sig("mdsm_set_device", _i)
_sig("mdsm_malloc", _P(_vp), C.c_size_t)
_sig("mdsm_free", _vp)
_sig("mdsm_host_malloc", _P(_vp), C.c_size_t)
_sig("mdsm_host_free", _vp)
_sig("mdsm_memcpy_htod", _vp, _vp, C.c_size_t, _vp)
_sig("mdsm_memcpy_dtoh", _vp, _vp, C.c_size_t, _vp)
_sig("mdsm_memcpy_dtod", _vp, _vp, C.c_size_t, _vp)
_sig("mdsm_memset", _vp, _i, C.c_size_t, _vp)
_sig("mdsm_stream_create", _P(_vp))
_sig("mdsm_stream_destroy", _vp)
_sig("mdsm_stream_synchronize", _vp)
_sig("mdsm_device_synchronize")
_sig("mdsm_event_create", _P(_vp))
_sig("mdsm_event_destroy", _vp)
_sig("mdsm_event_record", _vp, _vp)
_sig("mdsm_event_elapsed", _vp, _vp, _P(_f))
_sig("mdsm_delay_table", _P(Survey), _vp, _P(_u32))
_sig("mdsm_plan_create", _P(Survey), _i, _P(_vp))
_sig("mdsm_plan_destroy", _vp)
_sig("mdsm_plan_info", _vp, _P(_u32), _P(_u64))
_sig("mdsm_plan_device_table", _vp, _P(_vp))
_sig("mdsm_plan_set_variant", _vp, _i)
_sig("mdsm_plan_get_variant", _vp, _P(_i))
_sig("mdsm_tile_variants", _i, _P(_i))
_sig("mdsm_tile_preferred", _i, _P(_i))
_sig("mdsm_plan_variant_channels", _vp, _i, _P(_u32))
_sig("mdsm_plan_autotune", _vp, _vp, _u64, _vp, _vp, _P(_i), _P(_f))
_sig("mdsm_dedisperse", _vp, _vp, _u64, _vp, _vp)
_sig("mdsm_channel_stats", _vp, _u64, _u32, _u64, _vp, _vp, _vp)
_sig("mdsm_normalise", _vp, _u64, _u32, _u64, _vp, _vp, _vp, _i, _u64, _vp)
_sig("mdsm_block_stats", _vp, _u32, _u32, _u32, _vp, _vp, _vp)
_sig("mdsm_detect", _vp, _u32, _u32, _u32, _vp, _vp, _i, _d, _vp, _u32, _vp, _vp)
_sig("mdsm_detect_global", _vp, _u32, _u32, _u64, _u32, _f, _vp, _u32, _vp, _P(_d), _P(_d), _vp)
_sig("mdsm_threshold_global", _vp, _u32, _u32, _u64, _d, _d, _f, _vp, _u32, _vp, _vp)
_sig("mdsm_detect_global_noalloc", _vp, _u32, _u32, _u64, _u32, _f, _vp, _u32, _vp, _P(_d), _P(_d),
     _vp)
_sig("mdsm_detect_legacy", _vp, _u32, _u32, _f, _f, _f, _u64, _vp, _vp, _u32, _u32, _vp, _u32, _vp,
     _P(_f), _P(_f), _vp)
_sig("mdsm_median_filter5", _vp, _u32, _u32, _vp, _vp)
_sig("mdsm_detrend_normalise", _vp, _u64, _u32, _vp)
_sig("mdsm_sort_candidates", _vp, _u32)
_sig("mdsm_dbscan2d", _vp, _u32, _i, _i, _vp, _P(_u32))
_sig("mdsm_cluster_peaks", _vp, _vp, _u32, _u32, _vp, _vp)
_sig("mdsm_dbscan_ref", _vp, _vp, _vp, _u32, _f, _f, _f, _u32, _vp, _P(_u32))
_sig("mdsm_classify_cluster_ref", _vp, _vp, _vp, _u32, _f, _f, _u32, _f, _f, _f, _u32, _P(_f), _vp)
_sig("mdsm_batch_create", _P(Survey), _u32, _i, _P(_vp))
_sig("mdsm_batch_destroy", _vp)
_sig("mdsm_batch_info", _vp, _u32, _P(_u32), _P(_u64))
_sig("mdsm_batch_bind", _vp, _P(_vp), _P(_u64), _P(_vp))
_sig("mdsm_batch_launch", _vp, _vp)
_sig("mdsm_batch_dedisperse", _vp, _P(_vp), _P(_u64), _P(_vp), _vp)
_sig("mdsm_batch_set_variant", _vp, _i)
_sig("mdsm_batch_set_mode", _vp, _i)
_sig("mdsm_batch_autotune", _vp, _P(_vp), _P(_u64), _P(_vp), _vp, _P(_i), _P(_f))
BATCH_PERSISTENT, BATCH_FLAT = 0, 1
_sig("mdsm_generate", C.c_uint64, _u32, C.c_int64, _u64, _vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp,
     _u64, _vp)


def check(rc):
    if rc != MDSM_OK:
        msg = _lib.mdsm_last_error().decode(errors="replace")
        raise MdsmError(f"mdsm status {rc}: {msg}")


def lib():
    return _lib


def last_error() -> str:
    return _lib.mdsm_last_error().decode(errors="replace")


def device_count() -> int:
    n = _i(0)
    check(_lib.mdsm_device_count(C.byref(n)))
    return n.value


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


# ------------------------------------------------------------------ memory
class DeviceBuffer:
    """An owning HIP device allocation."""

    def __init__(self, nbytes: int):
        p = _vp()
        check(_lib.mdsm_malloc(C.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(_lib.mdsm_memcpy_htod(self.ptr + offset, _ptr(a), a.nbytes, None))

    def download(self, dtype, shape, offset: int = 0):
        out = np.empty(shape, dtype)
        assert offset + out.nbytes <= self.nbytes
        check(_lib.mdsm_memcpy_dtoh(_ptr(out), self.ptr + offset, out.nbytes, None))
        return out

    def zero(self, stream=None):
        check(_lib.mdsm_memset(self.ptr, 0, self.nbytes, stream))

    def free(self):
        if self.ptr:
            check(_lib.mdsm_free(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(_lib.mdsm_device_synchronize())


class Timer:
    """HIP-event timer on a stream (milliseconds)."""

    def __init__(self, stream=None):
        self.stream = stream
        a, b = _vp(), _vp()
        check(_lib.mdsm_event_create(C.byref(a)))
        check(_lib.mdsm_event_create(C.byref(b)))
        self.e0, self.e1 = a.value, b.value

    def start(self):
        check(_lib.mdsm_event_record(self.e0, self.stream))

    def stop(self) -> float:
        check(_lib.mdsm_event_record(self.e1, self.stream))
        ms = _f(0.0)
        check(_lib.mdsm_event_elapsed(self.e0, self.e1, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            _lib.mdsm_event_destroy(self.e0)
            _lib.mdsm_event_destroy(self.e1)
        except Exception:
            pass


# ------------------------------------------------------------- delay table
def delay_table(s: Survey):
    t = np.empty((s.nchans, s.ndms), np.int32)
    ms = _u32(0)
    check(_lib.mdsm_delay_table(C.byref(s), _ptr(t), C.byref(ms)))
    return t, ms.value


# -------------------------------------------------------------------- plan
class Plan:
    def __init__(self, s: Survey, dtype: int = F32):
        self.survey = s
        self.dtype = dtype
        p = _vp()
        check(_lib.mdsm_plan_create(C.byref(s), dtype, C.byref(p)))
        self.handle = p.value
        ms, ld = _u32(0), _u64(0)
        check(_lib.mdsm_plan_info(self.handle, C.byref(ms), C.byref(ld)))
        self.maxshift, self.min_ld = ms.value, ld.value

    def set_variant(self, v: int):
        check(_lib.mdsm_plan_set_variant(self.handle, v))

    @property
    def variant(self) -> int:
        v = _i(0)
        check(_lib.mdsm_plan_get_variant(self.handle, C.byref(v)))
        return v.value

    def variant_channels(self, v: int) -> int:
        """Channels tile shape v stages per LDS block here (0: cannot stage)."""
        n = _u32(0)
        check(_lib.mdsm_plan_variant_channels(self.handle, v, C.byref(n)))
        return n.value

    def dedisperse(self, d_in: int, ld: int, d_out: int, stream=None):
        check(_lib.mdsm_dedisperse(self.handle, d_in, ld, d_out, stream))

    def autotune(self, d_in: int, ld: int, d_out: int, stream=None):
        v, ms = _i(0), _f(0.0)
        check(_lib.mdsm_plan_autotune(self.handle, d_in, ld, d_out, stream, C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def close(self):
        if self.handle:
            check(_lib.mdsm_plan_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def preferred_variant(dtype=F32) -> int:
    v = _i(0)
    check(_lib.mdsm_tile_preferred(dtype, C.byref(v)))
    return v.value


def tile_variants(dtype=F32) -> int:
    n = _i(0)
    check(_lib.mdsm_tile_variants(dtype, C.byref(n)))
    return n.value


def dedisperse_host(s: Survey, x: np.ndarray, dtype: int = F32, variant: int = 0):
    """Upload x [nchans][ld], dedisperse on the GPU, return [ndms][nsamp] float32."""
    plan = Plan(s, dtype)
    if variant:
        plan.set_variant(variant)
    x = np.ascontiguousarray(x, dtype=_NP[dtype])
    assert x.shape[0] == s.nchans and x.shape[1] >= s.nsamp + plan.maxshift
    ld = max(x.shape[1], plan.min_ld)
    ld = -(-ld // ALIGN[dtype]) * ALIGN[dtype]
    if ld != x.shape[1]:
        xp = np.zeros((s.nchans, ld), x.dtype)
        xp[:, :x.shape[1]] = x
        x = xp
    din = DeviceBuffer.from_numpy(x)
    dout = DeviceBuffer(s.ndms * s.nsamp * 4)
    plan.dedisperse(din.ptr, ld, dout.ptr)
    synchronize()
    return dout.download(np.float32, (s.ndms, s.nsamp))


# ---------------------------------------------------- normalise / detect
def channel_stats(d_raw: int, ld: int, nchans: int, n: int, d_mean: int, d_std: int, stream=None):
    check(_lib.mdsm_channel_stats(d_raw, ld, nchans, n, d_mean, d_std, stream))


def normalise(d_raw, ld, nchans, n, d_mean, d_std, d_out, out_dtype, ldo, stream=None):
    check(_lib.mdsm_normalise(d_raw, ld, nchans, n, d_mean, d_std, d_out, out_dtype, ldo, stream))


def block_stats(d_x, nrows, nsamp, L, d_mean, d_std, stream=None):
    check(_lib.mdsm_block_stats(d_x, nrows, nsamp, L, d_mean, d_std, stream))


def detect(d_x, nrows, nsamp, L, d_mean, d_std, nwidths, thresh, d_cands, cap, d_count, stream=None):
    check(_lib.mdsm_detect(d_x, nrows, nsamp, L, d_mean, d_std, nwidths, thresh, d_cands, cap,
                           d_count, stream))


def detect_global(d_x, nrows, nsamp, thresh, d_cands, cap, d_count, stream=None, ld=None,
                  stats_nsamp=None):
    """Reference-semantics global detector; returns the (mean, rms) it used."""
    m, r = _d(0.0), _d(0.0)
    check(_lib.mdsm_detect_global(d_x, nrows, nsamp, nsamp if ld is None else ld,
                                  nsamp if stats_nsamp is None else stats_nsamp, thresh, d_cands,
                                  cap, d_count, C.byref(m), C.byref(r), stream))
    return m.value, r.value


def detect_global_noalloc(d_x, nrows, nsamp, thresh, d_cands, cap, d_count, stream=None, ld=None,
                          stats_nsamp=None):
    """mdsm_detect_global on the library's fixed device workspace (no
    allocation); returns the (mean, rms) it used."""
    m, r = _d(0.0), _d(0.0)
    check(_lib.mdsm_detect_global_noalloc(d_x, nrows, nsamp, nsamp if ld is None else ld,
                                          nsamp if stats_nsamp is None else stats_nsamp, thresh,
                                          d_cands, cap, d_count, C.byref(m), C.byref(r), stream))
    return m.value, r.value


def threshold_global(d_x, nrows, nsamp, mean, rms, thresh, d_cands, cap, d_count, stream=None,
                     ld=None):
    check(_lib.mdsm_threshold_global(d_x, nrows, nsamp, nsamp if ld is None else ld, mean, rms,
                                     thresh, d_cands, cap, d_count, stream))


def detect_legacy(d_x, ndms, nsamp, lowdm, dmstep, thresh, d_cands, cap, d_count, d_mean=None,
                  d_rms=None, samples_per_chunk=0, nstats=0, stream=None, nvals=0):
    """Legacy median-of-five detector; returns the (mean, rms) it used, or
    None with caller-supplied per-chunk statistics."""
    m, r = _f(0.0), _f(0.0)
    check(_lib.mdsm_detect_legacy(d_x, ndms, nsamp, lowdm, dmstep, thresh, nvals, d_mean, d_rms,
                                  samples_per_chunk, nstats, d_cands, cap, d_count, C.byref(m),
                                  C.byref(r), stream))
    return None if d_mean is not None else (m.value, r.value)


def median_filter5(d_x, nrows, nsamp, d_out, stream=None):
    check(_lib.mdsm_median_filter5(d_x, nrows, nsamp, d_out, stream))


def detrend_normalise(d_x, nseg, L, stream=None):
    check(_lib.mdsm_detrend_normalise(d_x, nseg, L, stream))


def read_candidates(d_cands: DeviceBuffer, d_count: DeviceBuffer, cap: int):
    """Download (count, sorted candidate array); raises on overflow."""
    n = int(d_count.download(np.uint32, (1,))[0])
    if n > cap:
        raise MdsmError(f"candidate overflow: {n} > capacity {cap}")
    c = d_cands.download(CAND_DTYPE, (n,)) if n else np.zeros(0, CAND_DTYPE)
    sort_candidates(c)
    return c


# ------------------------------------------------------------- clustering
def sort_candidates(c: np.ndarray):
    assert c.dtype == CAND_DTYPE and c.flags.c_contiguous
    check(_lib.mdsm_sort_candidates(_ptr(c), len(c)))


def dbscan2d(c: np.ndarray, eps: int = 3, minpts: int = 4):
    lab = np.empty(len(c), np.int32)
    n = _u32(0)
    check(_lib.mdsm_dbscan2d(_ptr(c), len(c), eps, minpts, _ptr(lab), C.byref(n)))
    return lab, n.value


def cluster_peaks(c: np.ndarray, lab: np.ndarray, nclusters: int):
    peaks = np.zeros(nclusters, CAND_DTYPE)
    sizes = np.zeros(nclusters, np.uint32)
    check(_lib.mdsm_cluster_peaks(_ptr(c), _ptr(np.ascontiguousarray(lab, np.int32)), len(c),
                                  nclusters, _ptr(peaks), _ptr(sizes)))
    return peaks, sizes


def dbscan_ref(time, dm, snr, min_time, min_dm, min_snr, min_points):
    time = np.ascontiguousarray(time, np.float64)
    dm = np.ascontiguousarray(dm, np.float64)
    snr = np.ascontiguousarray(snr, np.float64)
    lab = np.empty(len(time), np.int32)
    n = _u32(0)
    check(_lib.mdsm_dbscan_ref(_ptr(time), _ptr(dm), _ptr(snr), len(time), min_time, min_dm,
                               min_snr, min_points, _ptr(lab), C.byref(n)))
    return lab, n.value


# ------------------------------------------------------------------ batch
def classify_cluster_ref(time, dm, snr, min_dm, dm_step, num_dms, min_pulse_width, fch1, foff, nchans):
    """Reference cluster classification (score, [maxDM, maxSNR, width, position])."""
    t, d, s = (np.ascontiguousarray(a, np.float64) for a in (time, dm, snr))
    score = _f()
    info = np.zeros(4, np.float64)
    check(_lib.mdsm_classify_cluster_ref(_ptr(t), _ptr(d), _ptr(s), t.size, min_dm, dm_step, num_dms,
                                         min_pulse_width, fch1, foff, nchans, C.byref(score), _ptr(info)))
    return score.value, info


class Batch:
    def __init__(self, surveys, dtype=F32):
        arr = (Survey * len(surveys))(*surveys)
        p = _vp()
        check(_lib.mdsm_batch_create(arr, len(surveys), dtype, C.byref(p)))
        self.handle = p.value
        self.surveys = list(surveys)
        self.dtype = dtype
        self.min_ld = []
        self.maxshift = []
        for b in range(len(surveys)):
            ms, ld = _u32(0), _u64(0)
            check(_lib.mdsm_batch_info(self.handle, b, C.byref(ms), C.byref(ld)))
            self.maxshift.append(ms.value)
            self.min_ld.append(ld.value)

    def set_variant(self, v):
        check(_lib.mdsm_batch_set_variant(self.handle, v))

    def set_mode(self, mode):
        check(_lib.mdsm_batch_set_mode(self.handle, mode))

    def bind(self, d_in, lds, d_out):
        n = len(self.surveys)
        self._keep = ((_vp * n)(*d_in), (_u64 * n)(*lds), (_vp * n)(*d_out))
        check(_lib.mdsm_batch_bind(self.handle, *self._keep))

    def autotune(self, d_in, lds, d_out, stream=None):
        """Library tile autotuner; returns (variant, ms of its timed launch).
        Leaves the batch unbound."""
        n = len(self.surveys)
        keep = ((_vp * n)(*d_in), (_u64 * n)(*lds), (_vp * n)(*d_out))
        v, ms = _i(0), _f(0.0)
        check(_lib.mdsm_batch_autotune(self.handle, *keep, stream, C.byref(v), C.byref(ms)))
        return v.value, ms.value

    def launch(self, stream=None):
        check(_lib.mdsm_batch_launch(self.handle, stream))

    def dedisperse(self, d_in, lds, d_out, stream=None):
        self.bind(d_in, lds, d_out)
        self.launch(stream)

    def close(self):
        if self.handle:
            check(_lib.mdsm_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -------------------------------------------------------------- generator
@dataclass
class Pulse:
    dm: float
    start: int
    width: int
    amp: float


def pulse_delays(s: Survey, dms):
    """Per-channel delays (FP64 rounded formula) for arbitrary DMs, via the
    library's own table builder (a one-trial survey per DM)."""
    out = np.zeros((len(dms), s.nchans), np.int64)
    for i, dm in enumerate(dms):
        one = Survey.make(s.fch1, s.foff, s.nchans, s.tsamp, float(dm), 0.0, 1, 1)
        t, _ = delay_table(one)
        out[i] = t[:, 0]
    return out


def generate(seed, s: Survey, t0, n, d_raw, ld, gain, offset, pulses=(), stream=None):
    pulses = list(pulses)
    gain = np.ascontiguousarray(gain, np.float32)
    offset = np.ascontiguousarray(offset, np.float32)
    if pulses:
        st = np.array([p.start for p in pulses], np.int64)
        w = np.array([p.width for p in pulses], np.int32)
        a = np.array([p.amp for p in pulses], np.float32)
        dl = np.ascontiguousarray(pulse_delays(s, [p.dm for p in pulses]))
        args = (_ptr(st), _ptr(w), _ptr(a), _ptr(dl))
    else:
        args = (None, None, None, None)
    check(_lib.mdsm_generate(seed, s.nchans, t0, n, _ptr(gain), _ptr(offset), len(pulses), *args,
                             d_raw, ld, stream))


# --------------------------------------------------------------- pipeline
class DetectParams(C.Structure):
    _fields_ = [("stats_block", C.c_uint32), ("nwidths", C.c_int32), ("thresh", C.c_double),
                ("cand_cap", C.c_uint32)]


_sig("mdsm_pipeline_create", _P(Survey), _u32, _i, _P(DetectParams), _P(_vp))
_sig("mdsm_pipeline_destroy", _vp)
_sig("mdsm_pipeline_bind", _vp, _P(_vp), _P(_u64))
_sig("mdsm_pipeline_set_variant", _vp, _i)
_sig("mdsm_pipeline_run", _vp, _vp, _i)
_sig("mdsm_pipeline_candidates", _vp, _u32, _vp, _u32, _P(_u32))
_sig("mdsm_pipeline_output", _vp, _u32, _P(_vp))


class Pipeline:
    """normalise -> batched dedispersion -> block statistics -> detection,
    optionally recorded into a hipGraph and replayed."""

    def __init__(self, surveys, dtype=F16, stats_block=16384, nwidths=7, thresh=6.0, cand_cap=1 << 20):
        arr = (Survey * len(surveys))(*surveys)
        self.params = DetectParams(stats_block, nwidths, thresh, cand_cap)
        p = _vp()
        check(_lib.mdsm_pipeline_create(arr, len(surveys), dtype, C.byref(self.params), C.byref(p)))
        self.handle = p.value
        self.surveys = list(surveys)
        self.cap = cand_cap

    def bind(self, d_raw, lds):
        n = len(self.surveys)
        self._keep = ((_vp * n)(*d_raw), (_u64 * n)(*lds))
        check(_lib.mdsm_pipeline_bind(self.handle, *self._keep))

    def set_variant(self, v):
        check(_lib.mdsm_pipeline_set_variant(self.handle, v))

    def run(self, stream=None, graph=False):
        check(_lib.mdsm_pipeline_run(self.handle, stream, 1 if graph else 0))

    def candidates(self, beam):
        n = _u32(0)
        out = np.zeros(self.cap, CAND_DTYPE)
        check(_lib.mdsm_pipeline_candidates(self.handle, beam, _ptr(out), self.cap, C.byref(n)))
        return out[:n.value].copy()

    def output_ptr(self, beam):
        p = _vp()
        check(_lib.mdsm_pipeline_output(self.handle, beam, C.byref(p)))
        return p.value

    def close(self):
        if self.handle:
            check(_lib.mdsm_pipeline_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ SIGPROC ingest
class SigprocHeader(C.Structure):
    _fields_ = ([("source_name", C.c_char * 81), ("rawdatafile", C.c_char * 81)] +
                [(k, C.c_int32) for k in ("machine_id", "telescope_id", "data_type", "nchans", "nbits",
                                          "nifs", "nbeams", "ibeam", "barycentric", "pulsarcentric",
                                          "nbins", "nsamples")] +
                [(k, C.c_double) for k in ("tstart", "tsamp", "fch1", "foff", "refdm", "az_start",
                                           "za_start", "src_raj", "src_dej", "period")] +
                [("npuls", C.c_int64), ("nfchannel", C.c_int32), ("has_header", C.c_int32),
                 ("total_bytes", C.c_int32), ("data_offset", C.c_int64),
                 ("frequency_table", C.c_double * 4096)])


_sig("mdsm_sigproc_read_header", C.c_char_p, _P(SigprocHeader))
_sig("mdsm_unpack_corner_turn", _vp, _i, _u32, _u32, _u32, _vp, _u64, _u64, _u64, _vp)
_sig("mdsm_sigproc_open", C.c_char_p, _u32, _u32, _u32, _P(_vp), _P(SigprocHeader))
_sig("mdsm_sigproc_close", _vp)
_sig("mdsm_sigproc_read", _vp, _u32, _vp, _u64, _u64, _u64, _P(_u32), _vp)
_sig("mdsm_dm_writer_open", C.c_char_p, _vp, _u32, _d, _d, _d, _d, _P(_vp))
_sig("mdsm_dm_writer_write", _vp, _vp, _u32, _u32, _f, _f)
_sig("mdsm_dm_writer_close", _vp)


class DmWriter:
    """Per-DM SIGPROC files of dedispersed series (host rows)."""

    def __init__(self, prefix, dms, fch1, foff, tsamp, tstart=0.0):
        self.dms = np.ascontiguousarray(dms, np.float32)
        self.handle = _vp()
        check(_lib.mdsm_dm_writer_open(prefix.encode(), _ptr(self.dms), self.dms.size, fch1, foff, tsamp,
                                       tstart, C.byref(self.handle)))

    def write(self, series, lowdm, dmstep):
        series = np.ascontiguousarray(series, np.float32)
        check(_lib.mdsm_dm_writer_write(self.handle, _ptr(series), series.shape[0], series.shape[1], lowdm,
                                        dmstep))

    def close(self):
        if self.handle:
            check(_lib.mdsm_dm_writer_close(self.handle))
            self.handle = None


def sigproc_read_header(path: str) -> SigprocHeader:
    h = SigprocHeader()
    check(_lib.mdsm_sigproc_read_header(path.encode(), C.byref(h)))
    return h


def unpack_corner_turn(d_packed, nbits, nspectra, nchans, nbeams, d_out, ld, beam_stride=None,
                       col0=0, stream=None):
    beam_stride = nchans * ld if beam_stride is None else beam_stride
    check(_lib.mdsm_unpack_corner_turn(d_packed, nbits, nspectra, nchans, nbeams, d_out, ld,
                                       beam_stride, col0, stream))


class SigprocReader:
    """Streaming filterbank reader: packed spectra -> device [beam][chan][t] rows."""

    def __init__(self, path: str, nchans: int = 0, nbeams: int = 1, max_spectra: int = 65536):
        self.handle = _vp()
        self.header = SigprocHeader()
        check(_lib.mdsm_sigproc_open(path.encode(), nchans, nbeams, max_spectra, C.byref(self.handle),
                                     C.byref(self.header)))
        self.nbeams = nbeams
        self.nchans = nchans or self.header.nchans // nbeams

    def read(self, nspectra, d_out, ld, beam_stride=None, col0=0, stream=None) -> int:
        beam_stride = self.nchans * ld if beam_stride is None else beam_stride
        got = _u32(0)
        check(_lib.mdsm_sigproc_read(self.handle, nspectra, d_out, ld, beam_stride, col0,
                                     C.byref(got), stream))
        return got.value

    def close(self):
        if self.handle:
            check(_lib.mdsm_sigproc_close(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ bandpass / RFI
class RfiParams(C.Structure):
    _fields_ = [("ncoeffs", C.c_uint32), ("channel_block", C.c_uint32), ("channel_thresh", C.c_float),
                ("spectrum_thresh", C.c_float), ("nmasks", C.c_uint32), ("mask_from", _P(C.c_int32)),
                ("mask_to", _P(C.c_int32)), ("apply_channel_clip", C.c_int32)]

    @classmethod
    def make(cls, ncoeffs=8, channel_block=1024, channel_thresh=10.0, spectrum_thresh=6.0, masks=(),
             apply_channel_clip=False):
        p = cls(ncoeffs, channel_block, channel_thresh, spectrum_thresh, len(masks), None, None,
                1 if apply_channel_clip else 0)
        if masks:
            p._fr = (C.c_int32 * len(masks))(*[m[0] for m in masks])
            p._to = (C.c_int32 * len(masks))(*[m[1] for m in masks])
            p.mask_from = C.cast(p._fr, _P(C.c_int32))
            p.mask_to = C.cast(p._to, _P(C.c_int32))
        return p


class BandpassStats(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("mean", "std", "rmse", "channel_thresh", "spectrum_hi",
                                         "spectrum_lo")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_sig("mdsm_bandpass_sum", _vp, _u64, _u32, _u64, _u64, _vp, _vp)
_sig("mdsm_bandpass_fit", _vp, _u32, _P(RfiParams), _vp, _P(BandpassStats))
_sig("mdsm_rfi_channel_flag", _vp, _u64, _u32, _u64, _u64, _vp, _u32, _f, _vp, _vp)
_sig("mdsm_rfi_channel_clip", _vp, _u64, _u32, _u64, _u64, _vp, _u32, _vp, _vp)
_sig("mdsm_rfi_spectrum_clip", _vp, _u64, _u32, _u64, _u64, _vp, _f, _f, _vp, _vp)
_sig("mdsm_voltage_to_power", _vp, _u64, _u32, _u64, _u64, _vp)
_sig("mdsm_rfi_clean", _vp, _u64, _u32, _u64, _u64, _P(RfiParams), _vp, _vp, _P(BandpassStats), _vp)


def bandpass_fit(bandpass: np.ndarray, params: RfiParams):
    bp = np.ascontiguousarray(bandpass, np.float64)
    fitted = np.empty_like(bp)
    st = BandpassStats()
    check(_lib.mdsm_bandpass_fit(_ptr(bp), bp.size, C.byref(params), _ptr(fitted), C.byref(st)))
    return fitted, st.as_dict()


def rfi_clean(d_x, ld, nchans, col0, n, params: RfiParams, d_fitted, d_flags=None, stream=None):
    st = BandpassStats()
    check(_lib.mdsm_rfi_clean(d_x, ld, nchans, col0, n, C.byref(params), d_fitted, d_flags,
                              C.byref(st), stream))
    return st.as_dict()


# ------------------------------------------------------------ triage network
_sig("mdsm_cnn_create", _i, _i, _i, _P(_vp))
_sig("mdsm_cnn_destroy", _vp)
_sig("mdsm_cnn_param_count", _vp, _P(_u64))
_sig("mdsm_cnn_set_params", _vp, _vp)
_sig("mdsm_cnn_get_params", _vp, _vp)
_sig("mdsm_cnn_get_grads", _vp, _vp)
_sig("mdsm_cnn_init", _vp, C.c_uint64)
_sig("mdsm_cnn_forward", _vp, _vp, _u32, _vp, _vp)
_sig("mdsm_cnn_backward", _vp, _vp, _u32, _vp, _vp)
_sig("mdsm_cnn_adam_step", _vp, _f, _f, _vp)
_sig("mdsm_cnn_train_step", _vp, _vp, _vp, _u32, _f, _vp, _vp)


def to_bf16_bits(x) -> np.ndarray:
    """float32 -> BF16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(h) -> np.ndarray:
    return (np.asarray(h, np.uint16).astype(np.uint32) << 16).view(np.float32)


class TriageNet:
    """The BF16 triage CNN (forward, backward and Adam on the device)."""

    def __init__(self, height=64, width=256, max_batch=256):
        self.handle = _vp()
        check(_lib.mdsm_cnn_create(height, width, max_batch, C.byref(self.handle)))
        n = _u64()
        check(_lib.mdsm_cnn_param_count(self.handle, C.byref(n)))
        self.nparams = n.value
        self.H, self.W, self.max_batch = height, width, max_batch

    def init(self, seed):
        check(_lib.mdsm_cnn_init(self.handle, seed))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        check(_lib.mdsm_cnn_set_params(self.handle, _ptr(p)))

    def params(self):
        out = np.empty(self.nparams, np.float32)
        check(_lib.mdsm_cnn_get_params(self.handle, _ptr(out)))
        return out

    def grads(self):
        out = np.empty(self.nparams, np.float32)
        check(_lib.mdsm_cnn_get_grads(self.handle, _ptr(out)))
        return out

    def forward(self, d_x, batch, d_logits=None, stream=None):
        check(_lib.mdsm_cnn_forward(self.handle, d_x, batch, d_logits, stream))

    def backward(self, d_labels, batch, d_loss=None, stream=None):
        check(_lib.mdsm_cnn_backward(self.handle, d_labels, batch, d_loss, stream))

    def adam_step(self, lr, weight_decay=0.0, stream=None):
        check(_lib.mdsm_cnn_adam_step(self.handle, lr, weight_decay, stream))

    def train_step(self, d_x, d_labels, batch, lr, d_loss=None, stream=None):
        check(_lib.mdsm_cnn_train_step(self.handle, d_x, d_labels, batch, lr, d_loss, stream))

    def close(self):
        if self.handle:
            check(_lib.mdsm_cnn_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ triage data
class RfiItem(C.Structure):
    _fields_ = [("chan", C.c_uint32), ("kind", C.c_int32), ("start", C.c_int64), ("len", C.c_int64),
                ("amp", C.c_float), ("period", C.c_float)]


RFI_BURST, RFI_MAINS, RFI_TELEGRAPH = 0, 1, 2


class TriageConfig(C.Structure):
    _fields_ = [("survey", Survey), ("npulses", C.c_uint32), ("dm_lo", _d), ("dm_hi", _d),
                ("width_lo", _d), ("width_hi", _d), ("snr_lo", _d), ("snr_hi", _d),
                ("n_spikes", C.c_uint32), ("n_bursts", C.c_uint32), ("n_mains", C.c_uint32),
                ("n_telegraph", C.c_uint32), ("rfi_lo", _d), ("rfi_hi", _d), ("thresh", _d),
                ("stats_block", C.c_uint32), ("nwidths", C.c_int32), ("eps", C.c_int32),
                ("minpts", C.c_int32), ("snip_dms", C.c_int32), ("snip_samples", C.c_int32),
                ("snip_subbands", C.c_int32)]

    @classmethod
    def make(cls, survey, npulses=20, dm=(10.0, 800.0), width=(1.0, 64.0), snr=(6.0, 30.0),
             n_spikes=8, n_bursts=6, n_mains=2, n_telegraph=2, rfi=(8.0, 40.0), thresh=6.0,
             stats_block=16384, nwidths=7, eps=3, minpts=4, snip=(64, 256)):
        return cls(survey, npulses, dm[0], dm[1], width[0], width[1], snr[0], snr[1], n_spikes, n_bursts,
                   n_mains, n_telegraph, rfi[0], rfi[1], thresh, stats_block, nwidths, eps, minpts,
                   snip[0], snip[1], snip[0])


class TriageStats(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("chunks", "injected", "injected_found", "clusters",
                                          "false_clusters", "candidates")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


EVENT_DTYPE = np.dtype([("dm_idx", "<i4"), ("sample", "<i4"), ("width_log2", "<i4"), ("snr", "<f4"),
                        ("cluster_size", "<i4"), ("label", "<i4"), ("pulse", "<i4"), ("chunk", "<i4")])

_sig("mdsm_inject_rfi", _vp, _u32, _vp, C.c_uint64, _vp, _u64, _u64, _vp)
_sig("mdsm_triage_snippets", _vp, _u32, _u32, _vp, _u64, _u32, _vp, _vp, _vp, _u32, _i, _i, _i, _f, _f,
     _f, _vp, _vp)
_sig("mdsm_triage_sim_create", _P(TriageConfig), C.c_uint64, _P(_vp))
_sig("mdsm_triage_sim_destroy", _vp)
_sig("mdsm_triage_sim_run", _vp, _u32, _vp, _vp, _u32, _P(_u32), _P(TriageStats))
_sig("mdsm_cnn_train_epoch", _vp, _vp, _vp, _u32, _u32, _f, C.c_uint64, _P(_f), _vp)
_sig("mdsm_cnn_predict", _vp, _vp, _u32, _vp, _vp)


def inject_rfi(items, d_gain, seed, d_raw, ld, n, stream=None):
    arr = (RfiItem * len(items))(*items)
    check(_lib.mdsm_inject_rfi(C.cast(arr, _vp), len(items), d_gain, seed, d_raw, ld, n, stream))


def triage_snippets(d_out, ndms, nsamp, d_store, ld, nchans, d_table, ev_dm, ev_t, d_snip, SD=64, ST=256,
                    SB=64, scale0=1.0, scale1=1.0, clip=20.0, stream=None):
    ev_dm = np.ascontiguousarray(ev_dm, np.int32)
    ev_t = np.ascontiguousarray(ev_t, np.int32)
    check(_lib.mdsm_triage_snippets(d_out, ndms, nsamp, d_store, ld, nchans, d_table, _ptr(ev_dm),
                                    _ptr(ev_t), ev_dm.size, SD, ST, SB, scale0, scale1, clip, d_snip,
                                    stream))


class TriageSim:
    """Seeded synthetic search stream producing labelled events + snippets."""

    def __init__(self, config: TriageConfig, seed: int):
        self.config = config
        self.handle = _vp()
        check(_lib.mdsm_triage_sim_create(C.byref(config), seed, C.byref(self.handle)))
        self.stats = TriageStats()

    def run(self, nchunks, d_snip, max_events, events=None, n_events=0):
        if events is None:
            events = np.zeros(max_events, EVENT_DTYPE)
        n = _u32(n_events)
        check(_lib.mdsm_triage_sim_run(self.handle, nchunks, d_snip, _ptr(events), max_events, C.byref(n),
                                       C.byref(self.stats)))
        return events, n.value

    def close(self):
        if self.handle:
            check(_lib.mdsm_triage_sim_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def train_epoch(net, d_snip, d_labels, n, batch, lr, seed, stream=None):
    loss = _f()
    check(_lib.mdsm_cnn_train_epoch(net.handle, d_snip, d_labels, n, batch, lr, seed, C.byref(loss), stream))
    return loss.value


def predict(net, d_snip, n, d_prob, stream=None):
    check(_lib.mdsm_cnn_predict(net.handle, d_snip, n, d_prob, stream))

