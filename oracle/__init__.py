"""CPU oracle for the dedispersion + single-pulse detection path.

TEST INFRASTRUCTURE ONLY: this package is the oracle.  It may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py``; the product package ``mdsm_amd`` never imports it.

The arithmetic lives in ``oracle.c`` (built to ``liboracle.so`` by
``oracle/Makefile``); this module only marshals numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build() -> None:
    """Compile liboracle.so from oracle.c (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = _load()

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i = C.c_int
_i64 = C.c_int64
_d = C.c_double
_f = C.c_float
_u64 = C.c_uint64
_u32 = C.c_uint32


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("oracle_philox4x32_10", None, _u32p, _u32p, _u32p)
_sig("oracle_channel_bandpass", None, _u64, _i, _d, _d, _d, _d, _f32p, _f32p)
_sig("oracle_generate", None, _u64, _i, _d, _d, _d, _i64, _i64, _f32p, _f32p, _i,
     _f64p, _i64p, _i32p, _f32p, _f32p, _i64)
_sig("oracle_delay_table", None, _d, _d, _i, _d, _d, _d, _i, _i32p)
_sig("oracle_dm_shifts_ref", None, _f, _f, _i, _f32p)
_sig("oracle_delay_table_ref", None, _f, _f, _i, _f, _f, _f, _i, _i32p)
_sig("oracle_delay_table_legacy", None, _f, _f, _i, _f, _f, _f, _i, _i32p)
_sig("oracle_maxshift_ref", _u32, _f32p, _f32p, _i, _i, _f, _f, _f, _i)
_sig("oracle_shared_shift_ref", _u32, _f, _f, _f, _i, _f, _i)
_sig("oracle_calculate_nsamp_ref", _u32, _u32, _u32, _u32, _u32, _u64)
_sig("oracle_dedisperse_f64", None, _f32p, _i64, _i, _i, _i, _i32p, _f64p)
_sig("oracle_dedisperse_f32seq", None, _f32p, _i64, _i, _i, _i, _i32p, _f32p)
_sig("oracle_dedisperse_f16", None, _u16p, _i64, _i, _i, _i, _i32p, _f32p)
_sig("oracle_dedisperse_f16_pairs", None, _u16p, _i64, _i, _i, _i, _i32p, _f32p)
_sig("oracle_dedisperse_u8", None, _u8p, _i64, _i, _i, _i, _i32p, _f32p)
_sig("oracle_f32_to_f16_array", None, _f32p, _u16p, _i64)
_sig("oracle_f16_to_f32_array", None, _u16p, _f32p, _i64)
_sig("oracle_channel_stats", None, _f32p, _i64, _i, _i64, _f64p, _f64p)
_sig("oracle_normalise_f32", None, _f32p, _i64, _i, _i64, _f64p, _f64p, _f32p, _i64)
_sig("oracle_normalise_f16", None, _f32p, _i64, _i, _i64, _f64p, _f64p, _u16p, _i64)
_sig("oracle_normalise_u8", None, _f32p, _i64, _i, _i64, _f64p, _f64p, _u8p, _i64)
_sig("oracle_block_stats", None, _f64p, _i, _i, _i, _f64p, _f64p)
_sig("oracle_detect", _i64, _f64p, _i, _i, _i, _f64p, _f64p, _i, _d, _i64,
     _i32p, _i32p, _i32p, _f32p)
_sig("oracle_global_stats", None, _f32p, _i64, _i, _i, C.POINTER(_d), C.POINTER(_d))
_sig("oracle_global_stats_seq", None, _f32p, _i64, _i, _i, C.POINTER(_d), C.POINTER(_d))
_sig("oracle_threshold_ref", _i64, _f32p, _i64, _i, _i, _d, _d, _f, _i64, _i32p, _i32p, _f32p)
_sig("oracle_det_log", _d, _d)
_sig("oracle_det_sincos_turn", None, _d, C.POINTER(_d), C.POINTER(_d))
_sig("oracle_dbscan2d", _i, _i32p, _i32p, _i, _i, _i, _i32p)
_sig("oracle_dbscan_ref", _i, _f64p, _f64p, _f64p, _u32, _f, _f, _f, _u32, _i32p)
_sig("oracle_median5", _f, _f, _f, _f, _f, _f)
_sig("oracle_median_filter5", None, _f32p, _i, _i, _f32p)
_sig("oracle_process_brute_legacy", _i64, _f32p, _i, _i, _f, _f, _f, _i64, _i64, _i32p, _i32p,
     _f32p, C.POINTER(_f), C.POINTER(_f))
_sig("oracle_process_brute_legacy_chunks", _i64, _f32p, _i, _i, _f, _f, _f, _f32p, _f32p, _i,
     _i64, _i32p, _i32p, _f32p)
_sig("oracle_detrend_normalise", None, _f32p, _i64, _i, _f32p)
_sig("oracle_num_threads", _i)
_sig("oracle_sigproc_unpack", None, _u8p, _i, _i64, _f32p)
_sig("oracle_bandpass_sum", None, _f32p, _i64, _i, _i64, _i64, _f64p)
_sig("oracle_bandpass_mask", None, _f64p, _i, _i32p, _i32p, _i)
_sig("oracle_poly_eval", None, _f64p, _i, _i, _f64p)
_sig("oracle_bandpass_stats", None, _f64p, _f64p, _i, _f, _f, _f32p)
_sig("oracle_f64_to_f32_rz", _f, _d)
_sig("oracle_channel_flagger", None, _f32p, _i64, _i, _i64, _i64, _f64p, _i, _f, _u8p)
_sig("oracle_channel_clipper", None, _f32p, _i64, _i, _i64, _i64, _f64p, _i, _u8p)
_sig("oracle_spectrum_clipper", _i64, _f32p, _i64, _i, _i64, _i64, _f64p, _f, _f)
_sig("oracle_voltage_to_power", None, _f32p, _i64, _i, _i64, _i64)
_sig("oracle_cnn_param_count", _i64, _i, _i)
_sig("oracle_cnn_loss_grad", _d, _f64p, _i32p, _i, _i, _i, _f64p, C.c_void_p, C.c_void_p, _i)
_sig("oracle_bf16_round", _d, _d)
_sig("oracle_mulaw_power", None, _f32p, _i64, _u32, _u32, _f, _u8p)
_sig("oracle_mulaw_complex", None, np.ctypeslib.ndpointer(np.int16, flags="C_CONTIGUOUS"), _i64, _u32, _u32, _u8p)
_sig("oracle_classify_ref", _f, _f64p, _f64p, _f64p, _i, _f, _f, _u32, _f, _f, _f, _u32, _f64p)
_sig("oracle_rfi_inject", None, C.c_void_p, _i, _f32p, _u64, _f32p, _i64, _i64)
_sig("oracle_snippet", None, _f32p, _i, _i, _f32p, _i64, _i, _i32p, _i, _i, _i, _i, _i, _f, _f, _f, _f32p)
_sig("oracle_corner_turn", None, _f32p, _i64, _i, _i, _f32p, _i64, _i64, _i64)

KDM = 4148.808
U8_SCALE = 8.0 / 255.0
U8_OFFSET = -4.0


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def philox4x32_10(ctr, key):
    out = np.zeros(4, np.uint32)
    _lib.oracle_philox4x32_10(_c(ctr, np.uint32), _c(key, np.uint32), out)
    return out


def channel_bandpass(seed, nchans, gain=(0.5, 2.0), offset=(0.0, 10.0)):
    g = np.empty(nchans, np.float32)
    o = np.empty(nchans, np.float32)
    _lib.oracle_channel_bandpass(seed, nchans, gain[0], gain[1], offset[0], offset[1], g, o)
    return g, o


def generate(seed, nchans, fch1, foff, tsamp, t0, n, gain=None, offset=None, pulses=(), ld=None):
    """Raw filterbank block [nchans][ld] (float32).  pulses: iterable of
    (dm, start, width, amp) with amp in normalised units."""
    ld = n if ld is None else ld
    if gain is None:
        gain = np.ones(nchans, np.float32)
    if offset is None:
        offset = np.zeros(nchans, np.float32)
    p = list(pulses)
    p_dm = np.array([q[0] for q in p] or [0.0], np.float64)
    p_st = np.array([q[1] for q in p] or [0], np.int64)
    p_w = np.array([q[2] for q in p] or [0], np.int32)
    p_a = np.array([q[3] for q in p] or [0.0], np.float32)
    out = np.zeros((nchans, ld), np.float32)
    _lib.oracle_generate(seed, nchans, fch1, foff, tsamp, t0, n, _c(gain, np.float32),
                         _c(offset, np.float32), len(p), p_dm, p_st, p_w, p_a, out, ld)
    return out


def delay_table(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms):
    t = np.empty((nchans, ndms), np.int32)
    _lib.oracle_delay_table(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms, t)
    return t


def dm_shifts_ref(fch1, foff, nchans):
    s = np.empty(nchans, np.float32)
    _lib.oracle_dm_shifts_ref(fch1, foff, nchans, s)
    return s


def delay_table_ref(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms):
    t = np.empty((nchans, ndms), np.int32)
    _lib.oracle_delay_table_ref(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms, t)
    return t


def delay_table_legacy(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms):
    t = np.empty((nchans, ndms), np.int32)
    _lib.oracle_delay_table_legacy(fch1, foff, nchans, tsamp, lowdm, dmstep, ndms, t)
    return t


def maxshift_ref(fch1s, foffs, nchans, tsamp, lowdm, dmstep, ndms):
    return int(_lib.oracle_maxshift_ref(_c(fch1s, np.float32), _c(foffs, np.float32), len(fch1s),
                                        nchans, tsamp, lowdm, dmstep, ndms))


def shared_shift_ref(bottom_shift, lowdm, dmstep, ndms, tsamp, dedisp_dms=32):
    return int(_lib.oracle_shared_shift_ref(bottom_shift, lowdm, dmstep, ndms, tsamp, dedisp_dms))


def calculate_nsamp_ref(nsamp, maxshift, nchans, ndms, memory_mb):
    return int(_lib.oracle_calculate_nsamp_ref(nsamp, maxshift, nchans, ndms, memory_mb))


def dedisperse_f64(x, table, nsamp):
    x = _c(x, np.float32)
    nchans, ld = x.shape
    ndms = table.shape[1]
    out = np.empty((ndms, nsamp), np.float64)
    _lib.oracle_dedisperse_f64(x, ld, nchans, nsamp, ndms, _c(table, np.int32), out)
    return out


def dedisperse_f32seq(x, table, nsamp, out=None):
    x = _c(x, np.float32)
    nchans, ld = x.shape
    ndms = table.shape[1]
    if out is None:
        out = np.empty((ndms, nsamp), np.float32)
    _lib.oracle_dedisperse_f32seq(x, ld, nchans, nsamp, ndms, _c(table, np.int32), out)
    return out


def dedisperse_f16(h, table, nsamp):
    h = _c(h, np.uint16)
    nchans, ld = h.shape
    ndms = table.shape[1]
    out = np.empty((ndms, nsamp), np.float32)
    _lib.oracle_dedisperse_f16(h, ld, nchans, nsamp, ndms, _c(table, np.int32), out)
    return out


def dedisperse_f16_pairs(h, table, nsamp):
    h = _c(h, np.uint16)
    nchans, ld = h.shape
    ndms = table.shape[1]
    out = np.empty((ndms, nsamp), np.float32)
    _lib.oracle_dedisperse_f16_pairs(h, ld, nchans, nsamp, ndms, _c(table, np.int32), out)
    return out


def dedisperse_u8(u, table, nsamp):
    u = _c(u, np.uint8)
    nchans, ld = u.shape
    ndms = table.shape[1]
    out = np.empty((ndms, nsamp), np.float32)
    _lib.oracle_dedisperse_u8(u, ld, nchans, nsamp, ndms, _c(table, np.int32), out)
    return out


def f32_to_f16(x):
    x = _c(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    _lib.oracle_f32_to_f16_array(x.reshape(-1), out.reshape(-1), x.size)
    return out


def f16_to_f32(h):
    h = _c(h, np.uint16)
    out = np.empty(h.shape, np.float32)
    _lib.oracle_f16_to_f32_array(h.reshape(-1), out.reshape(-1), h.size)
    return out


def channel_stats(raw, n=None):
    raw = _c(raw, np.float32)
    nchans, ld = raw.shape
    n = ld if n is None else n
    m = np.empty(nchans, np.float64)
    s = np.empty(nchans, np.float64)
    _lib.oracle_channel_stats(raw, ld, nchans, n, m, s)
    return m, s


def normalise(raw, mean, std, dtype="f32", n=None):
    raw = _c(raw, np.float32)
    nchans, ld = raw.shape
    n = ld if n is None else n
    mean = _c(mean, np.float64)
    std = _c(std, np.float64)
    if dtype == "f32":
        out = np.empty((nchans, n), np.float32)
        _lib.oracle_normalise_f32(raw, ld, nchans, n, mean, std, out, n)
    elif dtype == "f16":
        out = np.empty((nchans, n), np.uint16)
        _lib.oracle_normalise_f16(raw, ld, nchans, n, mean, std, out, n)
    elif dtype == "u8":
        out = np.empty((nchans, n), np.uint8)
        _lib.oracle_normalise_u8(raw, ld, nchans, n, mean, std, out, n)
    else:
        raise ValueError(dtype)
    return out


def block_stats(x, L):
    x = _c(x, np.float64)
    nrows, nsamp = x.shape
    nb = (nsamp + L - 1) // L
    m = np.empty((nrows, nb), np.float64)
    s = np.empty((nrows, nb), np.float64)
    _lib.oracle_block_stats(x, nrows, nsamp, L, m, s)
    return m, s


def detect(x, L, nwidths, thresh, cap=1 << 20, stats=None):
    """Returns (d, t, k, snr) arrays sorted by (d, t) and the total count."""
    x = _c(x, np.float64)
    nrows, nsamp = x.shape
    m, s = block_stats(x, L) if stats is None else stats
    cd = np.empty(cap, np.int32)
    ct = np.empty(cap, np.int32)
    ck = np.empty(cap, np.int32)
    cs = np.empty(cap, np.float32)
    n = _lib.oracle_detect(x, nrows, nsamp, L, _c(m, np.float64), _c(s, np.float64), nwidths,
                           thresh, cap, cd, ct, ck, cs)
    k = min(n, cap)
    return cd[:k].copy(), ct[:k].copy(), ck[:k].copy(), cs[:k].copy(), int(n)


def global_stats(x, stats_nsamp=None, order="device"):
    """Reference-semantics (mean, rms) of the 2-D array x over its first
    stats_nsamp columns; order "device" is the product's fixed chunked
    order, "sequential" the reference's one loop."""
    x = _c(x, np.float32)
    nrows, ncols = x.shape
    sc = ncols if stats_nsamp is None else stats_nsamp
    m = C.c_double()
    r = C.c_double()
    fn = _lib.oracle_global_stats if order == "device" else _lib.oracle_global_stats_seq
    fn(x.reshape(-1), ncols, nrows, sc, C.byref(m), C.byref(r))
    return m.value, r.value


def threshold_ref(x, mean, rms, thresh, cap=1 << 20):
    """Candidates (row, sample, snr float32) of x - mean >= rms*thresh over
    the whole 2-D array x, in row-major order, and their total count."""
    x = _c(x, np.float32)
    ndms, nsamp = x.shape
    cd = np.empty(cap, np.int32)
    ct = np.empty(cap, np.int32)
    cs = np.empty(cap, np.float32)
    n = _lib.oracle_threshold_ref(x, nsamp, ndms, nsamp, mean, rms, thresh, cap, cd, ct, cs)
    k = min(n, cap)
    return cd[:k].copy(), ct[:k].copy(), cs[:k].copy(), int(n)


def det_log(x):
    return _lib.oracle_det_log(float(x))


def det_sincos_turn(u):
    s, c = C.c_double(), C.c_double()
    _lib.oracle_det_sincos_turn(float(u), C.byref(s), C.byref(c))
    return s.value, c.value


def dbscan2d(d, t, eps, minpts):
    d = _c(d, np.int32)
    t = _c(t, np.int32)
    lab = np.empty(len(d), np.int32)
    n = _lib.oracle_dbscan2d(d, t, len(d), eps, minpts, lab)
    return lab, int(n)


def dbscan_ref(time, dm, snr, min_time, min_dm, min_snr, min_points):
    time = _c(time, np.float64)
    dm = _c(dm, np.float64)
    snr = _c(snr, np.float64)
    lab = np.empty(len(time), np.int32)
    n = _lib.oracle_dbscan_ref(time, dm, snr, len(time), min_time, min_dm, min_snr, min_points, lab)
    return lab, int(n)


def median5(a, b, c, d, e):
    return float(_lib.oracle_median5(a, b, c, d, e))


def median_filter5(x):
    x = _c(x, np.float32)
    out = np.empty_like(x)
    _lib.oracle_median_filter5(x, x.shape[0], x.shape[1], out)
    return out


def process_brute_legacy(x, lowdm, dmstep, thresh, cap=1 << 20, nvals=0):
    x = _c(x, np.float32)
    ndms, nsamp = x.shape
    cd = np.empty(cap, np.int32)
    ct = np.empty(cap, np.int32)
    cs = np.empty(cap, np.float32)
    m = C.c_float()
    r = C.c_float()
    n = _lib.oracle_process_brute_legacy(x, ndms, nsamp, lowdm, dmstep, thresh, nvals, cap, cd, ct, cs,
                                         C.byref(m), C.byref(r))
    k = min(n, cap)
    return cd[:k].copy(), ct[:k].copy(), cs[:k].copy(), int(n), m.value, r.value


def process_brute_legacy_chunks(x, lowdm, dmstep, thresh, mean, rms, spc, cap=1 << 20):
    """Legacy detector with caller-supplied per-chunk mean and rms."""
    x = _c(x, np.float32)
    ndms, nsamp = x.shape
    mean = _c(mean, np.float32)
    rms = _c(rms, np.float32)
    assert spc > 0 and mean.size >= max(nsamp - 5, 0) // spc + 1 and rms.size == mean.size
    cd = np.empty(cap, np.int32)
    ct = np.empty(cap, np.int32)
    cs = np.empty(cap, np.float32)
    n = _lib.oracle_process_brute_legacy_chunks(x, ndms, nsamp, lowdm, dmstep, thresh, mean, rms,
                                                spc, cap, cd, ct, cs)
    k = min(n, cap)
    return cd[:k].copy(), ct[:k].copy(), cs[:k].copy(), int(n)


def detrend_normalise(x, L):
    x = _c(x, np.float32).reshape(-1)
    nseg = x.size // L
    out = np.empty(nseg * L, np.float32)
    _lib.oracle_detrend_normalise(x, nseg, L, out)
    return out


def num_threads():
    return int(_lib.oracle_num_threads())


def sigproc_unpack(raw_bytes, nbits, n):
    """Unpack n samples of width nbits from a byte string (read_block)."""
    b = np.frombuffer(bytes(raw_bytes), np.uint8).copy()
    need = (n * nbits + 7) // 8
    if b.size < need:
        raise ValueError("not enough bytes")
    out = np.empty(n, np.float32)
    _lib.oracle_sigproc_unpack(b if b.size else np.zeros(1, np.uint8), nbits, n, out)
    return out


def corner_turn(spectra, nchans, nbeams, ld=None, beam_stride=None, col0=0, out=None):
    """[t][nbeams*nchans] -> out[b*beam_stride + c*ld + col0 + t] (readBinaryData)."""
    spectra = _c(spectra, np.float32).reshape(-1, nchans * nbeams)
    nspec = spectra.shape[0]
    ld = nspec + col0 if ld is None else ld
    beam_stride = nchans * ld if beam_stride is None else beam_stride
    if out is None:
        out = np.zeros(nbeams * beam_stride, np.float32)
    _lib.oracle_corner_turn(spectra, nspec, nchans, nbeams, out, ld, beam_stride, col0)
    return out


# reference file_handler.cpp compiled from its own source (oracle/Makefile
# target ref); present only where the MDSM sources are.
_REF_FILE = os.path.join(_HERE, "_ref", "libref_file.so")


def ref_file_available():
    return os.path.exists(_REF_FILE)


def ref_read_header(path):
    """The reference read_header on `path`: dict or None (no HEADER_START)."""
    lib = C.CDLL(_REF_FILE)
    lib.ref_read_header.restype = C.c_int
    lib.ref_read_header.argtypes = [C.c_char_p] + [C.POINTER(_d)] * 4 + [C.POINTER(_i)] * 4 + [C.c_char_p]
    d = [_d() for _ in range(4)]
    i = [_i() for _ in range(4)]
    name = C.create_string_buffer(96)
    r = lib.ref_read_header(path.encode(), *[C.byref(x) for x in d], *[C.byref(x) for x in i], name)
    if r < 0:
        raise OSError(path)
    if r == 0:
        return None
    keys = ["fch1", "foff", "tsamp", "tstart"]
    out = {k: v.value for k, v in zip(keys, d)}
    out.update({k: v.value for k, v in zip(["nchans", "nbits", "nifs", "total_bytes"], i)})
    out["source_name"] = name.value.decode()
    return out


REF_HEADER_FIELDS = ["machine_id", "telescope_id", "data_type", "nchans", "nbits", "nifs", "nbeams",
                     "ibeam", "barycentric", "pulsarcentric", "nbins", "nsamples", "tstart", "tsamp",
                     "fch1", "foff", "refdm", "az_start", "za_start", "src_raj", "src_dej", "period",
                     "npuls", "total_bytes", "file_position"]


def ref_header_values(path, ntable=0):
    """All header fields the reference read_header fills (dict), or None."""
    lib = C.CDLL(_REF_FILE)
    lib.ref_header_values.restype = C.c_int
    lib.ref_header_values.argtypes = [C.c_char_p, _f64p, C.c_char_p, C.c_char_p, _f64p, _i]
    v = np.zeros(len(REF_HEADER_FIELDS), np.float64)
    src = C.create_string_buffer(96)
    raw = C.create_string_buffer(96)
    table = np.zeros(max(ntable, 1), np.float64)
    r = lib.ref_header_values(path.encode(), v, src, raw, table, ntable)
    if r < 0:
        raise OSError(path)
    if r == 0:
        return None
    out = dict(zip(REF_HEADER_FIELDS, v.tolist()))
    # unset strings hold uninitialised bytes in the reference
    out["source_name"] = src.value.decode(errors="replace")
    out["rawdatafile"] = raw.value.decode(errors="replace")
    out["frequency_table"] = table[:ntable]
    return out


def ref_read_block(path, byte_offset, nbits, nread):
    """The reference read_block: up to nread samples from byte_offset."""
    lib = C.CDLL(_REF_FILE)
    lib.ref_read_block.restype = C.c_long
    lib.ref_read_block.argtypes = [C.c_char_p, C.c_long, _i, _f32p, C.c_ulong]
    out = np.zeros(nread, np.float32)
    got = lib.ref_read_block(path.encode(), byte_offset, nbits, out, nread)
    return out[:got]


# ------------------------------------------------------------ bandpass / RFI
def bandpass_sum(x, col0, n):
    """Per-channel mean of x[c][col0:col0+n] in the reference's order."""
    x = _c(x, np.float32)
    out = np.empty(x.shape[0], np.float64)
    _lib.oracle_bandpass_sum(x, x.shape[1], x.shape[0], col0, n, out)
    return out


def bandpass_mask(bp, masks):
    bp = _c(bp, np.float64).copy()
    if masks:
        fr = np.array([m[0] for m in masks], np.int32)
        to = np.array([m[1] for m in masks], np.int32)
        _lib.oracle_bandpass_mask(bp, bp.size, fr, to, len(masks))
    return bp


def poly_fit(bp, ncoeffs):
    """Least squares fit of bp[i] by sum_j c_j x_i^j, x_i = i/nchans: SVD
    least squares (numpy lstsq), the method of gsl_multifit_linear."""
    n = bp.size
    xs = np.arange(n) / (1.0 * n)
    X = np.vander(xs, ncoeffs, increasing=True)
    return np.linalg.lstsq(X, bp, rcond=None)[0]


def poly_eval(coeffs, nchans):
    coeffs = _c(coeffs, np.float64)
    out = np.empty(nchans, np.float64)
    _lib.oracle_poly_eval(coeffs, coeffs.size, nchans, out)
    return out


def bandpass_fit(x, col0, n, ncoeffs=8, masks=(), channel_thresh=10.0, spectrum_thresh=6.0):
    """The reference's bandpass_fitting + rfi_clipping thresholds:
    returns (fitted bandpass, dict of mean/std/rmse/thresholds)."""
    bp = bandpass_mask(bandpass_sum(x, col0, n), list(masks))
    fitted = poly_eval(poly_fit(bp, ncoeffs), bp.size)
    return fitted, bandpass_stats(bp, fitted, channel_thresh, spectrum_thresh)


def bandpass_stats(bp, fitted, channel_thresh, spectrum_thresh):
    out = np.empty(6, np.float32)
    _lib.oracle_bandpass_stats(_c(bp, np.float64), _c(fitted, np.float64), bp.size, channel_thresh,
                               spectrum_thresh, out)
    return dict(zip(["mean", "std", "rmse", "channel_thresh", "spectrum_hi", "spectrum_lo"], out.tolist()))


def f64_to_f32_rz(d):
    return np.float32(_lib.oracle_f64_to_f32_rz(float(d)))


def channel_flagger(x, col0, n, fitted, cb, thresh):
    x = _c(x, np.float32)
    nb = (n + cb - 1) // cb
    flags = np.empty(x.shape[0] * nb, np.uint8)
    _lib.oracle_channel_flagger(x, x.shape[1], x.shape[0], col0, n, _c(fitted, np.float64), cb, thresh,
                                flags)
    return flags.reshape(x.shape[0], nb)


def channel_clipper(x, col0, n, fitted, cb, flags):
    """Returns a clipped copy of x."""
    x = _c(x, np.float32).copy()
    _lib.oracle_channel_clipper(x, x.shape[1], x.shape[0], col0, n, _c(fitted, np.float64), cb,
                                _c(flags, np.uint8).reshape(-1))
    return x


def spectrum_clipper(x, col0, n, fitted, hi, lo):
    """Returns (clipped copy of x, number of clipped spectra)."""
    x = _c(x, np.float32).copy()
    k = _lib.oracle_spectrum_clipper(x, x.shape[1], x.shape[0], col0, n, _c(fitted, np.float64), hi, lo)
    return x, int(k)


def voltage_to_power(x, col0, n):
    x = _c(x, np.float32).copy()
    _lib.oracle_voltage_to_power(x, x.shape[1], x.shape[0], col0, n)
    return x


# ------------------------------------------------------------ triage CNN
def cnn_param_count(H, W):
    return int(_lib.oracle_cnn_param_count(H, W))


def cnn_loss_grad(x, labels, params, bf16=False, grads=True):
    """Mean cross-entropy, gradient (flat, same layout as params) and logits
    of the triage CNN on x[B][2][H][W] (FP64 accumulation)."""
    x = _c(x, np.float64)
    B, _, H, W = x.shape
    params = _c(params, np.float64)
    assert params.size == cnn_param_count(H, W)
    g = np.empty_like(params) if grads else None
    logits = np.empty((B, 2), np.float64)
    loss = _lib.oracle_cnn_loss_grad(x, _c(labels, np.int32), B, H, W, params,
                                     g.ctypes.data if grads else None, logits.ctypes.data, 1 if bf16 else 0)
    return loss, g, logits


def bf16_round(v):
    return _lib.oracle_bf16_round(float(v))


def cnn_param_slices(H, W):
    """Names and [start, stop) of each parameter tensor in the flat layout."""
    ch = [2, 16, 32, 64]
    out, off = [], 0
    for k in range(3):
        n = ch[k + 1] * ch[k] * 9
        out.append((f"w{k + 1}", off, off + n)); off += n
        out.append((f"b{k + 1}", off, off + ch[k + 1])); off += ch[k + 1]
    fin = 64 * (H // 8) * (W // 8)
    for name, n in [("fw1", fin * 64), ("fb1", 64), ("fw2", 128), ("fb2", 2)]:
        out.append((name, off, off + n)); off += n
    return out


# ------------------------------------------------------------ triage data
RFI_ITEM_DTYPE = np.dtype([("chan", "<u4"), ("kind", "<i4"), ("start", "<i8"), ("len", "<i8"),
                           ("amp", "<f4"), ("period", "<f4")])


def rfi_inject(raw, items, gain, seed):
    """Returns a copy of raw [nchans][ld] with the RFI items added."""
    raw = _c(raw, np.float32).copy()
    items = np.ascontiguousarray(items, RFI_ITEM_DTYPE)
    _lib.oracle_rfi_inject(items.ctypes.data, items.size, _c(gain, np.float32), seed, raw, raw.shape[1],
                           raw.shape[1])
    return raw


def snippet(out, store, table, dm0, tc, SD=64, ST=256, SB=64, scale0=1.0, scale1=1.0, clip=20.0):
    """The two snippet planes [2][SD][ST] of one event (BF16 values)."""
    out = _c(out, np.float32)
    store = _c(store, np.float32)
    table = _c(table, np.int32)
    res = np.zeros((2, SD, ST), np.float32)
    _lib.oracle_snippet(out, out.shape[0], out.shape[1], store, store.shape[1], store.shape[0], table,
                        dm0, tc, SD, ST, SB, scale0, scale1, clip, res)
    return res


def classify_ref(time, dm, snr, min_dm, dm_step, num_dms, min_pulse_width, fch1, foff, nchans):
    """Reference cluster classification: (score, [maxDM, maxSNR, width, position])."""
    info = np.zeros(4, np.float64)
    t, d, s = (_c(a, np.float64) for a in (time, dm, snr))
    score = _lib.oracle_classify_ref(t, d, s, t.size, min_dm, dm_step, num_dms, min_pulse_width, fch1, foff,
                                     nchans, info)
    return float(score), info


def mulaw_power(x, bits, mu, max_value):
    x = _c(x, np.float32).reshape(-1)
    out = np.empty(x.size, np.uint8)
    _lib.oracle_mulaw_power(x, x.size, bits, mu, max_value, out)
    return out


def mulaw_complex(pairs, bits, mu):
    p = _c(pairs, np.int16).reshape(-1)
    out = np.empty(p.size // 2, np.uint8)
    _lib.oracle_mulaw_complex(p, out.size, bits, mu, out)
    return out

