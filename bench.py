"""Benchmark of the brute-force dedispersion hot path on one MI355X.

Headline workload (one step = one pass of the dedisperser over all beams):
32 beams of uneven sizes (nchans in {512, 1024}, nsamp uniform in
2^17..2^20, ndms uniform in 1024..4096, dmstep uniform in 0.1..0.5,
fch1 1500 MHz over a 300 MHz band, tsamp 64 us), Gaussian noise with
per-channel bandpass gain U(0.5, 2) and offset U(0, 10) plus 20 injected
pulses per beam, bandpass-normalised to FP16 storage, dedispersed by one
launch with FP32 accumulation.  The same beams are also run with U8
storage, through the plain path (FP32 storage, one launch per beam) for the
batched/FP16 vs plain/FP32 ratio, and batched with FP32 storage; a single
full beam (1024 channels, 2^20 samples, 4096 DMs at step 0.2) runs through
the plain FP32 path, and the per-chunk search pipeline (normalise, batched
dedispersion, block statistics, detection) on 16 beams is timed eager and
replayed from its hipGraph.  For four beams, sampled DM rows of every storage type
are compared with the oracle (skip with --no-precision): bit-exact against
the same-storage restatement, error against the FP64 sum of the
unquantised input, and the candidate lists of those rows.

All inputs are generated on the device and are resident in HBM before any
timed region; timing uses HIP events on the stream the kernels run on.
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from mdsm_amd import abi  # noqa: E402

HBM_PEAK = 8.0e12        # MI355X HBM3E, bytes/s
LDS_PEAK = 256 * 128 * 2.4e9  # bytes/s: 128 B per clock per CU, 256 CUs at 2.4 GHz
VALU_ADD_PEAK = 39.3e12  # FP32 vector adds/s, one per lane per clock (157.3 TFLOP/s counts
                         # packed FMA as 4 flops; packed FP32 is not used here)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_batch_fp16.json")
FCH1, BAND, TSAMP = 1500.0, 300.0, 64e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_beams(seed=2026, nbeams=32):
    rng = np.random.default_rng(seed)
    beams = []
    for b in range(nbeams):
        C = int(rng.choice([512, 1024]))
        N = int(rng.integers(1 << 17, (1 << 20) + 1))
        M = int(rng.integers(1024, 4097))
        step = float(rng.uniform(0.1, 0.5))
        s = abi.Survey.make(FCH1, -BAND / C, C, TSAMP, 0.0, step, M, N)
        dmmax = step * (M - 1)
        pulses = []
        for _ in range(20):
            dm = float(rng.uniform(10.0, min(800.0, 0.95 * dmmax)))
            w = int(round(np.exp(rng.uniform(0, np.log(64)))))
            snr = float(rng.uniform(8, 30))
            t0 = int(rng.integers(0, N - 64))
            pulses.append(abi.Pulse(dm, t0, w, snr / np.sqrt(C * w)))
        beams.append(dict(survey=s, pulses=pulses, seed=seed * 1000 + b))
    return beams


def single_beam():
    s = abi.Survey.make(FCH1, -BAND / 1024, 1024, TSAMP, 0.0, 0.2, 4096, 1 << 20)
    rng = np.random.default_rng(7)
    pulses = []
    for _ in range(20):
        w = int(round(np.exp(rng.uniform(0, np.log(64)))))
        pulses.append(abi.Pulse(float(rng.uniform(10, 800)), int(rng.integers(0, (1 << 20) - 64)), w,
                                float(rng.uniform(8, 30)) / np.sqrt(1024 * w)))
    return dict(survey=s, pulses=pulses, seed=77)


def bytes_per_beam(s, maxshift, in_bytes):
    """Algorithmic HBM bytes: input read once, output written once, delay table."""
    C, N, M = s.nchans, s.nsamp, s.ndms
    return C * (N + maxshift) * in_bytes + M * N * 4 + C * M * 4


class Stage:
    """Generates + normalises beams into device buffers of a storage dtype."""

    def __init__(self, stream):
        self.stream = stream
        self.raw = None
        self.raw_bytes = 0
        self.mean = abi.DeviceBuffer(1024 * 8)
        self.std = abi.DeviceBuffer(1024 * 8)

    def fill(self, beam, dst_ptr, ld, dtype):
        s = beam["survey"]
        need = s.nchans * ld * 4
        if need > self.raw_bytes:
            if self.raw is not None:
                self.raw.free()
            self.raw = abi.DeviceBuffer(need)
            self.raw_bytes = need
        rng = np.random.default_rng(beam["seed"])
        gain = rng.uniform(0.5, 2.0, s.nchans).astype(np.float32)
        off = rng.uniform(0.0, 10.0, s.nchans).astype(np.float32)
        abi.generate(beam["seed"], s, 0, ld, self.raw.ptr, ld, gain, off, beam["pulses"], self.stream)
        abi.channel_stats(self.raw.ptr, ld, s.nchans, ld, self.mean.ptr, self.std.ptr, self.stream)
        abi.normalise(self.raw.ptr, ld, s.nchans, ld, self.mean.ptr, self.std.ptr, dst_ptr, dtype, ld,
                      self.stream)

    def free(self):
        if self.raw is not None:
            self.raw.free()
            self.raw = None
            self.raw_bytes = 0


def time_steps(fn, steps, warmup, stream):
    for _ in range(warmup):
        fn()
    abi.synchronize()
    t = abi.Timer(stream)
    t.start()
    for _ in range(steps):
        fn()
    ms = t.stop()
    abi.synchronize()
    return ms / steps


def spot_check(beam, plan_like_table, d_in, ld, dtype, d_out, rows, want_error=False):
    """Bit-exact check of selected full-length DM rows against the oracle."""
    import oracle as O
    s = beam["survey"]
    np_dt = {abi.F32: np.float32, abi.F16: np.uint16, abi.U8: np.uint8, abi.F16_PAIRS: np.uint16}[dtype]
    x = d_in.download(np_dt, (s.nchans, ld)) if hasattr(d_in, "download") else None
    t = plan_like_table[:, rows]
    if dtype == abi.F32:
        ref = O.dedisperse_f32seq(x, np.ascontiguousarray(t), s.nsamp)
    elif dtype == abi.F16:
        ref = O.dedisperse_f16(x, np.ascontiguousarray(t), s.nsamp)
    elif dtype == abi.F16_PAIRS:
        ref = O.dedisperse_f16_pairs(x, np.ascontiguousarray(t), s.nsamp)
    else:
        ref = O.dedisperse_u8(x, np.ascontiguousarray(t), s.nsamp)
    ok = True
    got = np.stack([d_out.download(np.float32, (s.nsamp,), offset=int(d) * s.nsamp * 4) for d in rows])
    for i in range(len(rows)):
        ok &= bool(np.array_equal(got[i], ref[i]))
    if dtype == abi.F32 and want_error:
        # accumulation error of these rows against the FP64 sum of the same input
        y64 = O.dedisperse_f64(x, np.ascontiguousarray(t), s.nsamp)
        e = np.abs(got.astype(np.float64) - y64)
        return ok, {"max_err_over_max": float(e.max() / np.abs(y64).max()),
                    "max_err_over_rms": float(e.max() / np.sqrt((y64 * y64).mean()))}
    return ok


def detect_beam(beam, d_out, stream, L=16384, thr=6.0, cap=1 << 21):
    s = beam["survey"]
    M, N = s.ndms, s.nsamp
    nb = (N + L - 1) // L
    dm, ds = abi.DeviceBuffer(M * nb * 8), abi.DeviceBuffer(M * nb * 8)
    dc, dn = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    abi.block_stats(d_out.ptr, M, N, L, dm.ptr, ds.ptr, stream)
    abi.detect(d_out.ptr, M, N, L, dm.ptr, ds.ptr, 7, thr, dc.ptr, cap, dn.ptr, stream)
    abi.synchronize()
    c = abi.read_candidates(dc, dn, cap)
    lab, nc = abi.dbscan2d(c, 3, 4)
    peaks, sizes = abi.cluster_peaks(c, lab, nc)
    # recall: injected pulse matched by a cluster peak within 2% DM and one width in time
    found = 0
    for p in beam["pulses"]:
        di = p.dm / s.dmstep
        ok = [q for q in peaks if abs(q["dm_idx"] - di) <= max(2.0, 0.02 * di)
              and abs(int(q["sample"]) - p.start) <= max(2 * p.width, 4)]
        found += bool(ok)
    return len(c), nc, found


def cpu_model():
    """CPU model name and the cores this process may use (Linux)."""
    name = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return name, avail


def cpu_baseline(run_s=2.5, runs=5):
    """The oracle's multithreaded FP32 channel-order brute force (the
    reference's arithmetic, OpenMP over DM x time tiles, SIMD over time) on a
    bounded sample of the single-beam survey: 1024 channels x 2^17 samples x
    D DM trials (step 0.2, D <= 4096), D sized for about run_s seconds per
    run; the median of `runs` runs is reported, with every run's rate."""
    import oracle as O
    N = 1 << 17
    s = abi.Survey.make(FCH1, -BAND / 1024, 1024, TSAMP, 0.0, 0.2, 4096, N)
    table, S = abi.delay_table(s)
    x = np.random.default_rng(1).standard_normal((1024, N + S), dtype=np.float32)
    d0 = 128
    t0 = time.perf_counter()
    O.dedisperse_f32seq(x, np.ascontiguousarray(table[:, :d0]), N)
    el = time.perf_counter() - t0
    D = int(min(4096, max(d0, d0 * run_s / max(el, 1e-3))))
    D = max(64, (D // 64) * 64)
    tab = np.ascontiguousarray(table[:, :D])
    rates = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.dedisperse_f32seq(x, tab, N)
        rates.append(D * N / (time.perf_counter() - t0))
    model, avail = cpu_model()
    ref = None
    exe = os.path.join(ROOT, "oracle", "_ref", "cpu_dedisp")
    if os.path.exists(exe):
        # the reference's own single-threaded brute force (src/prototypes/cpu_dedisp.c)
        try:
            nref = 16384
            rr = []
            for _ in range(3):
                r = subprocess.run([exe, "-nchans", "1024", "-nsamp", str(nref), "-tdms", "64"],
                                   capture_output=True, text=True, timeout=120)
                ms = float([ln for ln in r.stdout.splitlines() if ln.startswith("Time:")][0].split()[1])
                rr.append(64 * nref / (ms / 1e3))
            ref = {"value": float(np.median(rr)), "unit": "DM*samples/s", "cores": 1, "kind": "reference",
                   "runs": rr, "cpu": model,
                   "sample": f"src/prototypes/cpu_dedisp.c, 1024 chans x 64 DMs x {nref} samples, time-major, "
                             "median of 3 runs"}
        except Exception as e:  # report, never fail the bench on the side measurement
            ref = {"error": str(e)}
    cpu = {"value": float(np.median(rates)), "unit": "DM*samples/s", "cores": O.num_threads(),
           "kind": "port", "runs": rates, "cpu": model, "cpus_available": avail,
           "sample": f"1024 chans x {D} DMs (step 0.2) x {N} samples of the single-beam survey, "
                     f"FP32 channel order, median of {runs} runs of about {run_s:.1f} s, "
                     f"{O.num_threads()} OpenMP threads"}
    return cpu, ref


def precision_rows(beam):
    """DM rows sampled for the precision report: first, last, and the rows of
    the first two injected pulses."""
    s = beam["survey"]
    rows = [0, s.ndms - 1] + [min(s.ndms - 1, int(round(p.dm / s.dmstep))) for p in beam["pulses"][:2]]
    return sorted(set(rows))


def precision_report(beams, prec_rows, L=16384, thr=6.0, nwidths=7):
    """Per beam and storage type, on the sampled rows: bit-exact against the
    oracle of the same storage, max and rms error against the FP64 sum of the
    unquantised normalised input (relative to the rms of that sum), and the
    6-sigma candidate lists of the rows (GPU detect on the GPU rows against
    the oracle's detect on the oracle's rows; also counted against the FP64
    rows).  The raw input is regenerated on the host by the oracle's
    generator, which the device generator matches bit for bit."""
    import oracle as O
    conv = {abi.F32: "f32", abi.F16: "f16", abi.U8: "u8", abi.F16_PAIRS: "f16_pairs"}
    out = []
    for b in sorted({b for _, b in prec_rows}):
        beam = beams[b]
        s = beam["survey"]
        rng = np.random.default_rng(beam["seed"])
        gain = rng.uniform(0.5, 2.0, s.nchans).astype(np.float32)
        off = rng.uniform(0.0, 10.0, s.nchans).astype(np.float32)
        lds = {dt: prec_rows[(dt, b)][1] for dt in conv if (dt, b) in prec_rows}
        ldmax = max(lds.values())
        t0 = time.perf_counter()
        raw = O.generate(beam["seed"], s.nchans, s.fch1, s.foff, s.tsamp, 0, ldmax, gain, off,
                         [(p.dm, p.start, p.width, p.amp) for p in beam["pulses"]], ldmax)
        table, _ = abi.delay_table(s)
        rec = {"beam": b, "nchans": s.nchans, "nsamp": s.nsamp, "ndms": s.ndms, "dmstep": s.dmstep}
        for dt, ld in lds.items():
            rows, _, got = prec_rows[(dt, b)]
            tab = np.ascontiguousarray(table[:, rows])
            m, sd = O.channel_stats(raw, ld)
            x32 = O.normalise(raw, m, sd, "f32", ld)
            y64 = O.dedisperse_f64(x32, tab, s.nsamp)
            if dt == abi.F32:
                ref = O.dedisperse_f32seq(x32, tab, s.nsamp)
            elif dt == abi.F16:
                ref = O.dedisperse_f16(O.normalise(raw, m, sd, "f16", ld), tab, s.nsamp)
            elif dt == abi.F16_PAIRS:
                ref = O.dedisperse_f16_pairs(O.normalise(raw, m, sd, "f16", ld), tab, s.nsamp)
            else:
                ref = O.dedisperse_u8(O.normalise(raw, m, sd, "u8", ld), tab, s.nsamp)
            err = got.astype(np.float64) - y64
            rms = float(np.sqrt((y64 * y64).mean()))
            d_rows = abi.DeviceBuffer.from_numpy(np.ascontiguousarray(got))
            nb = (s.nsamp + L - 1) // L
            dm_, ds_ = abi.DeviceBuffer(len(rows) * nb * 8), abi.DeviceBuffer(len(rows) * nb * 8)
            cap = 1 << 20
            dc, dn = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
            abi.block_stats(d_rows.ptr, len(rows), s.nsamp, L, dm_.ptr, ds_.ptr)
            abi.detect(d_rows.ptr, len(rows), s.nsamp, L, dm_.ptr, ds_.ptr, nwidths, thr, dc.ptr, cap, dn.ptr)
            abi.synchronize()
            c = abi.read_candidates(dc, dn, cap)
            o_d, o_t, o_k, _, _ = O.detect(ref, L, nwidths, thr)
            f_d, f_t, f_k, _, _ = O.detect(y64, L, nwidths, thr)
            gpu_set = set(zip(c["dm_idx"].tolist(), c["sample"].tolist(), c["width_log2"].tolist()))
            orc_set = set(zip(o_d.tolist(), o_t.tolist(), o_k.tolist()))
            f64_set = set(zip(f_d.tolist(), f_t.tolist(), f_k.tolist()))
            rec[conv[dt]] = {"rows": rows, "bitexact_vs_oracle": bool(np.array_equal(got, ref)),
                             "max_err_over_max": float(np.abs(err).max() / np.abs(y64).max()),
                             "max_err_over_rms": float(np.abs(err).max() / rms),
                             "rms_err_over_rms": float(np.sqrt((err * err).mean()) / rms),
                             "candidates": len(gpu_set), "candidates_equal_oracle": gpu_set == orc_set,
                             "candidates_fp64_rows": len(f64_set),
                             "candidates_differing_from_fp64": len(gpu_set ^ f64_set)}
        rec["host_seconds"] = time.perf_counter() - t0
        log(f"[precision] {json.dumps(rec)}")
        out.append(rec)
    return out


COMPAT_XML = """<observation>
  <dm lowDM="0" numDMs="{ndms}" dmStep="0.2" />
  <channels nchans="1024" npols="1" />
  <timing tsamp="0.000064" />
  <samples nsamp="{nsamp}" nbits="32" />
  <detection detectionThreshold="6" />
  <storage type="{storage}" />
  <series download="{download}" />
  <writer filePrefix="compat" baseDirectory="{dir}" singleFileMode="1" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-0.29296875" /></beams>
</observation>"""


def compat_throughput(nsamp=1 << 18, ndms=4096, nchunks=10, storage="fp32", download=True):
    """The drop-in manager API (get_buffer_pointer / next_chunk_multibeam /
    start_processing) driven as src/MDSM.cpp drives it, host buffers in and
    out: per chunk the input upload (PCIe), the overlap-save copy,
    dedispersion, the global-statistics threshold detection and, with
    download, the download of the dedispersed series into the caller's
    output buffers (PCIe; without it the series stay on the device, which
    src/MDSM.cpp, ignoring next_chunk_multibeam's return value, allows).
    The manager pipelines chunks across three streams, so the rate is taken
    over the whole stream: from the call of data chunk 3 (every staging
    buffer filled once by then) to the return of the final, empty
    start_processing call that drains the pipeline."""
    import ctypes as C
    import tempfile
    lib = abi.lib()
    lib.mdsm_compat_open.argtypes = [C.c_char_p, C.c_uint]
    lib.mdsm_compat_survey.argtypes = [C.POINTER(C.c_uint)] * 5
    lib.mdsm_compat_buffer.restype = C.POINTER(C.c_float)
    lib.mdsm_compat_next_chunk.argtypes = [C.c_uint, C.POINTER(C.c_uint), C.c_double, C.c_double]
    lib.mdsm_compat_next_chunk.restype = C.POINTER(C.POINTER(C.c_float))
    lib.mdsm_compat_start.argtypes = [C.c_uint]
    lib.mdsm_compat_last_error.restype = C.c_char_p
    with tempfile.TemporaryDirectory() as d:
        xml = os.path.join(d, "obs.xml")
        with open(xml, "w") as f:
            f.write(COMPAT_XML.format(ndms=ndms, nsamp=nsamp, dir=d, storage=storage,
                                      download=1 if download else 0))
        if lib.mdsm_compat_open(xml.encode(), 0) != 0:
            raise RuntimeError(lib.mdsm_compat_last_error().decode())
        nb, nch, ns, nd, ms = (C.c_uint() for _ in range(5))
        lib.mdsm_compat_survey(C.byref(nb), C.byref(nch), C.byref(ns), C.byref(nd), C.byref(ms))
        Cn, N, M, S = nch.value, ns.value, nd.value, ms.value
        x = np.random.default_rng(5).standard_normal((Cn, N)).astype(np.float32)
        # the three staging buffers are filled once (the manager uploads a
        # buffer's contents on every chunk whatever they are), so the wall
        # time per chunk is the manager's own rate, not the caller's copying
        filled = set()
        t_first = None
        for i in range(nchunks + 4):
            buf = lib.mdsm_compat_buffer()
            n = S if i == 0 else (N if i <= nchunks else 0)
            addr = C.cast(buf, C.c_void_p).value
            if n and (i == 0 or addr not in filled):
                np.ctypeslib.as_array(buf, shape=(Cn, n))[:] = x[:, :n]
                if i > 0:
                    filled.add(addr)
            samples = C.c_uint()
            if i == 3:
                t_first = time.perf_counter()  # steady state: every buffer filled
            lib.mdsm_compat_next_chunk(n, C.byref(samples), 64e-6 * i * N, 64e-6)
            r = lib.mdsm_compat_start(n)
            if r == 0:
                break
        t_end = time.perf_counter()
        lib.mdsm_compat_close()
    per = (t_end - t_first) / (nchunks - 2)  # chunks 3..nchunks and the drain
    pcie = Cn * N * 4 + (M * N * 4 if download else 0)
    return {"workload": f"1 beam, 1024 ch x {M} DMs (step 0.2) x {N} samples per chunk, host buffers, "
                        f"{storage} storage, series {'downloaded' if download else 'kept on the device'}",
            "s_per_chunk": per, "dm_samples_per_s": M * N / per,
            "pcie_bytes_per_chunk": pcie, "pcie_gbs": pcie / per / 1e9, "chunks_timed": nchunks - 2}


def detect_run(steps, rows=2048, nsamp=1 << 20):
    """The boxcar detector alone on a rows x nsamp FP32 array of device
    noise: block statistics (L = 16384) and detection (7 widths, 6 sigma)
    with the register-resident detector, then with a block length that is
    not a multiple of 16 (L = 16383, the LDS-tiled detector), and a device
    copy of the same bytes for the attainable rate.  Reports milliseconds
    and read rates (the workload of the detector's counter profiles)."""
    s = abi.Survey.make(FCH1, -BAND / rows, rows, TSAMP, 0.0, 0.2, 16, nsamp)
    B = rows * nsamp * 4
    d = abi.DeviceBuffer(B)
    abi.generate(3, s, 0, nsamp, d.ptr, nsamp, np.ones(rows, np.float32), np.zeros(rows, np.float32), [])
    cap = 1 << 22
    dc, dn = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    out = {"rows": rows, "nsamp": nsamp, "bytes": B}
    for L, name in ((16384, "register_resident"), (16383, "lds_tiled")):
        nb = (nsamp + L - 1) // L
        dm, ds = abi.DeviceBuffer(rows * nb * 8), abi.DeviceBuffer(rows * nb * 8)
        ms_bs = time_steps(lambda: abi.block_stats(d.ptr, rows, nsamp, L, dm.ptr, ds.ptr), steps, 1, None)
        ms = time_steps(lambda: abi.detect(d.ptr, rows, nsamp, L, dm.ptr, ds.ptr, 7, 6.0, dc.ptr, cap, dn.ptr),
                        steps, 1, None)
        out[name] = {"block": L, "detect_ms": ms, "detect_read_gbs": B / ms / 1e6, "block_stats_ms": ms_bs,
                     "block_stats_read_gbs": B / ms_bs / 1e6,
                     "candidates": int(dn.download(np.uint32, (1,))[0])}
    dst = abi.DeviceBuffer(B)
    ms = time_steps(lambda: abi.check(abi.lib().mdsm_memcpy_dtod(dst.ptr, d.ptr, B, None)), steps, 1, None)
    out["copy_ms"] = ms
    out["copy_gbs_read_plus_write"] = 2 * B / ms / 1e6
    return out


def tile_run(spec, steps):
    """One dedispersion tile shape on the single-beam survey truncated to
    2^18 samples (1024 channels x 4096 trials of 0.2), `steps` launches: the
    workload the per-tile rocprofv3 counter passes profile.  spec is
    '<f32|f16|f16p|u8>:<variant>'."""
    name, v = spec.split(":")
    dt = {"f32": abi.F32, "f16": abi.F16, "u8": abi.U8, "f16p": abi.F16_PAIRS}[name]
    s = abi.Survey.make(FCH1, -BAND / 1024, 1024, TSAMP, 0.0, 0.2, 4096, 1 << 18)
    plan = abi.Plan(s, dt)
    isz = {abi.F32: 4, abi.F16: 2, abi.U8: 1, abi.F16_PAIRS: 2}[dt]
    din = abi.DeviceBuffer(1024 * plan.min_ld * isz)
    din.zero()
    dout = abi.DeviceBuffer(4096 * (1 << 18) * 4)
    plan.set_variant(int(v))
    plan.dedisperse(din.ptr, plan.min_ld, dout.ptr)
    abi.synchronize()
    t = abi.Timer()
    t.start()
    for _ in range(steps):
        plan.dedisperse(din.ptr, plan.min_ld, dout.ptr)
    ms = t.stop() / steps
    adds = 1024 * 4096 * (1 << 18)
    return {"tile": spec, "ms": ms, "adds_per_s": adds / (ms / 1e3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--beams", type=int, default=32)
    ap.add_argument("--no-single", action="store_true", help="skip the single full-beam FP32 case")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain per-beam FP32 pass")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true", help="only the batched FP16 launches")
    ap.add_argument("--profile-pairs", action="store_true",
                    help="with --profile: the batched launches with FP16 channel-pair sums")
    ap.add_argument("--profile-u8", action="store_true",
                    help="with --profile: the batched launches with U8 storage")
    ap.add_argument("--variant", type=int, default=None,
                    help="tile variant of the batched FP16 launch instead of the autotuner's "
                         "(for profiling one tile shape)")
    ap.add_argument("--tile", default=None,
                    help="only time one tile shape ('<f32|f16|f16p|u8>:<variant>') at 1024 ch x 2^18 x 4096, "
                         "the workload of the per-tile counter profiles")
    ap.add_argument("--detect", action="store_true",
                    help="only time the boxcar detector on a 2048 x 2^20 array (both detector forms) "
                         "beside a device copy")
    ap.add_argument("--compat", action="store_true",
                    help="only measure the drop-in manager API with host buffers (PCIe included)")
    ap.add_argument("--no-pipeline", dest="pipeline", action="store_false",
                    help="skip the graph-replayed search pipeline (normalise, dedisperse, detect) on 16 beams")
    ap.add_argument("--no-precision", dest="precision", action="store_false",
                    help="skip the comparison of sampled DM rows of four beams with the oracle "
                         "(FP64 error, bit-exact same-storage rows, candidate lists)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this implementation has no multi-GPU path "
              "(the path would shard by beam, one process per GPU)", file=sys.stderr)
        sys.exit(2)
    if args.tile:
        res = tile_run(args.tile, args.steps)
        log(f"[tile] {res}")
        print(json.dumps({"tile": res}))
        return
    if args.detect:
        res = detect_run(args.steps)
        log(f"[detect] {res}")
        print(json.dumps({"detect": res}))
        return
    if args.compat:
        res = {st: compat_throughput(storage=st) for st in ("fp32", "fp16")}
        for st in ("fp32", "fp16", "fp16pairs"):
            res[st + "_series_on_device"] = compat_throughput(storage=st, download=False)
        log(f"[compat] {res}")
        print(json.dumps({"compat_host_buffers": res}))
        return
    if args.profile:
        args.no_single = args.no_plain = args.no_cpu = True
        args.precision = args.pipeline = False

    stream_p = abi.C.c_void_p()
    abi.check(abi.lib().mdsm_stream_create(abi.C.byref(stream_p)))
    stream = stream_p.value
    result = {}

    # ---------------------------------------------------------- single beam, plain FP32
    if not args.no_single:
        sb = single_beam()
        s = sb["survey"]
        plan = abi.Plan(s, abi.F32)
        ld = plan.min_ld
        log(f"[single] C=1024 N=2^20 M=4096 maxshift={plan.maxshift}")
        d_in = abi.DeviceBuffer(s.nchans * ld * 4)
        d_out = abi.DeviceBuffer(s.ndms * s.nsamp * 4)
        st = Stage(stream)
        st.fill(sb, d_in.ptr, ld, abi.F32)
        st.free()
        abi.synchronize()
        v, _ = plan.autotune(d_in.ptr, ld, d_out.ptr, stream)
        ms = time_steps(lambda: plan.dedisperse(d_in.ptr, ld, d_out.ptr, stream), args.steps,
                        args.warmup, stream)
        work = s.ndms * s.nsamp
        B = bytes_per_beam(s, plan.maxshift, 4)
        t, _ = abi.delay_table(s)
        ok, err64 = spot_check(sb, t, d_in, ld, abi.F32, d_out, [0, 1, 2047, 4095], want_error=True)
        ncand, ncl, found = detect_beam(sb, d_out, stream)
        log(f"[single] variant {v}: {ms:.2f} ms, {work / ms / 1e6:.3e} DM*samples/s, "
            f"spot check {'ok' if ok else 'FAILED'}, {ncand} cands, {ncl} clusters, {found}/20 pulses")
        result["single_fp32"] = {"ms": ms, "dm_samples_per_s": work / ms * 1e3, "variant": v,
                                 "hbm_frac": B / (ms / 1e3) / HBM_PEAK,
                                 "valu_frac": s.nchans * work / (ms / 1e3) / VALU_ADD_PEAK,
                                 "spot_check_bitexact": ok, "error_vs_fp64_rows": err64,
                                 "candidates": ncand,
                                 "clusters": ncl, "pulses_recovered": f"{found}/20"}
        del d_in, d_out, plan

    # ------------------------------------------------------------ batched beams
    beams = make_beams(nbeams=args.beams)
    surveys = [b["survey"] for b in beams]
    outs = [abi.DeviceBuffer(s.ndms * s.nsamp * 4) for s in surveys]
    total_work = sum(s.ndms * s.nsamp for s in surveys)
    total_adds = sum(s.ndms * s.nsamp * s.nchans for s in surveys)
    log(f"[batch] {len(beams)} beams, sum M*N = {total_work:.3e}, adds = {total_adds:.3e}")

    big = max(range(len(beams)), key=lambda i: surveys[i].nchans * surveys[i].ndms)
    by_adds = sorted(range(len(beams)), key=lambda i: surveys[i].nchans * surveys[i].ndms * surveys[i].nsamp)
    prec_beams = sorted({by_adds[-1], by_adds[0], by_adds[len(by_adds) // 3], by_adds[2 * len(by_adds) // 3]})

    prec_rows = {}  # (dtype, beam) -> (rows, ld, GPU output rows), with --precision

    def capture_rows(dtype, ins):
        for b in prec_beams:
            s = surveys[b]
            rows = precision_rows(beams[b])
            got = np.stack([outs[b].download(np.float32, (s.nsamp,), offset=int(d) * s.nsamp * 4)
                            for d in rows])
            prec_rows[(dtype, b)] = (rows, ins[b][1], got)

    def run_batch(dtype, tag, ins=None, v=None):
        batch = abi.Batch(surveys, dtype)
        isz = {abi.F32: 4, abi.F16: 2, abi.U8: 1, abi.F16_PAIRS: 2}[dtype]
        if ins is None:
            st = Stage(stream)
            ins = []
            for b, beam in enumerate(beams):
                ld = batch.min_ld[b]
                d = abi.DeviceBuffer(beam["survey"].nchans * ld * isz)
                st.fill(beam, d.ptr, ld, dtype)
                ins.append((d, ld))
            st.free()
            abi.synchronize()
        if v is None:  # the library's batch autotuner
            v, _ = batch.autotune([d.ptr for d, _ in ins], [l for _, l in ins], [o.ptr for o in outs],
                                  stream)
        batch.set_variant(v)
        # launch shape: one timed launch of each, keep the faster
        mode_ms = {}
        for mode in (abi.BATCH_PERSISTENT, abi.BATCH_FLAT):
            batch.set_mode(mode)
            batch.bind([d.ptr for d, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
            mode_ms[mode] = time_steps(lambda: batch.launch(stream), 1, 0, stream)
        best = min(mode_ms, key=mode_ms.get)
        batch.set_mode(best)
        batch.bind([d.ptr for d, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
        log(f"[batch {tag}] one launch: persistent {mode_ms[0]:.1f} ms, flat {mode_ms[1]:.1f} ms")
        ms = time_steps(lambda: batch.launch(stream), args.steps, args.warmup, stream)
        B = sum(bytes_per_beam(s, batch.maxshift[i], isz) for i, s in enumerate(surveys))
        log(f"[batch {tag}] variant {v}: {ms:.1f} ms/step, {total_work / ms * 1e3:.3e} DM*samples/s, "
            f"{B / (ms / 1e3) / 1e9:.1f} GB/s algorithmic")
        batch.mode_ms = {"persistent": mode_ms[0], "flat": mode_ms[1],
                         "chosen": "flat" if best == abi.BATCH_FLAT else "persistent"}
        if args.precision:
            capture_rows(dtype, ins)
        return batch, ins, ms, B, v

    # the headline storage (in --profile runs optionally the channel-pair sums)
    hdt = abi.F16
    if args.profile and args.profile_pairs:
        hdt = abi.F16_PAIRS
    elif args.profile and args.profile_u8:
        hdt = abi.U8
    htag = {abi.F16: "fp16", abi.F16_PAIRS: "fp16 pairs", abi.U8: "u8"}[hdt]
    batch16, ins16, ms16, B16, v16 = run_batch(hdt, htag, v=args.variant)
    modes16 = batch16.mode_ms
    checks = {}
    t0tab, _ = abi.delay_table(surveys[0])
    checks["beam0"] = spot_check(beams[0], t0tab, ins16[0][0], ins16[0][1], hdt, outs[0],
                                 [0, surveys[0].ndms // 2, surveys[0].ndms - 1])
    tl, _ = abi.delay_table(surveys[-1])
    checks["beam_last"] = spot_check(beams[-1], tl, ins16[-1][0], ins16[-1][1], hdt, outs[-1],
                                     [1, surveys[-1].ndms - 2])
    log(f"[batch {htag}] spot checks {checks}")
    det = {"candidates": 0, "clusters": 0, "pulses_found": 0, "pulses": 0}
    if not args.profile:
        tm = abi.Timer(stream)
        tm.start()
        for b, beam in enumerate(beams):
            n, nc, found = detect_beam(beam, outs[b], stream)
            det["candidates"] += n
            det["clusters"] += nc
            det["pulses_found"] += found
            det["pulses"] += len(beam["pulses"])
        det["ms_incl_host_clustering"] = tm.stop()
        log(f"[detect] {det}")
    for d, _ in ins16:
        d.free()
    del batch16
    if args.profile:
        # calibration for the PMC byte counters: one device-to-device copy of 1 GiB
        cal_src, cal_dst = abi.DeviceBuffer(1 << 30), abi.DeviceBuffer(1 << 30)
        abi.check(abi.lib().mdsm_memcpy_dtod(cal_dst.ptr, cal_src.ptr, 1 << 30, stream))
        abi.synchronize()
        cal_src.free()
        cal_dst.free()

    u8 = None
    if not args.profile:
        batch8, ins8, ms8, B8, v8 = run_batch(abi.U8, "u8")
        u8 = {"ms": ms8, "dm_samples_per_s": total_work / ms8 * 1e3, "tile_variant": v8,
              "launch_shape": batch8.mode_ms,
              "hbm_frac": B8 / (ms8 / 1e3) / HBM_PEAK,
              "adds_per_s": total_adds / (ms8 / 1e3)}
        for d, _ in ins8:
            d.free()
        del batch8

    # FP16 storage with channel-pair sums (one FP16 rounding per pair before
    # the FP32 accumulation): the same inputs as the FP16 run, regenerated
    pairs = None
    if not args.profile:
        batchp, insp, msp, Bp, vp = run_batch(abi.F16_PAIRS, "fp16 pairs")
        okp = spot_check(beams[0], t0tab, insp[0][0], insp[0][1], abi.F16_PAIRS, outs[0],
                         [0, surveys[0].ndms // 2, surveys[0].ndms - 1])
        pairs = {"ms": msp, "dm_samples_per_s": total_work / msp * 1e3, "tile_variant": vp,
                 "launch_shape": batchp.mode_ms, "hbm_frac": Bp / (msp / 1e3) / HBM_PEAK,
                 "adds_per_s": total_adds / (msp / 1e3), "spot_check_bitexact": okp,
                 "arithmetic": "channels 2k, 2k+1 added in FP16 (one IEEE rounding), "
                               "pair sums accumulated in FP32 in channel order"}
        log(f"[batch fp16 pairs] spot check {okp}")
        for d, _ in insp:
            d.free()
        del batchp

    plain = None
    if not args.no_plain:
        plans = [abi.Plan(s, abi.F32) for s in surveys]
        st = Stage(stream)
        ins32 = []
        for b, beam in enumerate(beams):
            ld = plans[b].min_ld
            d = abi.DeviceBuffer(surveys[b].nchans * ld * 4)
            st.fill(beam, d.ptr, ld, abi.F32)
            ins32.append((d, ld))
        st.free()
        abi.synchronize()
        # the same FP32 inputs batched into one launch (batching alone, no
        # precision change), tuned by the batch autotuner
        batch32, _, ms32b, B32b, v32b = run_batch(abi.F32, "fp32", ins=ins32)
        # the plain per-beam path, given its best configuration of two:
        # every plan with its own tile shape (the library's plan autotuner on
        # the beam truncated to 2^16 samples), or every plan with the shape
        # the batch autotuner chose for these FP32 inputs
        v32s = []
        for b, s in enumerate(surveys):
            probe = abi.Survey.make(s.fch1, s.foff, s.nchans, s.tsamp, s.lowdm, s.dmstep, s.ndms,
                                    min(s.nsamp, 1 << 16))
            pp = abi.Plan(probe, abi.F32)
            v32s.append(pp.autotune(ins32[b][0].ptr, ins32[b][1], outs[b].ptr, stream)[0])
            pp.close()

        def plain_step():
            for b, p in enumerate(plans):
                p.dedisperse(ins32[b][0].ptr, ins32[b][1], outs[b].ptr, stream)
        tried = {}
        for name, vs in (("per_beam", v32s), ("batch_shape", [v32b] * len(plans))):
            for p, v in zip(plans, vs):
                p.set_variant(v)
            tried[name] = time_steps(plain_step, args.steps, args.warmup, stream)
            log(f"[plain fp32] {name} variants {vs}: {tried[name]:.1f} ms/step")
        choice = min(tried, key=tried.get)
        ms32 = tried[choice]
        B32 = sum(bytes_per_beam(s, plans[i].maxshift, 4) for i, s in enumerate(surveys))
        plain = {"ms": ms32, "dm_samples_per_s": total_work / ms32 * 1e3, "configuration": choice,
                 "ms_per_configuration": tried, "tile_variants": v32s if choice == "per_beam" else [v32b],
                 "hbm_frac": B32 / (ms32 / 1e3) / HBM_PEAK, "launches_per_step": len(plans)}
        log(f"[plain fp32] {choice}: {ms32:.1f} ms/step, {total_work / ms32 * 1e3:.3e} DM*samples/s")
        plain["batched_fp32"] = {"ms": ms32b, "dm_samples_per_s": total_work / ms32b * 1e3,
                                 "tile_variant": v32b, "launch_shape": batch32.mode_ms,
                                 "hbm_frac": B32b / (ms32b / 1e3) / HBM_PEAK,
                                 "vs_plain_fp32": ms32 / ms32b}
        del batch32
        for d, _ in ins32:
            d.free()

    # measured device copy bandwidth (read + write bytes per second of a
    # 4 GiB device-to-device copy): the attainable HBM rate beside the 8 TB/s
    # nominal peak used for roofline.frac
    copy_gbs = None
    if not args.profile:
        nb_copy = 4 << 30
        src, dst = abi.DeviceBuffer(nb_copy), abi.DeviceBuffer(nb_copy)
        ms_copy = time_steps(lambda: abi.check(abi.lib().mdsm_memcpy_dtod(dst.ptr, src.ptr, nb_copy, stream)),
                             3, 1, stream)
        copy_gbs = 2 * nb_copy / (ms_copy / 1e3) / 1e9
        log(f"[peaks] device copy {copy_gbs:.0f} GB/s (read + write)")
        src.free()
        dst.free()

    pipeline = None
    if args.pipeline:
        # the per-chunk search sequence (normalise -> batched dedispersion ->
        # block statistics -> detection) on the first 16 beams, eager and
        # replayed from its hipGraph (the batch outputs are freed first)
        for o in outs:
            o.free()
        nb = min(16, len(beams))
        pipe = abi.Pipeline(surveys[:nb], abi.F16, stats_block=16384, nwidths=7, thresh=6.0,
                            cand_cap=1 << 21)
        raws = []
        for b in range(nb):
            sv = surveys[b]
            _, S = abi.delay_table(sv)
            ld = sv.nsamp + S
            d = abi.DeviceBuffer(sv.nchans * ld * 4)
            rng = np.random.default_rng(beams[b]["seed"])
            gain = rng.uniform(0.5, 2.0, sv.nchans).astype(np.float32)
            off = rng.uniform(0.0, 10.0, sv.nchans).astype(np.float32)
            abi.generate(beams[b]["seed"], sv, 0, ld, d.ptr, ld, gain, off, beams[b]["pulses"], stream)
            raws.append((d, ld))
        abi.synchronize()
        pipe.bind([d.ptr for d, _ in raws], [l for _, l in raws])
        pipe.set_variant(v16)
        eager = time_steps(lambda: pipe.run(stream, graph=False), args.steps, args.warmup, stream)
        pipe.run(stream, graph=True)  # records the graph
        graph = time_steps(lambda: pipe.run(stream, graph=True), args.steps, args.warmup, stream)
        ncand = sum(len(pipe.candidates(b)) for b in range(nb))
        work = sum(sv.ndms * sv.nsamp for sv in surveys[:nb])
        pipeline = {"beams": nb, "storage": "f16", "eager_ms": eager, "graph_ms": graph,
                    "dm_samples_per_s_graph": work / graph * 1e3, "candidates": ncand}
        log(f"[pipeline fp16, {nb} beams] eager {eager:.1f} ms, graph {graph:.1f} ms, {ncand} candidates")
        pipe.close()
        for d, _ in raws:
            d.free()

    cpu, cpu_ref = (None, None) if args.no_cpu else cpu_baseline()
    if cpu:
        log(f"[cpu] {cpu}")

    precision = precision_report(beams, prec_rows) if args.precision else None

    value = total_work / ms16 * 1e3
    achieved = B16 / (ms16 / 1e3)
    traffic = None
    traffic_source = None
    if hdt == abi.F16 and os.path.exists(TRAFFIC_FILE):  # measured HBM bytes per launch (rocprofv3 --pmc)
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        # counters of another tile shape do not describe this launch
        entry = tr.get("by_variant", {}).get(str(v16))
        if entry:
            traffic = entry.get("bytes_per_launch")
            traffic_source = (f"stored profile {os.path.relpath(TRAFFIC_FILE, ROOT)}: rocprofv3 --pmc "
                              f"FETCH_SIZE and WRITE_SIZE of this tile variant on this workload, "
                              "calibrated by a device copy; not measured in this run")
        else:
            log(f"no traffic profile of tile variant {v16} (profiled: {sorted(tr.get('by_variant', {}))}): "
                "roofline.traffic left null")
    line = {
        "metric": {abi.F16: "dedispersed DM x samples per second (batched, FP16 storage)",
                   abi.F16_PAIRS: "dedispersed DM x samples per second (batched, FP16 storage, channel-pair sums)",
                   abi.U8: "dedispersed DM x samples per second (batched, U8 storage)"}[hdt],
        "value": value,
        "unit": "DM*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms16,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": {abi.F16: "f16", abi.F16_PAIRS: "f16_pairs", abi.U8: "u8"}[hdt],
        "data": "synthetic (seeded Philox noise, bandpass, 20 injected pulses per beam), generated on the device",
        "config": {"workload": f"{len(beams)} uneven beams, nchans 512/1024, nsamp 2^17..2^20, "
                               "ndms 1024..4096, dmstep 0.1..0.5, one launch",
                   "beams": len(beams), "sum_dm_samples": total_work, "sum_adds": total_adds,
                   "tile_variant": v16, "launch_shape": modes16},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_source,
                     "algorithmic_bytes_per_launch": B16, "measured_copy_gbs": copy_gbs},
        "valu": {"adds_per_s": total_adds / (ms16 / 1e3), "peak_adds_per_s": VALU_ADD_PEAK,
                 "frac": total_adds / (ms16 / 1e3) / VALU_ADD_PEAK},
        # algorithmic LDS reads of the FP16 tiles: half a 4-byte word (one
        # 16-bit sample) per add
        "lds": {"achieved": 2.0 * total_adds / (ms16 / 1e3), "peak": LDS_PEAK, "unit": "B/s",
                "frac": 2.0 * total_adds / (ms16 / 1e3) / LDS_PEAK, "bytes_per_add": 2},
        "batched_u8": u8,
        "batched_fp16_pairs": pairs,
        "batched_fp16_pairs_vs_plain_fp32": (pairs["dm_samples_per_s"] / plain["dm_samples_per_s"])
        if (pairs and plain) else None,
        "batched_u8_vs_plain_fp32": (u8["dm_samples_per_s"] / plain["dm_samples_per_s"]) if (u8 and plain) else None,
        "plain_fp32_per_beam": plain,
        "batched_fp16_vs_plain_fp32": (value / plain["dm_samples_per_s"]) if plain else None,
        "single_beam_fp32": result.get("single_fp32"),
        "spot_checks_bitexact": checks,
        "precision": precision,
        "pipeline_fp16": pipeline,
        "detection": det,
        "cpu_baseline": cpu,
        "cpu_reference_1thread": cpu_ref,
    }
    line["roofline"]["achieved"] = achieved / 1e9
    line["roofline"]["peak"] = HBM_PEAK / 1e9
    print(json.dumps(line))


if __name__ == "__main__":
    main()
