"""GPU parity of normalisation, statistics, detection, filters, generator and
the end-to-end single-beam search against the CPU oracle.

Bars: per-channel and per-block FP64 statistics, normalised F32/F16/U8
arrays and candidate lists (dm_idx, sample, width) are bit-exact against the
oracle fed the same inputs (the kernels use the oracle's reduction order and
FP64 arithmetic without contraction).  Against the FP64 end-to-end oracle the
clustered candidate peaks must be identical.
"""
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _cands_tuple(c):
    return list(zip(c["dm_idx"].tolist(), c["sample"].tolist(), c["width_log2"].tolist()))


def test_channel_stats_and_normalise_bitexact():
    C, n, ld = 64, 50001, 50011
    g, o = O.channel_bandpass(3, C)
    raw = O.generate(3, C, 1500.0, -0.29296875, 64e-6, 0, n, g, o, ld=ld)
    _say(f"[normalise] C={C} n={n}")
    d_raw = abi.DeviceBuffer.from_numpy(raw)
    d_m, d_s = abi.DeviceBuffer(C * 8), abi.DeviceBuffer(C * 8)
    abi.channel_stats(d_raw.ptr, ld, C, n, d_m.ptr, d_s.ptr)
    m, s = d_m.download(np.float64, (C,)), d_s.download(np.float64, (C,))
    om, os_ = O.channel_stats(raw, n)
    assert np.array_equal(m, om) and np.array_equal(s, os_)
    for dt, name, npdt in [(abi.F32, "f32", np.float32), (abi.F16, "f16", np.uint16), (abi.U8, "u8", np.uint8)]:
        ldo = n + 7
        d_o = abi.DeviceBuffer(C * ldo * np.dtype(npdt).itemsize)
        abi.normalise(d_raw.ptr, ld, C, n, d_m.ptr, d_s.ptr, d_o.ptr, dt, ldo)
        got = d_o.download(npdt, (C, ldo))[:, :n]
        assert np.array_equal(got, O.normalise(raw, om, os_, name, n)), name


def _dedispersed(seed, C, M, N, pulses=()):
    s = abi.Survey.make(418.0, -0.078125 * 256 / C, C, 51.2e-6, 0.0, 0.1, M, N)
    t, S = abi.delay_table(s)
    x = O.generate(seed, C, s.fch1, s.foff, s.tsamp, 0, N + S, pulses=pulses)
    return s, t, x, abi.dedisperse_host(s, x)


@pytest.mark.parametrize("L", [16384, 4096, 1000])
def test_block_stats_and_detect_bitexact(L):
    s, t, x, y = _dedispersed(11, 128, 96, 40000, pulses=[(3.0, 5000, 16, 0.6), (7.7, 20000, 3, 1.5)])
    M, N = 96, 40000
    _say(f"[detect] M={M} N={N} L={L}")
    d_y = abi.DeviceBuffer.from_numpy(y)
    nb = (N + L - 1) // L
    d_m, d_s = abi.DeviceBuffer(M * nb * 8), abi.DeviceBuffer(M * nb * 8)
    abi.block_stats(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr)
    m, sd = d_m.download(np.float64, (M, nb)), d_s.download(np.float64, (M, nb))
    om, os_ = O.block_stats(y.astype(np.float64), L)
    assert np.array_equal(m, om) and np.array_equal(sd, os_)
    cap = 1 << 16
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    for thr in (6.0, 4.5):
        abi.detect(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr, 7, thr, d_c.ptr, cap, d_n.ptr)
        c = abi.read_candidates(d_c, d_n, cap)
        od, ot, ok, osnr, on = O.detect(y.astype(np.float64), L, 7, thr, stats=(om, os_))
        assert len(c) == on > 0
        assert _cands_tuple(c) == list(zip(od.tolist(), ot.tolist(), ok.tolist()))
        assert np.array_equal(c["snr"], osnr)


@pytest.mark.parametrize("N,L,nw", [(40000, 2048, 4), (913, 16, 7), (100, 64, 7), (1, 16, 3),
                                     (5001, 5000, 7), (70000, 8192, 1)])
def test_detect_row_tails_and_widths(N, L, nw):
    """Rows whose length is not a multiple of the detector's 912-sample wave
    segment or 16-sample runs, shorter than a run or a block, fewer boxcar
    widths, and block lengths on both detector paths (multiples of 16 use
    the register-resident kernel, others the LDS-tiled one)."""
    rng = np.random.default_rng(N + L + nw)
    M = 7
    y = rng.standard_normal((M, N)).astype(np.float32)
    y[:, N // 2: N // 2 + 8] += 3.0                      # a boxcar for the wider levels
    y[:, -1] += 9.0                                       # the row's last sample
    d_y = abi.DeviceBuffer.from_numpy(y)
    nb = (N + L - 1) // L
    d_m, d_s = abi.DeviceBuffer(M * nb * 8), abi.DeviceBuffer(M * nb * 8)
    abi.block_stats(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr)
    om, os_ = O.block_stats(y.astype(np.float64), L)
    cap = 1 << 16
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    for thr in (2.5, 4.0):
        abi.detect(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr, nw, thr, d_c.ptr, cap, d_n.ptr)
        c = abi.read_candidates(d_c, d_n, cap)
        od, ot, ok, osnr, on = O.detect(y.astype(np.float64), L, nw, thr, stats=(om, os_))
        assert len(c) == on
        assert _cands_tuple(c) == list(zip(od.tolist(), ot.tolist(), ok.tolist()))
        assert np.array_equal(c["snr"], osnr)


@pytest.mark.parametrize("L", [16, 15])
def test_detect_threshold_boundary(L):
    """S/N at the threshold to within rounding: one planted sample x per
    block, every other sample 0, and caller block statistics with the block
    mean m = x - thr * sd moved by -2..+2 ulps, so the exact test
    (x - m) / sd >= thr falls on both sides.  Candidates and S/N equal the
    oracle's on both detector forms (L = 16: register-resident; 15: LDS)."""
    rng = np.random.default_rng(L)
    M, nblocks, thr = 5, 96, 4.0
    N = L * nblocks
    y = np.zeros((M, N), np.float32)
    m = np.zeros((M, nblocks))
    sd = np.zeros((M, nblocks))
    for r in range(M):
        for b in range(nblocks):
            x = np.float32(rng.uniform(1.0, 3.0))
            s = float(rng.uniform(0.05, 0.2))
            y[r, b * L + int(rng.integers(0, L))] = x
            mb = float(x) - thr * s
            for _ in range(abs((r + b) % 5 - 2)):
                mb = np.nextafter(mb, np.inf if (r + b) % 5 > 2 else -np.inf)
            m[r, b], sd[r, b] = mb, s
    d_y = abi.DeviceBuffer.from_numpy(y)
    d_m, d_s = abi.DeviceBuffer.from_numpy(m), abi.DeviceBuffer.from_numpy(sd)
    cap = 1 << 14
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    abi.detect(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr, 7, thr, d_c.ptr, cap, d_n.ptr)
    c = abi.read_candidates(d_c, d_n, cap)
    od, ot, ok, osnr, on = O.detect(y.astype(np.float64), L, 7, thr, stats=(m, sd))
    _say(f"[detect boundary L={L}] {on} of {M * nblocks} planted samples at or above {thr}")
    assert 0 < on < M * nblocks
    assert len(c) == on
    assert _cands_tuple(c) == list(zip(od.tolist(), ot.tolist(), ok.tolist()))
    assert np.array_equal(c["snr"], osnr)


def test_detect_overflow_is_reported():
    s, t, x, y = _dedispersed(12, 64, 32, 8192)
    d_y = abi.DeviceBuffer.from_numpy(y)
    d_m, d_s = abi.DeviceBuffer(32 * 8), abi.DeviceBuffer(32 * 8)
    abi.block_stats(d_y.ptr, 32, 8192, 8192, d_m.ptr, d_s.ptr)
    d_c, d_n = abi.DeviceBuffer(16 * 16), abi.DeviceBuffer(4)
    abi.detect(d_y.ptr, 32, 8192, 8192, d_m.ptr, d_s.ptr, 7, 1.0, d_c.ptr, 16, d_n.ptr)
    with pytest.raises(abi.MdsmError):
        abi.read_candidates(d_c, d_n, 16)


def _gpu_search(s, x, L=16384, thr=6.0):
    y = abi.dedisperse_host(s, x)
    M, N = s.ndms, s.nsamp
    nb = (N + L - 1) // L
    d_y = abi.DeviceBuffer.from_numpy(y)
    d_m, d_s = abi.DeviceBuffer(M * nb * 8), abi.DeviceBuffer(M * nb * 8)
    abi.block_stats(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr)
    cap = 1 << 20
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    abi.detect(d_y.ptr, M, N, L, d_m.ptr, d_s.ptr, 7, thr, d_c.ptr, cap, d_n.ptr)
    c = abi.read_candidates(d_c, d_n, cap)
    lab, nc = abi.dbscan2d(c, 3, 4)
    peaks, sizes = abi.cluster_peaks(c, lab, nc)
    return y, c, peaks


def test_256ch_pipeline_matches_fp64_oracle():
    """Single beam, C=256, N=65536, M=256 (DM 0..25.5 step 0.1), N(0,1) noise
    and three width-8 pulses of +0.5 per channel at DM 5.0, 12.3, 20.7."""
    s = abi.Survey.make(418.0, -0.078125, 256, 51.2e-6, 0.0, 0.1, 256, 65536)
    t, S = abi.delay_table(s)
    pulses = [(5.0, 9000, 8, 0.5), (12.3, 30000, 8, 0.5), (20.7, 52000, 8, 0.5)]
    x = O.generate(2024, 256, s.fch1, s.foff, s.tsamp, 0, 65536 + S, pulses=pulses)
    _say("[256-ch pipeline] GPU search")
    y, c, peaks = _gpu_search(s, x)
    _say("[256-ch pipeline] FP64 oracle search")
    y64 = O.dedisperse_f64(x, t, 65536)
    od, ot, ok, osnr, on = O.detect(y64, 16384, 7, 6.0)
    assert _cands_tuple(c) == list(zip(od.tolist(), ot.tolist(), ok.tolist()))
    olab, onc = O.dbscan2d(od, ot, 3, 4)
    assert len(peaks) == onc
    for k in range(onc):
        m = olab == k
        i = np.flatnonzero(m)[np.argmax(osnr[m])]
        assert (peaks[k]["dm_idx"], peaks[k]["sample"], peaks[k]["width_log2"]) == (od[i], ot[i], ok[i])
    # every injected pulse is the peak of a cluster at its DM, start and width
    for dm, t0, w, _ in pulses:
        hit = [p for p in peaks if abs(p["dm_idx"] - round(dm / 0.1)) <= 1 and abs(p["sample"] - t0) <= 2]
        assert hit and hit[0]["width_log2"] == 3 and hit[0]["snr"] > 15, (dm, hit)


@pytest.mark.parametrize("case", ["contiguous", "multi_chunk", "strided_prefix"])
def test_detect_global_reference_semantics(case):
    """Global FP64 mean/rms and the x - mean >= rms*thr threshold: statistics
    and candidates (row, sample, snr) bit-exact against the oracle's
    restatement of the device order."""
    if case == "contiguous":
        s, t, x, y = _dedispersed(13, 64, 48, 20000, pulses=[(2.0, 3000, 4, 1.0)])
        nsamp, ld, sc, thr = 20000, 20000, 20000, 4.0
    else:
        rng = np.random.default_rng(21)
        ncols = 150000 if case == "multi_chunk" else 70001
        y = (rng.standard_normal((5, ncols)) * 3 + 2).astype(np.float32)
        y[3, 1234] += 40.0
        y[4, ncols - 1] += 40.0  # past the thresholded columns in the strided case
        ld = ncols
        nsamp = ncols if case == "multi_chunk" else 69000
        sc = nsamp if case == "multi_chunk" else 66001
        thr = 4.5
    nrows = y.shape[0]
    _say(f"[detect_global] case={case} rows={nrows} nsamp={nsamp} ld={ld} stats={sc}")
    d_y = abi.DeviceBuffer.from_numpy(y)
    cap = 1 << 16
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    mean, rms = abi.detect_global(d_y.ptr, nrows, nsamp, thr, d_c.ptr, cap, d_n.ptr, ld=ld,
                                  stats_nsamp=sc)
    om, orms = O.global_stats(y, sc) if ld == nsamp else O.global_stats(y[:, :nsamp], sc)
    assert (mean, rms) == (om, orms)
    c = abi.read_candidates(d_c, d_n, cap)
    od, ot, osnr, on = O.threshold_ref(np.ascontiguousarray(y[:, :nsamp]), om, orms, thr)
    assert len(c) == on > 0
    assert list(zip(c["dm_idx"], c["sample"])) == list(zip(od, ot))
    assert np.array_equal(c["snr"], osnr)
    # the same threshold at caller-given statistics
    fm, fr = om + 0.25 * orms, 1.125 * orms
    abi.threshold_global(d_y.ptr, nrows, nsamp, fm, fr, thr, d_c.ptr, cap, d_n.ptr, ld=ld)
    c = abi.read_candidates(d_c, d_n, cap)
    od, ot, osnr, on = O.threshold_ref(np.ascontiguousarray(y[:, :nsamp]), fm, fr, thr)
    assert on > 0
    assert len(c) == on and list(zip(c["dm_idx"], c["sample"])) == list(zip(od, ot))
    assert np.array_equal(c["snr"], osnr)


def _legacy_gpu(x, lowdm, dmstep, thresh, mean=None, rms=None, spc=0, offset=0, nvals=0):
    ndms, nsamp = x.shape
    # offset shifts the device copy by whole floats (16-byte misalignment)
    d_x = abi.DeviceBuffer.from_numpy(np.concatenate([np.zeros(offset, np.float32), x.ravel()]))
    cap = 1 << 18
    d_c, d_n = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    d_m = abi.DeviceBuffer.from_numpy(mean) if mean is not None else None
    d_r = abi.DeviceBuffer.from_numpy(rms) if rms is not None else None
    st = abi.detect_legacy(d_x.ptr + 4 * offset, ndms, nsamp, lowdm, dmstep, thresh, d_c.ptr, cap,
                           d_n.ptr, d_m.ptr if d_m else None, d_r.ptr if d_r else None, spc,
                           0 if mean is None else mean.size, nvals=nvals)
    return st, abi.read_candidates(d_c, d_n, cap)


@pytest.mark.parametrize("case", ["pulse", "odd_small", "quirk", "misaligned", "short_rows", "prefix"])
def test_detect_legacy_bitexact(case):
    """Global-statistics branch of the legacy detector: mean and rms equal
    the oracle's FP32 values exactly, candidates (row, sample, snr) too."""
    rng = np.random.default_rng(21)
    off = 0
    if case == "pulse":
        _, _, _, x = _dedispersed(13, 64, 48, 20000, pulses=[(2.0, 3000, 4, 1.0)])
        args = (0.0, 0.1, 3.0)
    elif case == "odd_small":
        x = rng.standard_normal((3, 1001)).astype(np.float32)
        args = (0.3, 0.35, 1.5)
    elif case == "quirk":
        x = np.zeros((3, 32768), np.float32)
        x[0], x[1], x[2] = 1.0, 3.0, 100.0
        args = (0.0, 2.0, 6.0)
    elif case == "misaligned":
        x = (rng.standard_normal((9, 32771)) * 3 + 1).astype(np.float32)
        args = (0.5, 0.25, 2.5)
        off = 1
    elif case == "short_rows":
        x = rng.standard_normal((4, 5)).astype(np.float32) * 4
        args = (1.5, 1.0, 0.0)
    else:
        # statistics over a prefix of the array (a short last chunk:
        # read_nsamp * ndms values), threshold over every row
        x = (rng.standard_normal((5, 40000)) + np.arange(5)[:, None]).astype(np.float32)
        args = (0.0, 0.5, 2.0)
        nv = 3 * 40000 + 777
    nv = nv if case == "prefix" else 0
    _say("legacy", case, x.shape, nv)
    (m, r), c = _legacy_gpu(x, *args, offset=off, nvals=nv)
    od, ot, osnr, on, om, orms = O.process_brute_legacy(x, *args, nvals=nv)
    assert (m, r) == (om, orms), (m, r, om, orms)
    assert len(c) == on
    assert c["dm_idx"].tolist() == od.tolist() and c["sample"].tolist() == ot.tolist()
    assert np.array_equal(c["snr"], osnr)
    if case == "quirk":
        assert (m, r) == (2.0, 1.0) and on > 0


def test_detect_legacy_chunk_statistics_bitexact():
    rng = np.random.default_rng(22)
    ndms, nsamp, spc = 40, 12345, 1000
    x = rng.standard_normal((ndms, nsamp)).astype(np.float32)
    x[:, 5000:5010] += 6.0
    nstats = (nsamp - 5) // spc + 1
    mean = (rng.random(nstats) * 0.4 - 0.2).astype(np.float32)
    rms = (0.8 + rng.random(nstats) * 0.4).astype(np.float32)
    st, c = _legacy_gpu(x, 0.2, 0.05, 2.0, mean, rms, spc)
    assert st is None
    od, ot, osnr, on = O.process_brute_legacy_chunks(x, 0.2, 0.05, 2.0, mean, rms, spc)
    assert on > 100 and len(c) == on
    assert c["dm_idx"].tolist() == od.tolist() and c["sample"].tolist() == ot.tolist()
    assert np.array_equal(c["snr"], osnr)


def test_detect_legacy_argument_errors():
    x = np.zeros((4, 100), np.float32)
    d_x = abi.DeviceBuffer.from_numpy(x)
    d_c, d_n = abi.DeviceBuffer(16 * 16), abi.DeviceBuffer(4)
    d_m = abi.DeviceBuffer.from_numpy(np.zeros(2, np.float32))
    with pytest.raises(abi.MdsmError):   # mean without rms
        abi.detect_legacy(d_x.ptr, 4, 100, 0.0, 1.0, 3.0, d_c.ptr, 16, d_n.ptr, d_m.ptr, None, 50, 2)
    with pytest.raises(abi.MdsmError):   # (100 - 5) // 40 + 1 = 3 entries needed
        abi.detect_legacy(d_x.ptr, 4, 100, 0.0, 1.0, 3.0, d_c.ptr, 16, d_n.ptr, d_m.ptr, d_m.ptr,
                          40, 2)
    with pytest.raises(abi.MdsmError):   # samples_per_chunk 0
        abi.detect_legacy(d_x.ptr, 4, 100, 0.0, 1.0, 3.0, d_c.ptr, 16, d_n.ptr, d_m.ptr, d_m.ptr,
                          0, 2)
    with pytest.raises(abi.MdsmError):   # the value count must fit 32 bits
        abi.detect_legacy(d_x.ptr, 65535, 65538, 0.0, 1.0, 3.0, d_c.ptr, 16, d_n.ptr)


def test_median_and_detrend():
    rng = np.random.default_rng(14)
    x = rng.standard_normal((7, 3001)).astype(np.float32)
    d_x, d_o = abi.DeviceBuffer.from_numpy(x), abi.DeviceBuffer(x.nbytes)
    abi.median_filter5(d_x.ptr, 7, 3001, d_o.ptr)
    assert np.array_equal(d_o.download(np.float32, x.shape), O.median_filter5(x))
    L = 4096
    z = (0.002 * np.arange(3 * L) + 5 + rng.standard_normal(3 * L)).astype(np.float32)
    d_z = abi.DeviceBuffer.from_numpy(z)
    abi.detrend_normalise(d_z.ptr, 3, L)
    assert np.allclose(d_z.download(np.float32, z.shape), O.detrend_normalise(z, L), atol=2e-6)


def test_device_generator_matches_oracle():
    C, n, t0 = 48, 10003, 777
    s = abi.Survey.make(1500.0, -0.29296875 * 1024 / C, C, 64e-6, 0.0, 1.0, 4, 1024)
    g, o = O.channel_bandpass(9, C)
    pulses = [abi.Pulse(300.0, 2000, 12, 1.25), abi.Pulse(55.5, 9000, 1, 3.0)]
    d_raw = abi.DeviceBuffer(C * n * 4)
    abi.generate(9, s, t0, n, d_raw.ptr, n, g, o, pulses)
    got = d_raw.download(np.float32, (C, n))
    ref = O.generate(9, C, s.fch1, s.foff, s.tsamp, t0, n, g, o,
                     pulses=[(p.dm, p.start, p.width, p.amp) for p in pulses])
    # Box-Muller from IEEE operations only on both sides: identical bits
    assert np.array_equal(got, ref), np.abs(got - ref).max()
