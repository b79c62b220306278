"""The benchmark workloads at their real size on the GPU.

* The 32-beam uneven batch (nchans 512/1024, nsamp 2^17..2^20, ndms
  1024..4096, dmstep 0.1..0.5, 1500 MHz top over 300 MHz, 64 us; seeded
  noise, bandpass and 20 pulses per beam, generated and bandpass-normalised
  on the device) dedispersed by one batched launch with FP32, FP16, FP16
  with channel-pair sums and U8 storage, in both launch shapes (persistent
  queues and the flat grid):
  - sampled full-length DM rows of four beams equal the oracle of the same
    storage bit for bit, and the candidate lists of those rows (device
    detector) equal the oracle's;
  - FP16 and U8 rows against the FP64 sum of the unquantised normalised
    input: the error is printed for every beam checked, and FP16 (also
    with pair sums) stays within 1e-3 of the series maximum; FP32 and FP16
    candidate lists equal those of the FP64 rows;
  - sampled rows of every beam agree between the two launch shapes;
  - with a channel-constant input (x[c][t] = c mod 7) every sampled row of
    every beam is the exact channel sum, whatever the delays.
* The full single-beam search (1024 channels x 2^20 samples x 4096 DM
  trials of 0.2, 20 pulses): normalise -> FP32 dedispersion -> block
  statistics -> boxcar detection -> DBSCAN; sampled rows and their
  candidates equal the oracle's, and the injected pulses are recovered.

Device memory: the 32 beams' outputs take 196 GB and the FP32 inputs 60 GB
(one storage type's inputs are resident at a time).
"""
import sys

import numpy as np
import pytest

import bench
import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu

ISZ = {abi.F32: 4, abi.F16: 2, abi.U8: 1, abi.F16_PAIRS: 2}
NAME = {abi.F32: "f32", abi.F16: "f16", abi.U8: "u8", abi.F16_PAIRS: "f16_pairs"}
NPDT = {abi.F32: np.float32, abi.F16: np.uint16, abi.U8: np.uint8, abi.F16_PAIRS: np.uint16}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def batch32():
    beams = bench.make_beams()
    surveys = [b["survey"] for b in beams]
    outs = [abi.DeviceBuffer(s.ndms * s.nsamp * 4) for s in surveys]
    by_adds = sorted(range(len(beams)), key=lambda i: surveys[i].nchans * surveys[i].ndms * surveys[i].nsamp)
    checked = sorted({by_adds[-1], by_adds[0], by_adds[len(by_adds) // 3], by_adds[2 * len(by_adds) // 3]})
    yield beams, surveys, outs, checked
    for o in outs:
        o.free()


def _sample_rows(s):
    return sorted({0, s.ndms // 3, (2 * s.ndms) // 3, s.ndms - 1})


def _rows(out, s, rows):
    return np.stack([out.download(np.float32, (s.nsamp,), offset=int(d) * s.nsamp * 4) for d in rows])


def _oracle_rows(dtype, x, table, nsamp):
    t = np.ascontiguousarray(table)
    if dtype == abi.F32:
        return O.dedisperse_f32seq(x, t, nsamp)
    if dtype == abi.F16:
        return O.dedisperse_f16(x, t, nsamp)
    if dtype == abi.F16_PAIRS:
        return O.dedisperse_f16_pairs(x, t, nsamp)
    return O.dedisperse_u8(x, t, nsamp)


def _device_candidates(rows, L=16384, nwidths=7, thr=6.0, cap=1 << 20):
    R, N = rows.shape
    nb = (N + L - 1) // L
    dx = abi.DeviceBuffer.from_numpy(np.ascontiguousarray(rows, np.float32))
    dm, ds = abi.DeviceBuffer(R * nb * 8), abi.DeviceBuffer(R * nb * 8)
    dc, dn = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    dn.zero()
    abi.block_stats(dx.ptr, R, N, L, dm.ptr, ds.ptr)
    abi.detect(dx.ptr, R, N, L, dm.ptr, ds.ptr, nwidths, thr, dc.ptr, cap, dn.ptr)
    abi.synchronize()
    return abi.read_candidates(dc, dn, cap)


def _cand_set(d, t, k):
    return set(zip(np.asarray(d).tolist(), np.asarray(t).tolist(), np.asarray(k).tolist()))


def _fill_channel_constant(ins, surveys, dtype):
    """x[c][t] = c mod 7 in every beam's input rows (device-to-device copies
    from seven pattern rows)."""
    ldmax = max(ld for _, ld in ins)
    pat_np = {abi.F32: np.float32, abi.F16: np.float16, abi.U8: np.uint8, abi.F16_PAIRS: np.float16}[dtype]
    pat = np.repeat(np.arange(7, dtype=pat_np)[:, None], ldmax, axis=1)
    dpat = abi.DeviceBuffer.from_numpy(pat)
    isz = ISZ[dtype]
    for (d, ld), s in zip(ins, surveys):
        for c in range(s.nchans):
            abi.check(abi.lib().mdsm_memcpy_dtod(d.ptr + c * ld * isz, dpat.ptr + (c % 7) * ldmax * isz,
                                                 ld * isz, None))
    abi.synchronize()
    dpat.free()


def _channel_constant_expect(dtype, C):
    v = np.arange(C) % 7
    if dtype == abi.U8:
        return np.float32(float(v.sum()) * (8.0 / 255.0) - 4.0 * C)
    return np.float32(v.sum())


@pytest.mark.parametrize("dtype,variant", [(abi.F32, None), (abi.F16, None), (abi.U8, None),
                                           (abi.F16_PAIRS, None), (abi.F16_PAIRS, 15), (abi.U8, 2)],
                         ids=["f32", "f16", "u8", "f16_pairs", "f16_pairs_packed_tile", "u8_packed_tile"])
def test_batch_32_beams_full_size(batch32, dtype, variant):
    """The 32-beam batch in each storage type with its untuned shape (for
    the FP16 channel-pair sums and U8 a pair-sum tile shape, which takes
    every beam of this batch), and the FP16 channel-pair sums and U8 also
    with the shapes that take every survey (15 and 2)."""
    beams, surveys, outs, checked = batch32
    batch = abi.Batch(surveys, dtype)
    if variant is not None:
        batch.set_variant(variant)
    isz = ISZ[dtype]
    st = bench.Stage(None)
    ins = []
    for b, beam in enumerate(beams):
        ld = batch.min_ld[b]
        d = abi.DeviceBuffer(surveys[b].nchans * ld * isz)
        st.fill(beam, d.ptr, ld, dtype)
        ins.append((d, ld))
    abi.synchronize()
    args = ([d.ptr for d, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
    log(f"[32-beam batch {NAME[dtype]}] {len(beams)} beams generated and normalised on the device, "
        f"tile shape {variant or 'untuned'}")

    sampled = {}
    for mode in (abi.BATCH_PERSISTENT, abi.BATCH_FLAT):
        for o in outs:
            o.zero()
        batch.set_mode(mode)
        log(f"[32-beam batch {NAME[dtype]}] launch mode {mode}")
        batch.dedisperse(*args)
        abi.synchronize()
        sampled[mode] = [_rows(outs[b], s, _sample_rows(s)) for b, s in enumerate(surveys)]
    for b in range(len(surveys)):
        assert np.array_equal(sampled[abi.BATCH_PERSISTENT][b], sampled[abi.BATCH_FLAT][b]), b

    # oracle rows of four beams (flat launch outputs), their candidates, FP64 error
    for b in checked:
        s = surveys[b]
        ld = ins[b][1]
        rows = bench.precision_rows(beams[b])
        table, _ = abi.delay_table(s)
        x = ins[b][0].download(NPDT[dtype], (s.nchans, ld))
        got = _rows(outs[b], s, rows)
        ref = _oracle_rows(dtype, x, table[:, rows], s.nsamp)
        assert np.array_equal(got, ref), (NAME[dtype], b, rows)
        del x
        # candidates of these rows: device detector == oracle detector
        c = _device_candidates(got)
        od, ot, ok, osnr, on = O.detect(ref.astype(np.float64), 16384, 7, 6.0)
        assert len(c) == on and _cand_set(c["dm_idx"], c["sample"], c["width_log2"]) == _cand_set(od, ot, ok)
        assert np.array_equal(c["snr"], osnr)
        # against the FP64 sum of the unquantised normalised input
        if dtype == abi.F32:
            x32 = ins[b][0].download(np.float32, (s.nchans, ld))
        else:
            tmp = abi.DeviceBuffer(s.nchans * ld * 4)
            st.fill(beams[b], tmp.ptr, ld, abi.F32)
            abi.synchronize()
            x32 = tmp.download(np.float32, (s.nchans, ld))
            tmp.free()
        y64 = O.dedisperse_f64(x32, np.ascontiguousarray(table[:, rows]), s.nsamp)
        del x32
        err = np.abs(got.astype(np.float64) - y64)
        e_max = float(err.max() / np.abs(y64).max())
        e_rms = float(err.max() / np.sqrt((y64 * y64).mean()))
        fd, ft, fk, _, _ = O.detect(y64, 16384, 7, 6.0)
        ndiff = len(_cand_set(od, ot, ok) ^ _cand_set(fd, ft, fk))
        log(f"[32-beam batch {NAME[dtype]}] beam {b} (C={s.nchans} N={s.nsamp} M={s.ndms} step={s.dmstep:.3f}) "
            f"rows {rows}: bit-exact vs oracle, {on} candidates; vs FP64: max err {e_max:.2e} of max, "
            f"{e_rms:.2e} of rms; candidates differing from the FP64 rows' {ndiff}")
        if dtype in (abi.F16, abi.F16_PAIRS):
            assert e_max <= 1e-3, e_max
        if dtype in (abi.F32, abi.F16):
            assert ndiff == 0
        if dtype == abi.U8:
            assert e_max <= 3e-2, e_max

    # channel-constant known answer on every beam (flat launch)
    _fill_channel_constant(ins, surveys, dtype)
    for o in outs:
        o.zero()
    batch.dedisperse(*args)
    abi.synchronize()
    for b, s in enumerate(surveys):
        got = _rows(outs[b], s, _sample_rows(s))
        assert (got == _channel_constant_expect(dtype, s.nchans)).all(), (NAME[dtype], b)
    log(f"[32-beam batch {NAME[dtype]}] channel-constant known answer on {len(surveys)} beams")
    st.free()
    for d, _ in ins:
        d.free()
    batch.close()


def test_single_beam_full_search():
    sb = bench.single_beam()
    s = sb["survey"]
    M, N = s.ndms, s.nsamp
    plan = abi.Plan(s, abi.F32)
    ld = plan.min_ld
    log(f"[single beam] C=1024 N=2^20 M=4096 maxshift={plan.maxshift}")
    d_in = abi.DeviceBuffer(s.nchans * ld * 4)
    d_out = abi.DeviceBuffer(M * N * 4)
    st = bench.Stage(None)
    st.fill(sb, d_in.ptr, ld, abi.F32)
    st.free()
    plan.dedisperse(d_in.ptr, ld, d_out.ptr)
    L, thr, cap = 16384, 6.0, 1 << 22
    nb = (N + L - 1) // L
    dm, ds = abi.DeviceBuffer(M * nb * 8), abi.DeviceBuffer(M * nb * 8)
    dc, dn = abi.DeviceBuffer(cap * 16), abi.DeviceBuffer(4)
    dn.zero()
    abi.block_stats(d_out.ptr, M, N, L, dm.ptr, ds.ptr)
    abi.detect(d_out.ptr, M, N, L, dm.ptr, ds.ptr, 7, thr, dc.ptr, cap, dn.ptr)
    abi.synchronize()
    c = abi.read_candidates(dc, dn, cap)
    assert len(c) < cap
    lab, nc = abi.dbscan2d(c, 3, 4)
    peaks, _ = abi.cluster_peaks(c, lab, nc)
    log(f"[single beam] {len(c)} candidates, {nc} clusters")

    # sampled rows (first, last, the rows of every fifth pulse) vs the oracle
    table, _ = abi.delay_table(s)
    rows = sorted({0, M - 1} | {min(M - 1, int(round(p.dm / s.dmstep))) for p in sb["pulses"][::5]})
    x = d_in.download(np.float32, (s.nchans, ld))
    ref = O.dedisperse_f32seq(x, np.ascontiguousarray(table[:, rows]), N)
    del x
    got = _rows(d_out, s, rows)
    assert np.array_equal(got, ref)
    od, ot, ok, osnr, on = O.detect(ref.astype(np.float64), L, 7, thr)
    sel = np.isin(c["dm_idx"], rows)
    cs = c[sel]
    remap = {r: i for i, r in enumerate(rows)}
    assert _cand_set([remap[int(v)] for v in cs["dm_idx"]], cs["sample"], cs["width_log2"]) == _cand_set(od, ot, ok)

    # recall of the injected pulses: a cluster peak within 2% in DM and
    # max(2 widths, 4 samples) in time
    missed = []
    for p in sb["pulses"]:
        di = p.dm / s.dmstep
        ok_ = [q for q in peaks if abs(q["dm_idx"] - di) <= max(2.0, 0.02 * di)
               and abs(int(q["sample"]) - p.start) <= max(2 * p.width, 4)]
        if not ok_:
            near = sorted(peaks, key=lambda q: abs(q["dm_idx"] - di) / max(2.0, 0.02 * di)
                          + abs(int(q["sample"]) - p.start) / max(2 * p.width, 4))[:1]
            missed.append((p.dm, p.start, p.width, p.amp * np.sqrt(1024 * p.width),
                           [(int(q["dm_idx"]), int(q["sample"]), float(q["snr"])) for q in near]))
    for m in missed:
        log(f"[single beam] missed pulse DM {m[0]:.2f} start {m[1]} width {m[2]} injected S/N {m[3]:.1f}; "
            f"nearest peak (dm_idx, sample, snr) {m[4]}")
    log(f"[single beam] recovered {20 - len(missed)}/20 pulses")
    assert len(missed) <= 1
    plan.close()
