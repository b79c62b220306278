"""The MDSM multibeam manager API (initialiseMDSM / get_buffer_pointer /
next_chunk_multibeam / start_processing / tearDownMDSM) driven the way
src/MDSM.cpp drives it: a priming read of maxshift samples, then chunks of
nsamp.  Checks, against the oracle: the overlap-save stream semantics (every
returned series equals the one-shot dedispersion of the concatenated stream,
bit for bit), the two-call output latency, the reference delay convention,
and the candidate CSV file."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu

lib = abi.lib()
lib.mdsm_compat_open.argtypes = [C.c_char_p, C.c_uint]
lib.mdsm_compat_open.restype = C.c_int
lib.mdsm_compat_survey.argtypes = [C.POINTER(C.c_uint)] * 5
lib.mdsm_compat_buffer.restype = C.POINTER(C.c_float)
lib.mdsm_compat_next_chunk.argtypes = [C.c_uint, C.POINTER(C.c_uint), C.c_double, C.c_double]
lib.mdsm_compat_next_chunk.restype = C.POINTER(C.POINTER(C.c_float))
lib.mdsm_compat_start.argtypes = [C.c_uint]
lib.mdsm_compat_last_error.restype = C.c_char_p
lib.mdsm_compat_set_last_nsamp.argtypes = [C.c_uint]
lib.mdsm_compat_storage.restype = C.c_int

XML = """<observation>
  <dm lowDM="0" numDMs="40" dmStep="1.0" />
  <channels nchans="64" npols="1" />
  <timing tsamp="0.000064" />
  <samples nsamp="4096" nbits="32" />
  <detection detectionThreshold="6" />
  <clustering applyClustering="1" minPoints="2" timeRange="8" dmRange="3" snrRange="100" />
  <writer filePrefix="compat" baseDirectory="{dir}" singleFileMode="1" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-1.0" /></beams>
</observation>"""


def test_streaming_manager_matches_one_shot(tmp_path):
    xml = tmp_path / "obs.xml"
    xml.write_text(XML.format(dir=str(tmp_path)))
    assert lib.mdsm_compat_open(str(xml).encode(), 0) == 0, lib.mdsm_compat_last_error()
    nb, nch, ns, nd, ms = (C.c_uint() for _ in range(5))
    lib.mdsm_compat_survey(C.byref(nb), C.byref(nch), C.byref(ns), C.byref(nd), C.byref(ms))
    C_, N, M, S = nch.value, ns.value, nd.value, ms.value
    print(f"[compat] C={C_} N={N} M={M} maxshift={S}", file=sys.stderr, flush=True)
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    assert S >= table.max() and S % 256 == 0
    nchunks = 5
    total = S + nchunks * N
    stream = O.generate(77, C_, 1500.0, -1.0, 64e-6, 0, total,
                        pulses=[(20.0, 6000, 8, 1.0), (31.0, 15000, 4, 1.5)])
    # one-shot reference over the whole stream
    full = O.dedisperse_f32seq(stream, table, total - S)
    got, reads = {}, []
    for i in range(nchunks + 3):
        buf = lib.mdsm_compat_buffer()
        if i == 0:
            n = S
            view = np.ctypeslib.as_array(buf, shape=(C_, S))
            view[:] = stream[:, :S]
        elif i <= nchunks:
            n = N
            view = np.ctypeslib.as_array(buf, shape=(C_, N))
            view[:] = stream[:, S + (i - 1) * N:S + i * N]
        else:
            n = 0
        samples = C.c_uint()
        reads.append(n)
        outs = lib.mdsm_compat_next_chunk(n, C.byref(samples), 64e-6 * i * N, 64e-6)
        if i < 1:
            assert not outs and samples.value == 0xFFFFFFFF
        if i >= 3 and i - 2 <= nchunks:
            arr = np.ctypeslib.as_array(outs[0], shape=(M, N)).copy()
            # the reference reports the previous call's data_read with the
            # buffer of the chunk two calls back (0 for the last chunk)
            assert samples.value == reads[i - 1]
            got[i - 2] = (arr, samples.value)
        r = lib.mdsm_compat_start(n)
        if r == 0:
            break
    lib.mdsm_compat_close()
    assert sorted(got) == list(range(1, nchunks + 1))
    assert [got[k][1] for k in sorted(got)] == [N] * (nchunks - 1) + [0]
    for k, (arr, samples) in got.items():
        assert np.array_equal(arr, full[:, (k - 1) * N:k * N]), k
    lines = (tmp_path / "compat.dat").read_text().strip().splitlines()
    assert lines, "no candidates written"
    rows = [tuple(float(v) for v in ln.split(",")) for ln in lines]
    assert all(len(r) == 5 for r in rows)
    # both pulses appear as clusters near their DM
    dms = {round(r[1]) for r in rows if r[3] > 0}
    assert any(abs(d - 20) <= 2 for d in dms) and any(abs(d - 31) <= 2 for d in dms), dms


XML_RFI = """<observation>
  <dm lowDM="0" numDMs="24" dmStep="1.0" />
  <channels nchans="64" npols="1" ncoeffs="6" />
  <timing tsamp="0.000064" />
  <samples nsamp="4096" nbits="32" />
  <rfi applyRFIClipper="1" channelBlock="1024" spectrumThreshold="4" channelThreshold="10" channelMask="5,20-22"/>
  <detection detectionThreshold="6" applyMedianFilter="1" applyDetrending="1" />
  <writer filePrefix="rfi" baseDirectory="{dir}" singleFileMode="1" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-1.0" /></beams>
</observation>"""


def _clean(block, masks):
    """The oracle's bandpass fit + spectrum clipping of one block (the
    reference clears the channel flags before clipping them)."""
    fitted, st = O.bandpass_fit(block, 0, block.shape[1], 6, masks, 10.0, 4.0)
    out, k = O.spectrum_clipper(block, 0, block.shape[1], fitted, st["spectrum_hi"], st["spectrum_lo"])
    return out, k


def test_streaming_manager_rfi_median_detrend(tmp_path):
    xml = tmp_path / "obs.xml"
    xml.write_text(XML_RFI.format(dir=str(tmp_path)))
    assert lib.mdsm_compat_open(str(xml).encode(), 0) == 0, lib.mdsm_compat_last_error()
    nb, nch, ns, nd, ms = (C.c_uint() for _ in range(5))
    lib.mdsm_compat_survey(C.byref(nb), C.byref(nch), C.byref(ns), C.byref(nd), C.byref(ms))
    C_, N, M, S = nch.value, ns.value, nd.value, ms.value
    print(f"[compat rfi] C={C_} N={N} M={M} maxshift={S}", file=sys.stderr, flush=True)
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    nchunks = 3
    total = S + nchunks * N
    gain, offset = O.channel_bandpass(5, C_)  # a bandpass to fit: gains 0.5-2, offsets 0-10
    stream = O.generate(5, C_, 1500.0, -1.0, 64e-6, 0, total, gain=gain, offset=offset,
                        pulses=[(12.0, 5000, 8, 2.0)])
    stream[:, S + 1000] += 40.0  # broadband RFI in chunk 1
    stream[:, S + N + 300:S + N + 303] += 60.0  # and in chunk 2
    masks = [(5, 5), (20, 22)]
    cleaned = [_clean(stream[:, :S], masks)[0]]
    clipped = 0
    for k in range(nchunks):
        block, n = _clean(stream[:, S + k * N:S + (k + 1) * N], masks)
        cleaned.append(block)
        clipped += n
    assert clipped >= 4
    cleaned = np.concatenate(cleaned, axis=1)
    full = O.dedisperse_f32seq(cleaned, table, total - S)
    got = {}
    for i in range(nchunks + 3):
        buf = lib.mdsm_compat_buffer()
        if i == 0:
            n = S
            np.ctypeslib.as_array(buf, shape=(C_, S))[:] = stream[:, :S]
        elif i <= nchunks:
            n = N
            np.ctypeslib.as_array(buf, shape=(C_, N))[:] = stream[:, S + (i - 1) * N:S + i * N]
        else:
            n = 0
        samples = C.c_uint()
        outs = lib.mdsm_compat_next_chunk(n, C.byref(samples), 64e-6 * i * N, 64e-6)
        if i >= 3 and i - 2 <= nchunks:
            got[i - 2] = np.ctypeslib.as_array(outs[0], shape=(M, N)).copy()
        if lib.mdsm_compat_start(n) == 0:
            break
    lib.mdsm_compat_close()
    assert sorted(got) == list(range(1, nchunks + 1))
    for k, arr in got.items():
        series = full[:, (k - 1) * N:k * N]
        want = O.detrend_normalise(O.median_filter5(series), N).reshape(M, N)
        # replacement values come from two least-squares solvers (1 ulp at most)
        np.testing.assert_allclose(arr, want, rtol=1e-4, atol=1e-4, err_msg=f"chunk {k}")


XML_CLASSIFY = """<observation>
  <dm lowDM="0" numDMs="64" dmStep="1.0" />
  <channels nchans="64" npols="1" />
  <timing tsamp="0.000064" />
  <samples nsamp="8192" nbits="32" />
  <detection detectionThreshold="5" />
  <clustering applyClustering="1" minPoints="3" timeRange="8" dmRange="3" snrRange="100"
              applyClassification="1" minPulseWidth="0.1" />
  <writer filePrefix="cls" baseDirectory="{dir}" singleFileMode="1" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-1.0" /></beams>
</observation>"""


def test_streaming_manager_classification(tmp_path):
    xml = tmp_path / "obs.xml"
    xml.write_text(XML_CLASSIFY.format(dir=str(tmp_path)))
    assert lib.mdsm_compat_open(str(xml).encode(), 0) == 0, lib.mdsm_compat_last_error()
    nb, nch, ns, nd, ms = (C.c_uint() for _ in range(5))
    lib.mdsm_compat_survey(C.byref(nb), C.byref(nch), C.byref(ns), C.byref(nd), C.byref(ms))
    C_, N, M, S = nch.value, ns.value, nd.value, ms.value
    print(f"[compat classify] C={C_} N={N} M={M} maxshift={S}", file=sys.stderr, flush=True)
    stream = O.generate(9, C_, 1500.0, -1.0, 64e-6, 0, S + N,
                        pulses=[(40.0, S + 2000, 64, 1.2), (0.0, S + 6000, 4, 6.0)])
    for i, n in enumerate([S, N, 0, 0]):
        buf = lib.mdsm_compat_buffer()
        if n:
            np.ctypeslib.as_array(buf, shape=(C_, n))[:] = stream[:, :S] if i == 0 else stream[:, S:S + N]
        samples = C.c_uint()
        lib.mdsm_compat_next_chunk(n, C.byref(samples), 0.0, 64e-6)
        r = lib.mdsm_compat_start(n)
        if i == 2:  # the first stop call finishes the last chunk
            t, d, s, cl = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_int)()
            lib.mdsm_compat_candidates.restype = C.c_uint
            cnt = lib.mdsm_compat_candidates(0, C.byref(t), C.byref(d), C.byref(s), C.byref(cl))
            tm = np.ctypeslib.as_array(t, (cnt,)).copy()
            dm = np.ctypeslib.as_array(d, (cnt,)).copy()
            sn = np.ctypeslib.as_array(s, (cnt,)).copy()
            lab = np.ctypeslib.as_array(cl, (cnt,)).copy()
        if r == 0:
            break
    lib.mdsm_compat_close()
    assert lab.max() >= 2, "expected at least two clusters"
    expect = {}
    for cid in range(1, lab.max() + 1):
        m = lab == cid
        score, _ = O.classify_ref(tm[m], dm[m], sn[m], 0.0, 1.0, M, 0.1, 1500.0, -1.0, C_)
        expect[cid] = cid if score < 0.1 else -2
    rows = [ln.split(",") for ln in (tmp_path / "cls.dat").read_text().strip().splitlines()]
    written = [int(r[3]) for r in rows if int(r[3]) != -1]
    want = [expect[cid] for cid in range(1, lab.max() + 1) for _ in range(int((lab == cid).sum()))]
    assert written == want
    assert any(v != -2 for v in want)  # the wide pulse at DM 40 is kept


# ------------------------------------------------------------------ helpers
def _open(tmp_path, text, env=None):
    xml = tmp_path / "obs.xml"
    xml.write_text(text.format(dir=str(tmp_path)))
    old = {}
    for k, v in (env or {}).items():
        old[k] = os.environ.get(k)
        os.environ[k] = v
    try:
        assert lib.mdsm_compat_open(str(xml).encode(), 0) == 0, lib.mdsm_compat_last_error()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k)
            else:
                os.environ[k] = v
    nb, nch, ns, nd, ms = (C.c_uint() for _ in range(5))
    lib.mdsm_compat_survey(C.byref(nb), C.byref(nch), C.byref(ns), C.byref(nd), C.byref(ms))
    return nch.value, ns.value, nd.value, ms.value


def _candidates(beam=0):
    lib.mdsm_compat_candidates.restype = C.c_uint
    t, d, s, cl = (C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)(),
                   C.POINTER(C.c_int)())
    n = lib.mdsm_compat_candidates(beam, C.byref(t), C.byref(d), C.byref(s), C.byref(cl))
    if n == 0:
        return np.zeros(0), np.zeros(0), np.zeros(0)
    return (np.ctypeslib.as_array(t, (n,)).copy(), np.ctypeslib.as_array(d, (n,)).copy(),
            np.ctypeslib.as_array(s, (n,)).copy())


def _drive(stream, C_, S, N, M, reads, ts=lambda i: 0.0, block=64e-6, before_stop=None):
    """Feeds the priming read and the chunks of `reads` (samples per chunk,
    stream column offsets follow), then two stop calls.  Returns the series
    returned for each chunk and the candidates finished by each call."""
    got, cands, col = {}, {}, S
    calls = [S] + list(reads) + [0, 0]
    for i, n in enumerate(calls):
        buf = lib.mdsm_compat_buffer()
        if i == 0:
            np.ctypeslib.as_array(buf, shape=(C_, S))[:] = stream[:, :S]
        elif n:
            # the buffer holds nsamp columns per channel; a short read fills
            # the first n of them
            np.ctypeslib.as_array(buf, shape=(C_, N))[:, :n] = stream[:, col:col + n]
            col += n
        if i == len(calls) - 2 and before_stop:
            before_stop()
        samples = C.c_uint()
        outs = lib.mdsm_compat_next_chunk(n, C.byref(samples), ts(i), block)
        if i >= 3 and i - 2 <= len(reads):
            # NULL when the series stay on the device
            got[i - 2] = np.ctypeslib.as_array(outs[0], shape=(M, N)).copy() if outs else None
        r = lib.mdsm_compat_start(n)
        if i >= 2:
            cands[i - 1] = _candidates()
        if r == 0:
            break
    return got, cands


XML_PLAIN = """<observation>
  <dm lowDM="0" numDMs="40" dmStep="1.0" />
  <channels nchans="64" npols="1" />
  <timing tsamp="0.000064" />
  <samples nsamp="4096" nbits="32" />
  <detection detectionThreshold="5" {detect} />
  {storage}
  <writer filePrefix="plain" baseDirectory="{{dir}}" singleFileMode="{single}" secondsPerFile="{spf}"
          usePCTime="0" />
  <beams><beam beamId="0" topFrequency="1500" frequencyOffset="-1.0" /></beams>
</observation>"""


def _plain(detect="", storage="", single=1, spf=600):
    return XML_PLAIN.format(detect=detect, storage=storage, single=single, spf=spf)


def _check_cands(cands, series, ts, block, stats_cols=None, mean_rms=None, thr=5.0):
    """Candidates of one chunk against the oracle's threshold on its series."""
    t, d, s = cands
    if mean_rms is None:
        sc = series.shape[1] if stats_cols is None else stats_cols
        mean, rms = O.global_stats(series, sc)
    else:
        mean, rms = mean_rms
    od, ot, osnr, on = O.threshold_ref(series, mean, rms, thr)
    assert len(t) == on and on > 0, (len(t), on)
    assert np.array_equal(np.rint((t - ts) / block).astype(int), ot)
    assert np.array_equal(d, od.astype(np.float32).astype(np.float64))
    assert np.array_equal(s.astype(np.float32), osnr)


def test_streaming_manager_short_last_buffer(tmp_path):
    C_, N, M, S = _open(tmp_path, _plain())
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    reads = [N, N, 1500]
    total = S + sum(reads)
    stream = O.generate(31, C_, 1500.0, -1.0, 64e-6, 0, total,
                        pulses=[(20.0, 3000, 8, 1.2), (17.0, N + 1900, 8, 1.5),
                                (25.0, 2 * N + 200, 8, 1.5)])
    print(f"[compat short] C={C_} N={N} M={M} S={S} reads={reads}", file=sys.stderr, flush=True)
    got, cands = _drive(stream, C_, S, N, M, reads, ts=lambda i: 10.0 * i)
    lib.mdsm_compat_close()
    padded = np.concatenate([stream, np.zeros((C_, N - 1500), np.float32)], axis=1)
    full = O.dedisperse_f32seq(padded, table, 3 * N)
    for k in (1, 2, 3):
        n = reads[k - 1]
        assert np.array_equal(got[k][:, :n], full[:, (k - 1) * N:(k - 1) * N + n]), k
    # every chunk's candidates: statistics and threshold over its valid columns
    for k in (1, 2, 3):
        n = reads[k - 1]
        series = np.ascontiguousarray(full[:, (k - 1) * N:(k - 1) * N + n])
        _check_cands(cands[k], series, 10.0 * k, 64e-6)


def test_streaming_manager_last_nsamp_statistics(tmp_path):
    """The reference's last buffer (src/MDSM.cpp:93-103 sets last_nsamp and
    still reports a full chunk): statistics over the first last_nsamp
    samples of each row, threshold over all of them
    (multibeam_dedispersion_output.cpp:216-229)."""
    C_, N, M, S = _open(tmp_path, _plain())
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    reads = [N, N]
    stream = O.generate(32, C_, 1500.0, -1.0, 64e-6, 0, S + 2 * N, pulses=[(19.0, 1300, 8, 1.5), (22.0, N + 900, 8, 1.5)])
    last = 2100
    got, cands = _drive(stream, C_, S, N, M, reads,
                        before_stop=lambda: lib.mdsm_compat_set_last_nsamp(last))
    lib.mdsm_compat_close()
    full = O.dedisperse_f32seq(stream, table, 2 * N)
    assert np.array_equal(got[2], full[:, N:])
    _check_cands(cands[1], np.ascontiguousarray(full[:, :N]), 0.0, 64e-6)
    _check_cands(cands[2], np.ascontiguousarray(full[:, N:]), 0.0, 64e-6, stats_cols=last)


def test_streaming_manager_detrended_threshold(tmp_path):
    """Detrended series are thresholded at mean 0 and rms 1 (process_brute,
    multibeam_dedispersion_output.cpp:165-166), checked on the series the
    manager returns."""
    C_, N, M, S = _open(tmp_path, _plain(detect='applyDetrending="1"'))
    stream = O.generate(33, C_, 1500.0, -1.0, 64e-6, 0, S + 2 * N, pulses=[(15.0, 700, 8, 1.5), (27.0, N + 2500, 8, 1.5)])
    got, cands = _drive(stream, C_, S, N, M, [N, N])
    lib.mdsm_compat_close()
    for k in (1, 2):
        _check_cands(cands[k], got[k], 0.0, 64e-6, mean_rms=(0.0, 1.0))


def test_streaming_manager_candidate_overflow(tmp_path):
    """More candidates than the buffer holds: the manager grows it and
    writes every one (a cap of 100 against thousands at threshold 2)."""
    C_, N, M, S = _open(tmp_path, _plain().replace('detectionThreshold="5"', 'detectionThreshold="2"'),
                        env={"MDSM_COMPAT_CAND_CAP": "100"})
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    stream = O.generate(34, C_, 1500.0, -1.0, 64e-6, 0, S + 2 * N)
    got, cands = _drive(stream, C_, S, N, M, [N, N])
    lib.mdsm_compat_close()
    full = O.dedisperse_f32seq(stream, table, 2 * N)
    for k in (1, 2):
        series = np.ascontiguousarray(full[:, (k - 1) * N:k * N])
        assert len(cands[k][0]) > 1000
        _check_cands(cands[k], series, 0.0, 64e-6, thr=2.0)
    lines = (tmp_path / "plain.dat").read_text().strip().splitlines()
    assert len(lines) == len(cands[1][0]) + len(cands[2][0])


def test_streaming_manager_file_rotation(tmp_path):
    """Multi-file mode: files named by the chunk timestamp are opened when a
    period starts and closed once secondsPerFile of data have been written
    (multibeam_dedispersion_output.cpp:21-86,403-406)."""
    C_, N, M, S = _open(tmp_path, _plain(single=0, spf=0))
    stream = O.generate(35, C_, 1500.0, -1.0, 64e-6, 0, S + 3 * N,
                        pulses=[(12.0, 500, 8, 2.0), (12.0, N + 500, 8, 2.0), (12.0, 2 * N + 500, 8, 2.0)])
    base = 1.7e9
    got, cands = _drive(stream, C_, S, N, M, [N, N, N], ts=lambda i: base + 100.0 * i)
    lib.mdsm_compat_close()
    import time as _t
    files = sorted(p.name for p in tmp_path.glob("plain_beam_0_*.dat"))
    want = sorted(f"plain_beam_0_{_t.strftime('%Y-%m-%d_%H:%M:%S', _t.localtime(int(base + 100.0 * i)))}.dat"
                  for i in (1, 2, 3))
    assert files == want
    for i in (1, 2, 3):
        name = f"plain_beam_0_{_t.strftime('%Y-%m-%d_%H:%M:%S', _t.localtime(int(base + 100.0 * i)))}.dat"
        lines = (tmp_path / name).read_text().strip().splitlines()
        assert len(lines) == len(cands[i][0]) > 0


@pytest.mark.parametrize("storage", ["fp16", "u8", "fp16pairs"])
def test_streaming_manager_reduced_storage(tmp_path, storage):
    """<storage type="fp16"/>, "u8" or "fp16pairs" (FP16 rows, channel pairs
    added in FP16 before the FP32 sum): each chunk's rows (overlap tail and
    new samples) are normalised per channel over their valid samples, then
    quantised; the returned series equal the oracle's dedispersion of the
    same quantised rows bit for bit, and the candidates its threshold."""
    C_, N, M, S = _open(tmp_path, _plain(storage=f'<storage type="{storage}" />'))
    dt = {"fp16": abi.F16, "u8": abi.U8, "fp16pairs": abi.F16_PAIRS}[storage]
    assert lib.mdsm_compat_storage() == dt
    table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
    gain, offset = O.channel_bandpass(36, C_)
    reads = [N, N, 2000]
    stream = O.generate(36, C_, 1500.0, -1.0, 64e-6, 0, S + sum(reads), gain=gain, offset=offset,
                        pulses=[(18.0, 1200, 8, 2.0), (24.0, N + 3000, 8, 2.0),
                                (30.0, 2 * N + 300, 4, 2.5)])
    got, cands = _drive(stream, C_, S, N, M, reads)
    lib.mdsm_compat_close()
    ld = (N + S + 3) // 4 * 4
    col = S
    for k, n in enumerate(reads, start=1):
        row = np.zeros((C_, ld), np.float32)
        row[:, :S] = stream[:, col - S:col]
        row[:, S:S + n] = stream[:, col:col + n]
        col += n
        m, sd = O.channel_stats(row, S + n)
        q = O.normalise(row, m, sd, {"fp16": "f16", "u8": "u8", "fp16pairs": "f16"}[storage])
        want = {"fp16": O.dedisperse_f16, "u8": O.dedisperse_u8,
                "fp16pairs": O.dedisperse_f16_pairs}[storage](q, table, N)
        assert np.array_equal(got[k][:, :n], want[:, :n]), k
        _check_cands(cands[k], np.ascontiguousarray(want[:, :n]), 0.0, 64e-6)


def test_streaming_manager_series_on_device(tmp_path):
    """<series download="0"/>: next_chunk_multibeam returns NULL once it
    would return series, no series cross the bus, and every chunk's
    candidates equal those of the same stream with the series downloaded
    (and the oracle's threshold on the oracle's series)."""
    reads = [4096, 4096, 4096]
    runs = {}
    for dl in (1, 0):
        d = tmp_path / f"dl{dl}"
        d.mkdir()
        C_, N, M, S = _open(d, _plain().replace("</observation>", f'  <series download="{dl}" />\n</observation>'))
        table = O.delay_table_ref(1500.0, -1.0, C_, 64e-6, 0.0, 1.0, M)
        stream = O.generate(44, C_, 1500.0, -1.0, 64e-6, 0, S + sum(reads),
                            pulses=[(20.0, 1500, 8, 1.5), (12.0, N + 2500, 8, 1.5), (30.0, 2 * N + 900, 8, 1.5)])
        runs[dl] = _drive(stream, C_, S, N, M, reads, ts=lambda i: 5.0 * i)
        lib.mdsm_compat_close()
    got1, c1 = runs[1]
    got0, c0 = runs[0]
    assert all(g is not None for g in got1.values()) and all(g is None for g in got0.values())
    assert sorted(c1) == sorted(c0)
    for k in sorted(c1):
        for a, b in zip(c1[k], c0[k]):
            assert np.array_equal(a, b), k
    full = O.dedisperse_f32seq(stream, table, 3 * N)
    for k in (1, 2, 3):
        _check_cands(c0[k], np.ascontiguousarray(full[:, (k - 1) * N:k * N]), 5.0 * k, 64e-6)
