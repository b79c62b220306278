"""The single-beam (legacy) manager API of libmdsm_legacy.so on the GPU,
driven as the Pelican module drives it: four chunks through the input
buffer's overlap-save rows (the first chunk fills whole rows, later chunks
the nsamp samples after the first maxshift), next_chunk returning the
previous chunk's output, then the two-call stop.  Bars: every returned
output equals the oracle's FP32 channel-order sum over the concatenated
stream with the legacy delay table, bit for bit; every chunk's candidates
(row, sample, snr) equal the oracle's legacy detector on that output, with
the chunk's own statistics or with the caller's per-chunk mean/rms; the
output file holds one reference-format line per candidate."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi  # noqa: F401  (loads libmdsm_amd first)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FCH1, FOFF, NCH, TSAMP, LOWDM, STEP, NDM, NSAMP, THR, SPC = 420.0, -0.25, 64, 0.000256, 0.0, 0.25, 96, 4096, 3.5, 1024


def _lib():
    lib = C.CDLL(os.path.join(ROOT, "mdsm_amd", "libmdsm_legacy.so"))
    lib.mdsm_legacy_open.argtypes = [C.c_char_p, C.c_uint, C.c_uint]
    lib.mdsm_legacy_info.argtypes = [C.POINTER(C.c_uint)] * 4
    lib.mdsm_legacy_buffer.restype = C.POINTER(C.c_float)
    lib.mdsm_legacy_next_chunk.restype = C.POINTER(C.c_float)
    lib.mdsm_legacy_next_chunk.argtypes = [C.c_uint, C.POINTER(C.c_uint), C.c_double, C.c_double,
                                           C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.mdsm_legacy_start.argtypes = [C.c_uint]
    lib.mdsm_legacy_candidates.restype = C.c_uint
    lib.mdsm_legacy_candidates.argtypes = [C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.POINTER(C.c_int32)),
                                           C.POINTER(C.POINTER(C.c_float))]
    lib.mdsm_legacy_last_error.restype = C.c_char_p
    return lib


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def test_legacy_manager_stream(tmp_path):
    xml = tmp_path / "obs.xml"
    xml.write_text(f'<observation><frequencies top="{FCH1}" offset="{FOFF}"/>'
                   f'<dm lowDM="{LOWDM}" numDMs="{NDM}" dmStep="{STEP}" useBruteForce="1"/>'
                   f'<channels number="{NCH}"/><timing tsamp="{TSAMP}"/><samples number="{NSAMP}"/>'
                   f'<detection threshold="{THR}"/>'
                   f'<output filePrefix="legacy_run" baseDirectory="{tmp_path}" singleFileMode="1"/>'
                   '</observation>')
    lib = _lib()
    assert lib.mdsm_legacy_open(str(xml).encode(), 0, SPC) == 0, lib.mdsm_legacy_last_error()
    try:
        info = [C.c_uint() for _ in range(4)]
        assert lib.mdsm_legacy_info(*[C.byref(v) for v in info]) == 0
        nch, nsamp, S, ndm = (v.value for v in info)
        assert (nch, nsamp, ndm) == (NCH, NSAMP, NDM)
        table = O.delay_table_legacy(FCH1, FOFF, NCH, TSAMP, LOWDM, STEP, NDM)
        shifts = O.dm_shifts_ref(FCH1, FOFF, NCH)
        hi = np.float32(np.float32(LOWDM) + np.float32(STEP) * np.float32(NDM - 1))
        assert S == int(np.float32(np.float32(shifts[-1] * hi) / np.float32(TSAMP)))
        print(f"[legacy stream] C={nch} M={ndm} N={nsamp} maxshift={S} table max={table.max()}",
              file=sys.stderr, flush=True)
        K = 4
        rng = np.random.default_rng(40)
        width = S + K * nsamp
        x = rng.standard_normal((nch, width + 8)).astype(np.float32)
        for t0, d in ((1500, 80), (nsamp + 1500, 60), (2 * nsamp + 300, 30), (3 * nsamp + 2000, 45)):
            for c in range(nch):
                x[c, t0 + table[c, d]: t0 + table[c, d] + 4] += 1.5
        buf = np.ctypeslib.as_array(lib.mdsm_legacy_buffer(), shape=(nch, nsamp + S))
        pad = max(0, int(table.max()) - S)
        outs, total = [], 0
        for k in range(K + 1):
            if k < K:
                if k == 0:
                    buf[:, :] = x[:, :nsamp + S]
                else:
                    buf[:, S:] = x[:, S + k * nsamp: S + (k + 1) * nsamp]
                stats = k == 2
                nst = (nsamp - 5) // SPC + 1
                mean = (rng.random(nst) * 0.2).astype(np.float32) if stats else None
                rms = (7.5 + rng.random(nst)).astype(np.float32) if stats else None
                got = C.c_uint()
                out = lib.mdsm_legacy_next_chunk(nsamp, C.byref(got), 100.0 + k, TSAMP, _fptr(mean), _fptr(rms))
            else:
                got = C.c_uint()
                out = lib.mdsm_legacy_next_chunk(0, C.byref(got), 0.0, 0.0, None, None)
            if k == 0:
                assert not out and got.value == 0xFFFFFFFF
            else:
                assert got.value == nsamp
                y = np.ctypeslib.as_array(out, shape=(ndm, nsamp))
                assert np.array_equal(y, outs[k - 1]), f"output of chunk {k - 1}"
            if k == K:
                break
            assert lib.mdsm_legacy_start(nsamp) == k + 1, lib.mdsm_legacy_last_error()
            win = np.concatenate([x[:, k * nsamp: k * nsamp + nsamp + S], np.zeros((nch, pad), np.float32)], 1)
            yk = O.dedisperse_f32seq(win, table, nsamp)
            outs.append(yk)
            if mean is not None:
                od, ot, osnr, on = O.process_brute_legacy_chunks(yk, LOWDM, STEP, THR, mean, rms, SPC)
            else:
                od, ot, osnr, on, _, _ = O.process_brute_legacy(yk, LOWDM, STEP, THR, nvals=nsamp * ndm)
            pd, ps, pn = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_float)()
            n = lib.mdsm_legacy_candidates(C.byref(pd), C.byref(ps), C.byref(pn))
            print(f"  chunk {k}: {n} candidates (oracle {on})", file=sys.stderr, flush=True)
            assert n == on and n > 0
            assert np.ctypeslib.as_array(pd, (n,)).tolist() == od.tolist()
            assert np.ctypeslib.as_array(ps, (n,)).tolist() == ot.tolist()
            assert np.array_equal(np.ctypeslib.as_array(pn, (n,)), osnr)
            total += n
        assert lib.mdsm_legacy_start(0) != 0          # first stop call: last output was returned
        got = C.c_uint()
        lib.mdsm_legacy_next_chunk(0, C.byref(got), 0.0, 0.0, None, None)
        assert lib.mdsm_legacy_start(0) == 0          # second stop call ends the run
    finally:
        lib.mdsm_legacy_close()
    lines = (tmp_path / "legacy_run.dat").read_text().splitlines()
    assert len(lines) == total
    assert len(lines[0].split(", ")) in (3, 5)


def test_legacy_manager_file_rotation(tmp_path):
    """Without singleFileMode the candidates go to
    <basedir>/<fileprefix>_<time>_<secondsPerFile>.dat, named from the data
    timestamp (usePCTime 0), and a new file starts once the samples written
    span more than secondsPerFile (dedispersion_output.cpp:254-281, 296-302)."""
    xml = tmp_path / "obs.xml"
    xml.write_text(f'<observation><frequencies top="{FCH1}" offset="{FOFF}"/>'
                   f'<dm lowDM="{LOWDM}" numDMs="{NDM}" dmStep="{STEP}" useBruteForce="1"/>'
                   f'<channels number="{NCH}"/><timing tsamp="{TSAMP}"/><samples number="{NSAMP}"/>'
                   f'<detection threshold="{THR}"/>'
                   f'<output filePrefix="rot" baseDirectory="{tmp_path}" singleFileMode="0" usePCTime="0" '
                   'secondsPerFile="1"/></observation>')
    lib = _lib()
    assert lib.mdsm_legacy_open(str(xml).encode(), 0, SPC) == 0, lib.mdsm_legacy_last_error()
    try:
        info = [C.c_uint() for _ in range(4)]
        lib.mdsm_legacy_info(*[C.byref(v) for v in info])
        nch, nsamp, S, _ = (v.value for v in info)
        buf = np.ctypeslib.as_array(lib.mdsm_legacy_buffer(), shape=(nch, nsamp + S))
        rng = np.random.default_rng(41)
        print(f"[legacy rotation] C={nch} N={nsamp} maxshift={S}", file=sys.stderr, flush=True)
        for k in range(3):
            buf[:, :] = (rng.standard_normal((nch, nsamp + S)) * 3).astype(np.float32)
            got = C.c_uint()
            # 4096 samples x 0.000256 s = 1.05 s per chunk > secondsPerFile: one file per chunk,
            # named from the chunk's timestamp 1000, 2000, 3000 s
            lib.mdsm_legacy_next_chunk(nsamp, C.byref(got), 1000.0 * (k + 1), TSAMP, None, None)
            assert lib.mdsm_legacy_start(nsamp) == k + 1, lib.mdsm_legacy_last_error()
    finally:
        lib.mdsm_legacy_close()
    files = sorted(p.name for p in tmp_path.glob("rot_*_1.dat"))
    assert len(files) == 3, files
