"""The per-chunk search pipeline (normalise -> batched dedispersion ->
block statistics -> detection) against the oracle, eager and replayed from
a hipGraph.  Graph replays are interleaved with eager runs and host
synchronisation, and every replay must give the eager result bit for bit."""
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


def _beams():
    out = []
    for C, M, N, step in [(64, 48, 12000, 0.3), (128, 40, 9000, 0.5), (32, 70, 15000, 0.2)]:
        out.append(abi.Survey.make(1500.0, -300.0 / C, C, 64e-6, 0.0, step, M, N))
    return out


@pytest.mark.parametrize("dtype,name", [(abi.F16, "f16"), (abi.U8, "u8"), (abi.F32, "f32"),
                                        (abi.F16_PAIRS, "f16_pairs")])
def test_pipeline_eager_and_graph_match_oracle(dtype, name):
    beams = _beams()
    pipe = abi.Pipeline(beams, dtype, stats_block=4096, nwidths=7, thresh=5.0, cand_cap=1 << 18)
    print(f"[pipeline {name}] {len(beams)} beams", file=sys.stderr, flush=True)
    raws, keep, refs = [], [], []
    for b, s in enumerate(beams):
        t, S = abi.delay_table(s)
        ld = s.nsamp + S
        g, o = O.channel_bandpass(10 + b, s.nchans)
        x = O.generate(10 + b, s.nchans, s.fch1, s.foff, s.tsamp, 0, ld, g, o,
                       pulses=[(s.dmstep * (s.ndms // 2), s.nsamp // 3, 8, 0.4)])
        d = abi.DeviceBuffer.from_numpy(x)
        keep.append(d)
        raws.append((d.ptr, ld))
        # oracle: channel stats over the stored samples (nsamp + maxshift, rounded
        # up to the storage alignment, within the raw row), normalise, dedisperse, detect
        n = min(ld, -(-ld // abi.ALIGN[dtype]) * abi.ALIGN[dtype])
        m, sd = O.channel_stats(x, n)
        xs = O.normalise(x, m, sd, "f16" if name == "f16_pairs" else name, n)
        y = {"f32": O.dedisperse_f32seq, "f16": O.dedisperse_f16, "u8": O.dedisperse_u8,
             "f16_pairs": O.dedisperse_f16_pairs}[name](xs, t, s.nsamp)
        refs.append((y, O.detect(y.astype(np.float64), 4096, 7, 5.0)))
    pipe.bind([p for p, _ in raws], [l for _, l in raws])

    def check(tag):
        for b, s in enumerate(beams):
            y, (od, ot, ok, osnr, on) = refs[b]
            got = np.empty((s.ndms, s.nsamp), np.float32)
            abi.check(abi.lib().mdsm_memcpy_dtoh(got.ctypes.data, pipe.output_ptr(b), got.nbytes, None))
            assert np.array_equal(got, y), (tag, b)
            c = pipe.candidates(b)
            assert len(c) == on > 0, (tag, b)
            assert list(zip(c["dm_idx"], c["sample"], c["width_log2"])) == list(zip(od, ot, ok)), (tag, b)

    pipe.run()
    abi.synchronize()
    check("eager")
    for rep in range(3):
        pipe.run(graph=True)
        abi.synchronize()
        check(f"graph {rep}")
        pipe.run()
        abi.synchronize()
        check(f"eager {rep}")
