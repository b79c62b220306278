"""GPU parity of the multi-beam dedisperser (persistent queue and flat
XCD-ordered launch): every beam of an uneven batch (different nchans,
nsamp, ndms, dmstep, band) must equal the oracle bit for bit, for F32, F16,
F16 channel-pair sums and U8 storage."""
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu


def _beams(seed, n, odd=False):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        C = int(rng.choice([32, 64, 96, 128])) - (1 if odd and k % 2 else 0)
        N = int(rng.integers(1000, 9000))
        M = int(rng.integers(10, 140))
        step = float(rng.uniform(0.1, 0.5))
        fch1 = float(rng.choice([1500.0, 418.0]))
        foff = -(0.29296875 if fch1 > 1000 else 0.078125) * 512 / C
        out.append(abi.Survey.make(fch1, foff, C, 64e-6, 0.0, step, M, N))
    return out


@pytest.mark.parametrize("dtype", [abi.F32, abi.F16, abi.U8, abi.F16_PAIRS])
def test_batch_uneven_beams_bitexact(dtype):
    # the pair sums also get odd channel counts (a zero partner row)
    beams = _beams(100 + dtype, 9, odd=dtype == abi.F16_PAIRS)
    batch = abi.Batch(beams, dtype)
    print(f"[batch dtype {dtype}] {len(beams)} beams", flush=True, file=sys.stderr)
    rng = np.random.default_rng(dtype)
    ins, outs, refs, keep = [], [], [], []
    for b, s in enumerate(beams):
        ld = batch.min_ld[b] + 8
        x = rng.standard_normal((s.nchans, ld)).astype(np.float32)
        t, _ = abi.delay_table(s)
        if dtype == abi.F32:
            h = x
            refs.append(O.dedisperse_f32seq(x, t, s.nsamp))
        elif dtype == abi.F16:
            h = O.f32_to_f16(x)
            refs.append(O.dedisperse_f16(h, t, s.nsamp))
        elif dtype == abi.F16_PAIRS:
            h = O.f32_to_f16(x)
            refs.append(O.dedisperse_f16_pairs(h, t, s.nsamp))
        else:
            h = rng.integers(0, 256, x.shape).astype(np.uint8)
            refs.append(O.dedisperse_u8(h, t, s.nsamp))
        di = abi.DeviceBuffer.from_numpy(h)
        do = abi.DeviceBuffer(s.ndms * s.nsamp * 4)
        keep += [di, do]
        ins.append((di.ptr, ld))
        outs.append(do)
    for mode in (abi.BATCH_PERSISTENT, abi.BATCH_FLAT):
        batch.set_mode(mode)
        for v in (0, 1, 4, 9, 12, 13, 15, "tuned"):
            for b, s in enumerate(beams):  # every tile must be written again
                outs[b].upload(np.full(s.ndms * s.nsamp, np.nan, np.float32))
            print(f"[batch dtype {dtype}] mode {mode} variant {v}", flush=True, file=sys.stderr)
            if v == "tuned":  # the library's batch autotuner picks and keeps a shape
                got_v, ms = batch.autotune([p for p, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
                assert 1 <= got_v <= abi.tile_variants(dtype) and ms > 0
            else:
                batch.set_variant(v)
            batch.dedisperse([p for p, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
            abi.synchronize()
            for b, s in enumerate(beams):
                got = outs[b].download(np.float32, (s.ndms, s.nsamp))
                assert np.array_equal(got, refs[b]), (mode, v, b)


@pytest.mark.parametrize("dtype", [abi.F16_PAIRS, abi.U8])
def test_batch_pairsum_tiles_bitexact(dtype):
    """F16 channel-pair sums and U8 over an uneven batch of narrow-channel
    beams (odd channel counts among them) that the pair-sum tile shapes
    take: every such shape, in both launch forms, and the autotuner's pick
    equal the oracle bit for bit."""
    rng = np.random.default_rng(515)
    beams = []
    for k in range(7):
        C = int(rng.choice([256, 384, 512])) - (1 if k % 2 else 0)
        N = int(rng.integers(1500, 7000))
        M = int(rng.integers(30, 200))
        beams.append(abi.Survey.make(418.0, -0.078125 * 256 / C, C, 64e-6, 0.0,
                                     float(rng.uniform(0.05, 0.2)), M, N))
    batch = abi.Batch(beams, dtype)
    print(f"[batch pair sums] dtype {dtype}, {len(beams)} beams", flush=True, file=sys.stderr)
    ins, outs, refs, keep = [], [], [], []
    for b, s in enumerate(beams):
        ld = batch.min_ld[b] + 4
        t, _ = abi.delay_table(s)
        if dtype == abi.U8:
            h = rng.integers(0, 256, (s.nchans, ld)).astype(np.uint8)
            refs.append(O.dedisperse_u8(h, t, s.nsamp))
        else:
            h = O.f32_to_f16(rng.standard_normal((s.nchans, ld)).astype(np.float32))
            refs.append(O.dedisperse_f16_pairs(h, t, s.nsamp))
        di = abi.DeviceBuffer.from_numpy(h)
        do = abi.DeviceBuffer(s.ndms * s.nsamp * 4)
        keep += [di, do]
        ins.append((di.ptr, ld))
        outs.append(do)
    args = ([p for p, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
    shapes = list(range(abi.tile_variants() + 1, abi.tile_variants(dtype) + 1))
    for mode in (abi.BATCH_PERSISTENT, abi.BATCH_FLAT):
        batch.set_mode(mode)
        for v in shapes + ["tuned"]:
            for b, s in enumerate(beams):
                outs[b].upload(np.full(s.ndms * s.nsamp, np.nan, np.float32))
            print(f"[batch pair sums] mode {mode} variant {v}", flush=True, file=sys.stderr)
            if v == "tuned":
                got_v, ms = batch.autotune(*args)
                assert 1 <= got_v <= abi.tile_variants(dtype) and ms > 0
            else:
                batch.set_variant(v)
            try:
                batch.dedisperse(*args)
            except abi.MdsmError:
                # wider DM tiles take only some of these beams: a clean error;
                # the narrowest takes every beam
                assert v not in (shapes[0], "tuned"), v
                continue
            abi.synchronize()
            for b, s in enumerate(beams):
                assert np.array_equal(outs[b].download(np.float32, (s.ndms, s.nsamp)), refs[b]), (mode, v, b)


@pytest.mark.parametrize("dtype", [abi.F32, abi.F16])
def test_batch_autotune_skips_shapes_that_do_not_fit(dtype):
    """A beam with about 1000 samples of delay per DM step (64 channels
    400 -> 274 MHz, DM step 2.2): the wide tile shapes, the untuned default
    among them, stage rows longer than the LDS holds.  The untuned plan and
    the untuned batch must fall back to a shape that fits, the autotuner
    must choose among those, and all three equal the oracle bit for bit."""
    beams = [abi.Survey.make(400.0, -2.0, 64, 64e-6, 0.0, 2.2, 100, 4096),
             abi.Survey.make(1500.0, -0.29296875 * 8, 64, 64e-6, 0.0, 0.3, 60, 3000)]
    # premise: the default shape cannot stage the wide beam
    wide = abi.Plan(beams[0], dtype)
    pref = abi.preferred_variant(dtype)
    assert wide.variant_channels(pref) == 0
    fallback = wide.variant
    assert fallback != pref and wide.variant_channels(fallback) > 0
    batch = abi.Batch(beams, dtype)
    print(f"[batch autotune, wide spread, dtype {dtype}] maxshift {batch.maxshift}", flush=True,
          file=sys.stderr)
    rng = np.random.default_rng(77)
    ins, outs, refs, keep = [], [], [], []
    for b, s in enumerate(beams):
        ld = batch.min_ld[b]
        x = rng.standard_normal((s.nchans, ld)).astype(np.float32)
        t, _ = abi.delay_table(s)
        if dtype == abi.F32:
            h = x
            refs.append(O.dedisperse_f32seq(x, t, s.nsamp))
        else:
            h = O.f32_to_f16(x)
            refs.append(O.dedisperse_f16(h, t, s.nsamp))
        di = abi.DeviceBuffer.from_numpy(h)
        do = abi.DeviceBuffer(s.ndms * s.nsamp * 4)
        keep += [di, do]
        ins.append((di.ptr, ld))
        outs.append(do)
    # (a) the untuned single-beam plan on the wide beam
    outs[0].upload(np.full(beams[0].ndms * beams[0].nsamp, np.nan, np.float32))
    wide.dedisperse(ins[0][0], ins[0][1], outs[0].ptr)
    abi.synchronize()
    assert np.array_equal(outs[0].download(np.float32, (beams[0].ndms, beams[0].nsamp)), refs[0])
    # (b) the untuned batch, then (c) the autotuned batch
    for tuned in (False, True):
        if tuned:
            v, ms = batch.autotune([p for p, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
            assert 1 <= v <= abi.tile_variants() and ms > 0
        else:
            v = "untuned"
        for b, s in enumerate(beams):
            outs[b].upload(np.full(s.ndms * s.nsamp, np.nan, np.float32))
        batch.dedisperse([p for p, _ in ins], [l for _, l in ins], [o.ptr for o in outs])
        abi.synchronize()
        for b, s in enumerate(beams):
            assert np.array_equal(outs[b].download(np.float32, (s.ndms, s.nsamp)), refs[b]), (v, b)
