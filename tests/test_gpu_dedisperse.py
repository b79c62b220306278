"""GPU parity of the dedispersion kernels against the CPU oracle.

Bar: bit-exact against the oracle's FP32 channel-order sum (the kernels add
in channel order 0..C-1, one rounding per add) and against the oracle's
exact-integer U8 sum; max error against the FP64 oracle stated per case
(tolerance: max |y - y64| <= 1e-5 * rms(y64) * sqrt(C) for FP32 inputs).
"""
import sys

import numpy as np
import pytest

import oracle as O
from mdsm_amd import abi

pytestmark = pytest.mark.gpu

CFG0 = dict(fch1=418.0, foff=-0.078125, nchans=256, tsamp=51.2e-6, lowdm=0.0, dmstep=0.1, ndms=256)


def _say(*a):
    print(*a, flush=True, file=sys.stderr)


def _survey(nsamp, mode=abi.DELAY_FP64_ROUND, **kw):
    p = dict(CFG0)
    p.update(kw)
    return abi.Survey.make(p["fch1"], p["foff"], p["nchans"], p["tsamp"], p["lowdm"], p["dmstep"],
                           p["ndms"], nsamp, mode)


def _table(s):
    t, _ = abi.delay_table(s)
    return t


def _err_stats(y, y64):
    e = np.abs(y.astype(np.float64) - y64)
    return e.max(), e.max() / np.sqrt((y64 * y64).mean())


@pytest.mark.parametrize("variant", list(range(1, abi.tile_variants() + 1)))
def test_f32_256ch_bitexact(variant):
    s = _survey(65536)
    t = _table(s)
    S = int(t.max())
    _say(f"[f32 256-ch variant {variant}] C=256 N=65536 M=256 S={S}")
    x = O.generate(1, 256, s.fch1, s.foff, s.tsamp, 0, 65536 + S)
    y = abi.dedisperse_host(s, x, abi.F32, variant)
    ref = O.dedisperse_f32seq(x, t, 65536)
    assert np.array_equal(y, ref)
    if variant == 1:
        y64 = O.dedisperse_f64(x, t, 65536)
        mx, rel = _err_stats(y, y64)
        _say(f"    max|y-y64| = {mx:.3e}, / rms = {rel:.3e}")
        assert rel < 1e-5


@pytest.mark.parametrize("shape", [(37, 45, 1000, 2.5, 1.7), (1, 1, 1, 0.0, 1.0), (130, 129, 777, 0.0, 9.0),
                                   (64, 300, 4096, 10.0, 0.05)])
def test_f32_ragged_shapes(shape):
    C, M, N, lo, step = shape
    s = _survey(N, nchans=C, ndms=M, lowdm=lo, dmstep=step, foff=-0.078125 * 256 / max(C, 1) / 4)
    t = _table(s)
    S = int(t.max())
    _say(f"[f32 ragged] C={C} M={M} N={N} S={S}")
    x = np.random.default_rng(C * 1000 + M).standard_normal((C, N + S + 3)).astype(np.float32)
    ref = O.dedisperse_f32seq(x, t, N)
    for v in range(1, abi.tile_variants() + 1):
        assert np.array_equal(abi.dedisperse_host(s, x, abi.F32, v), ref), v


def test_known_answer_channel_index():
    s = _survey(4096)
    t = _table(s)
    S = int(t.max())
    x = np.repeat(np.arange(256, dtype=np.float32)[:, None], 4096 + S, axis=1)
    y = abi.dedisperse_host(s, x)
    assert (y == 256 * 255 / 2).all()


def test_f16_u8_bitexact_and_error():
    s = _survey(16384)
    t = _table(s)
    S = int(t.max())
    _say(f"[f16/u8] C=256 N=16384 M=256 S={S}")
    x = O.generate(5, 256, s.fch1, s.foff, s.tsamp, 0, 16384 + S)
    h = O.f32_to_f16(x)
    y16 = abi.dedisperse_host(s, h, abi.F16)
    assert np.array_equal(y16, O.dedisperse_f16(h, t, 16384))
    m, sd = O.channel_stats(x)
    u = O.normalise(x, m, sd, "u8")
    y8 = abi.dedisperse_host(s, u, abi.U8)
    assert np.array_equal(y8, O.dedisperse_u8(u, t, 16384))
    y16p = abi.dedisperse_host(s, h, abi.F16_PAIRS)
    assert np.array_equal(y16p, O.dedisperse_f16_pairs(h, t, 16384))
    # errors against FP64 of the unquantised input, reported
    y64 = O.dedisperse_f64(x, t, 16384)
    mx, rel = _err_stats(y16, y64)
    _say(f"    f16 storage: max|y-y64|/rms = {rel:.3e}")
    assert rel < 2e-3
    mxp, relp = _err_stats(y16p, y64)
    _say(f"    f16 pair sums: max|y-y64|/rms = {relp:.3e}")
    assert relp < 4e-3


@pytest.mark.parametrize("shape", [(37, 45, 1000, 2.5, 1.7), (130, 129, 777, 0.0, 9.0), (600, 70, 3001, 0.0, 0.3)])
def test_packed_ragged_shapes_all_variants(shape):
    """F16 and U8 storage, every tile variant, on shapes that leave partial
    channel blocks, DM tiles and time tiles (odd channel counts for the pair
    copies; 600 channels cross the U8 16-bit fold every 256 channels)."""
    C, M, N, lo, step = shape
    s = _survey(N, nchans=C, ndms=M, lowdm=lo, dmstep=step, foff=-0.078125 * 256 / C / 4)
    t = _table(s)
    S = int(t.max())
    _say(f"[packed ragged] C={C} M={M} N={N} S={S}")
    rng = np.random.default_rng(C * 7 + M)
    h = O.f32_to_f16(rng.standard_normal((C, N + S + 5)).astype(np.float32))
    u = rng.integers(0, 256, (C, N + S + 5)).astype(np.uint8)
    ref16 = O.dedisperse_f16(h, t, N)
    ref16p = O.dedisperse_f16_pairs(h, t, N)
    ref8 = O.dedisperse_u8(u, t, N)
    p16, p16p = abi.Plan(s, abi.F16), abi.Plan(s, abi.F16_PAIRS)
    for v in range(1, abi.tile_variants() + 1):
        assert np.array_equal(abi.dedisperse_host(s, h, abi.F16, v), ref16), ("f16", v)
        if p16p.variant_channels(v) == 0:
            # a channel pair per LDS block does not fit this delay spread,
            # where F16 stages one row at a time: a clean error
            assert p16.variant_channels(v) <= 1, v
            with pytest.raises(abi.MdsmError):
                abi.dedisperse_host(s, h, abi.F16_PAIRS, v)
        else:
            assert p16p.variant_channels(v) % 2 == 0, v
            assert np.array_equal(abi.dedisperse_host(s, h, abi.F16_PAIRS, v), ref16p), ("f16 pairs", v)
        assert np.array_equal(abi.dedisperse_host(s, u, abi.U8, v), ref8), ("u8", v)
    # pair-sum tiles (F16 pair sums, U8): equal where they take the survey,
    # a clean error where not
    p8 = abi.Plan(s, abi.U8)
    for v in range(abi.tile_variants() + 1, abi.tile_variants(abi.F16_PAIRS) + 1):
        for plan, dt, arr, ref in ((p16p, abi.F16_PAIRS, h, ref16p), (p8, abi.U8, u, ref8)):
            if plan.variant_channels(v) == 0:
                with pytest.raises(abi.MdsmError):
                    abi.dedisperse_host(s, arr, dt, v)
            else:
                assert np.array_equal(abi.dedisperse_host(s, arr, dt, v), ref), ("pair-sum tile", dt, v)


# (nchans, ndms, nsamp, lowdm, dmstep, foff): surveys whose channel pairs
# differ in delay by at most three values per DM tile, so the pair-sum tile
# shapes take them (odd channel counts, partial DM tiles and time tiles, one
# channel summed with zero, one trial shorter than a time tile)
PAIRSUM_SURVEYS = [(256, 100, 1300, 0.0, 0.1, -0.078125), (255, 70, 1000, 0.0, 0.1, -0.078125),
                   (128, 64, 2100, 5.0, 0.1, -0.078125), (99, 49, 777, 0.0, 0.1, -0.078125),
                   (512, 40, 600, 0.0, 0.25, -0.0390625), (1024, 300, 5000, 0.0, 0.2, -0.0390625),
                   (1, 5, 300, 0.0, 0.1, -0.078125), (64, 1, 100, 0.0, 0.1, -0.078125)]


@pytest.mark.parametrize("shape", PAIRSUM_SURVEYS)
def test_pairsum_tiles_bitexact(shape):
    """The pair-sum tile shapes (each pair's FP16 sums built once per delay
    difference, then read by the trials) equal the oracle's channel-pair
    sums bit for bit, like the F16_PAIRS forms of the F16 shapes."""
    C, M, N, lo, step, foff = shape
    s = _survey(N, nchans=C, ndms=M, lowdm=lo, dmstep=step, foff=foff)
    t = _table(s)
    S = int(t.max())
    _say(f"[pair-sum tiles] C={C} M={M} N={N} S={S}")
    rng = np.random.default_rng(C + M)
    h = O.f32_to_f16((rng.standard_normal((C, N + S + 6)) * 3).astype(np.float32))
    u = rng.integers(0, 256, (C, N + S + 8)).astype(np.uint8)
    for dt, arr, ref in ((abi.F16_PAIRS, h, O.dedisperse_f16_pairs(h, t, N)),
                         (abi.U8, u, O.dedisperse_u8(u, t, N))):
        p = abi.Plan(s, dt)
        for v in range(abi.tile_variants() + 1, abi.tile_variants(dt) + 1):
            if p.variant_channels(v) == 0:
                continue
            assert np.array_equal(abi.dedisperse_host(s, arr, dt, v), ref), (dt, v)
        # the first (narrowest) shape takes every survey of this list
        assert p.variant_channels(abi.tile_variants() + 1) > 0


@pytest.mark.parametrize("C", [2, 3, 10, 65])
def test_f16_pairs_rounding_cases_all_variants(C):
    """The FP16 pair sums at their rounding edges: ties to even (2048 + 1,
    1 + 2^-11, 1 + 3*2^-11), the overflow tie (65504 + 16 -> inf, and its
    negative), subnormal sums, exact cancellation, next to random halves;
    odd channel counts pair the last channel with zero.  Every tile shape
    equals the oracle bit for bit."""
    N, M = 1024, 3
    s = _survey(N, nchans=C, ndms=M, lowdm=0.0, dmstep=0.5, foff=-0.078125 * 256 / C / 4)
    t = _table(s)
    S = int(t.max())
    rng = np.random.default_rng(C)
    x = (rng.standard_normal((C, N + S + 4)) * 3).astype(np.float32)
    edge = [(2048.0, 1.0), (1.0, 2.0 ** -11), (1.0, 3 * 2.0 ** -11), (65504.0, 16.0), (-65504.0, -16.0),
            (2.0 ** -24, 2.0 ** -24), (-2.0 ** -14, 2.0 ** -24), (5.5, -5.5), (1000.0, 0.25)]
    for e, (a, b) in enumerate(edge):  # planted along DM 0 (rows 0 and 1 at the same column)
        col = 7 + 97 * e
        x[0, col + t[0, 0]] = a
        x[1, col + t[1, 0]] = b
    h = O.f32_to_f16(x)
    ref = O.dedisperse_f16_pairs(h, t, N)
    _say(f"[f16 pairs edges] C={C} M={M} N={N} S={S}")
    p = abi.Plan(s, abi.F16_PAIRS)
    for v in range(1, abi.tile_variants(abi.F16_PAIRS) + 1):
        if p.variant_channels(v) == 0:
            continue
        assert np.array_equal(abi.dedisperse_host(s, h, abi.F16_PAIRS, v), ref), v
    if C >= 10:  # the pair-sum shapes take these (adjacent delays differ by <= 1 over 3 trials)
        assert p.variant_channels(abi.tile_variants() + 1) > 0


@pytest.mark.parametrize("dtype,fallback", [(abi.F16_PAIRS, 15), (abi.U8, 2)])
def test_untuned_default_pairsum_or_fallback(dtype, fallback):
    """Untuned plans take the pair-sum shape where it takes the survey (many
    narrow channels) and the storage type's shape for every survey where it
    does not (few wide channels); both equal the oracle."""
    pref = abi.preferred_variant(dtype)
    assert pref > abi.tile_variants()  # a pair-sum shape
    for C, foff, want in ((512, -0.0390625, pref), (32, -0.625, fallback)):
        s = _survey(2000, nchans=C, ndms=60, dmstep=0.25, foff=foff)
        t = _table(s)
        S = int(t.max())
        p = abi.Plan(s, dtype)
        _say(f"[untuned default] dtype {dtype} C={C} S={S} -> variant {p.variant}")
        assert p.variant == want
        rng = np.random.default_rng(C)
        if dtype == abi.U8:
            x = rng.integers(0, 256, (C, 2000 + S + 4)).astype(np.uint8)
            ref = O.dedisperse_u8(x, t, 2000)
        else:
            x = O.f32_to_f16(rng.standard_normal((C, 2000 + S + 4)).astype(np.float32))
            ref = O.dedisperse_f16_pairs(x, t, 2000)
        assert np.array_equal(abi.dedisperse_host(s, x, dtype), ref)


def test_u8_near_saturated_sums_all_variants():
    """Bytes of 250..255 over 1500 channels: sums up to 382500 need the
    high-part accumulator after every 256-channel window, in every variant."""
    C, M, N = 1500, 40, 600
    s = _survey(N, nchans=C, ndms=M, lowdm=0.0, dmstep=0.5, foff=-0.078125 * 256 / C / 4)
    t = _table(s)
    S = int(t.max())
    _say(f"[u8 saturated] C={C} M={M} N={N} S={S}")
    u = np.random.default_rng(31).integers(250, 256, (C, N + S + 4)).astype(np.uint8)
    u[:, ::7] = 255
    ref8 = O.dedisperse_u8(u, t, N)
    p8 = abi.Plan(s, abi.U8)
    for v in range(1, abi.tile_variants(abi.U8) + 1):
        if v > abi.tile_variants():  # pair-sum tiles: pair sums up to 510, folded every 128 pairs
            assert p8.variant_channels(v) > 0, v
        assert np.array_equal(abi.dedisperse_host(s, u, abi.U8, v), ref8), ("u8", v)


def test_reference_delay_mode():
    s = _survey(8192, mode=abi.DELAY_REFERENCE)
    t = O.delay_table_ref(s.fch1, s.foff, 256, s.tsamp, s.lowdm, s.dmstep, 256)
    S = int(t.max())
    x = np.random.default_rng(3).standard_normal((256, 8192 + S)).astype(np.float32)
    assert np.array_equal(abi.dedisperse_host(s, x), O.dedisperse_f32seq(x, t, 8192))


def test_rejects_short_rows_and_autotune():
    s = _survey(8192)
    plan = abi.Plan(s)
    din = abi.DeviceBuffer(256 * plan.min_ld * 4)
    dout = abi.DeviceBuffer(256 * 8192 * 4)
    with pytest.raises(abi.MdsmError):
        plan.dedisperse(din.ptr, plan.min_ld - 1, dout.ptr)
    x = np.random.default_rng(4).standard_normal((256, plan.min_ld)).astype(np.float32)
    din.upload(x)
    v, ms = plan.autotune(din.ptr, plan.min_ld, dout.ptr)
    _say(f"[autotune] best variant {v} {ms:.3f} ms")
    assert plan.variant == v
    plan.dedisperse(din.ptr, plan.min_ld, dout.ptr)
    abi.synchronize()
    t = _table(s)
    assert np.array_equal(dout.download(np.float32, (256, 8192)), O.dedisperse_f32seq(x, t, 8192))
