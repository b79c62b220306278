"""Pins the CPU oracle: RNG, delay tables, dedispersion, FP16 conversion,
normalisation statistics.  Every check is independent of the oracle's own
code path (published known answers, closed forms, high-precision evaluation,
numpy brute force) so that a dropped term, a wrong constant, a transposed
index or a wrong rounding mode fails.  CPU only."""
from decimal import Decimal, getcontext

import numpy as np
import pytest

import oracle as O

getcontext().prec = 50


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expect", [
    # Random123 known-answer vectors for philox4x32_10 (kat_vectors).
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert list(O.philox4x32_10(ctr, key)) == expect


def test_generator_is_standard_normal_and_places_pulses():
    from scipy import stats
    x = O.generate(7, 4, 1500.0, -0.29296875, 64e-6, 0, 50000)
    assert stats.kstest(x.reshape(-1), "norm").pvalue > 1e-3
    # counter-based: a sub-block equals the matching slice of a longer block
    y = O.generate(7, 4, 1500.0, -0.29296875, 64e-6, 1234, 777)
    assert np.array_equal(y, x[:, 1234:1234 + 777])
    # a pulse adds its amplitude exactly on the dispersed track
    fch1, foff, ts, dm = 1500.0, -0.29296875, 64e-6, 300.0
    base = O.generate(3, 64, fch1, foff, ts, 0, 4096)
    pul = O.generate(3, 64, fch1, foff, ts, 0, 4096, pulses=[(dm, 100, 5, 2.0)])
    diff = pul - base
    for c in range(64):
        f = Decimal(fch1) + Decimal(foff) * c
        dly = int((Decimal("4148.808") * Decimal(dm) * (1 / (f * f) - 1 / (Decimal(fch1) ** 2))
                   / Decimal(ts)).to_integral_value(rounding="ROUND_HALF_UP"))
        hit = np.nonzero(np.abs(diff[c]) > 1.0)[0]
        assert list(hit) == list(range(100 + dly, 105 + dly)), c


def test_bandpass_ranges():
    g, o = O.channel_bandpass(11, 1024)
    assert g.min() >= 0.5 and g.max() <= 2.0 and o.min() >= 0.0 and o.max() <= 10.0
    assert abs(g.mean() - 1.25) < 0.1 and abs(o.mean() - 5.0) < 0.5


# ---------------------------------------------------------------- delays
def _exact_delay(fch1, foff, c, tsamp, dm, k="4148.808"):
    f = Decimal(repr(fch1)) + Decimal(repr(foff)) * c
    ft = Decimal(repr(fch1))
    v = Decimal(k) * Decimal(repr(dm)) * (1 / (f * f) - 1 / (ft * ft)) / Decimal(repr(tsamp))
    return v


def test_delay_table_hand_values():
    # fch1 = 1000 MHz, channel 1 at 500 MHz: f^-2 - ftop^-2 = 3e-6 MHz^-2.
    # DM 1000, tsamp 1e-4 s: 4148.808*1000*3e-6/1e-4 = 124464.24 -> 124464.
    t = O.delay_table(1000.0, -500.0, 2, 1e-4, 0.0, 1000.0, 2)
    assert t.tolist() == [[0, 0], [0, 124464]]
    # The reference constant 4148.741601 with truncation gives 124462.248 -> 124462.
    r = O.delay_table_ref(1000.0, -500.0, 2, 1e-4, 0.0, 1000.0, 2)
    assert r.tolist() == [[0, 0], [0, 124462]]


@pytest.mark.parametrize("cfg", [
    (418.0, -0.078125, 256, 51.2e-6, 0.0, 0.1, 256),
    (1500.0, -0.29296875, 1024, 64e-6, 0.0, 0.2, 4096),
])
def test_delay_table_matches_high_precision(cfg):
    fch1, foff, C, ts, lo, step, M = cfg
    t = O.delay_table(*cfg)
    assert t.shape == (C, M)
    assert (t[:, 0] == 0).all() and (t[0, :] == 0).all()     # DM 0 and top channel
    assert (np.diff(t, axis=0) >= 0).all() and (np.diff(t, axis=1) >= 0).all()
    rng = np.random.default_rng(0)
    for c, d in zip(rng.integers(0, C, 300), rng.integers(0, M, 300)):
        v = _exact_delay(fch1, foff, int(c), ts, lo + step * int(d))
        # FP64 rounding only matters within 1e-9 of a .5 boundary
        frac = v - int(v)
        if abs(frac - Decimal("0.5")) > Decimal("1e-6"):
            assert t[c, d] == int(v.to_integral_value(rounding="ROUND_HALF_UP")), (c, d)


@pytest.mark.parametrize("cfg", [
    (418.0, -0.078125, 256, 51.2e-6, 0.0, 0.1, 256),
    (1500.0, -0.29296875, 1024, 64e-6, 0.0, 0.2, 4096),
    (1400.0, -0.5, 512, 100e-6, 3.0, 0.37, 999),
])
def test_legacy_delay_table_is_floor_of_the_float_shift(cfg):
    """The legacy table is floor(dm_shifts[c] * DM / tsamp): against the
    exact value of that product (the float dm_shifts and DM taken as given)
    it may differ only where float rounding reaches an integer boundary; and
    its operation order (DM / tsamp first) differs from the multibeam one on
    some cells of a large table."""
    fch1, foff, C, ts, lo, step, M = cfg
    t = O.delay_table_legacy(*cfg)
    r = O.delay_table_ref(*cfg)
    sh = O.dm_shifts_ref(fch1, foff, C)
    assert (t[0, :] == 0).all() and (np.diff(t, axis=0) >= 0).all() and (np.diff(t, axis=1) >= 0).all()
    rng = np.random.default_rng(1)
    ts32 = Decimal(float(np.float32(ts)))
    for c, d in zip(rng.integers(0, C, 400), rng.integers(0, M, 400)):
        dm = np.float32(np.float32(lo) + np.float32(d) * np.float32(step))
        v = Decimal(float(sh[c])) * Decimal(float(dm)) / ts32
        near = abs(v - v.to_integral_value()) <= v * Decimal("1e-6") + Decimal("1e-6")
        if not near:
            assert t[c, d] == int(v), (c, d, v)
    assert abs(t.astype(np.int64) - r).max() <= 1
    if C * M >= 1 << 20:
        assert (t != r).any()


def test_survey_maxshift_values():
    # maxshift S = D[C-1][M-1] at the two single-beam workloads
    t0 = O.delay_table(418.0, -0.078125, 256, 51.2e-6, 0.0, 0.1, 256)
    t1 = O.delay_table(1500.0, -0.29296875, 1024, 64e-6, 0.0, 0.2, 4096)
    assert int(t0[-1, -1]) == 1213 and int(t1[-1, -1]) == 13255


def test_reference_compat_helpers():
    s = O.dm_shifts_ref(418.0, -0.078125, 256)
    assert s[0] == 0.0 and (np.diff(s) > 0).all()
    # float32 restatement agrees with an independent numpy float32 evaluation
    F1 = np.float32(418.0) + np.float32(-0.078125) * np.arange(256, dtype=np.float32)
    F1d = F1.astype(np.float64)
    ref = (4148.741601 * ((1.0 / F1d / F1d) - (1.0 / 418.0 / 418.0))).astype(np.float32)
    assert np.array_equal(s, ref)
    # maxshift rounding: (x/256)*256+256, applied once per beam inside the loop
    one = O.maxshift_ref([418.0], [-0.078125], 256, 51.2e-6, 0.0, 0.1, 256)
    two = O.maxshift_ref([418.0, 418.0], [-0.078125, -0.078125], 256, 51.2e-6, 0.0, 0.1, 256)
    raw = int(np.float32(s[-1]) * np.float32(25.5) / np.float32(51.2e-6))
    assert one == (raw // 256) * 256 + 256 and two == one + 256
    # nsamp: explicit value kept unless below maxshift (then next power of two)
    assert O.calculate_nsamp_ref(65536, 1280, 256, 256, 1000) == 65536
    assert O.calculate_nsamp_ref(1000, 1280, 256, 256, 1000) == 2048


# ---------------------------------------------------------------- dedispersion
def _small_case(C=32, N=256, M=24, seed=0):
    table = O.delay_table(418.0, -0.078125 * 8, C, 51.2e-6, 0.0, 1.0, M)
    S = int(table.max())
    x = np.random.default_rng(seed).standard_normal((C, N + S)).astype(np.float32)
    return x, table, N


def test_known_answer_channel_index():
    # in[c][t] = c  =>  every output equals sum_c c (prototype known answer)
    x, table, N = _small_case()
    C = x.shape[0]
    x[:] = np.arange(C, dtype=np.float32)[:, None]
    expect = C * (C - 1) / 2
    assert (O.dedisperse_f32seq(x, table, N) == expect).all()
    assert (O.dedisperse_f64(x, table, N) == expect).all()


def test_constant_and_dm0():
    x, table, N = _small_case()
    C = x.shape[0]
    x2 = np.full_like(x, 0.1)
    out = O.dedisperse_f64(x2, table, N)
    assert (out == C * float(np.float32(0.1))).all()
    # DM 0 row is the plain channel sum
    out = O.dedisperse_f64(x, table, N)
    assert np.allclose(out[0], x[:, :N].astype(np.float64).sum(0), rtol=0, atol=1e-12)


def test_impulse_on_dm_track():
    x, table, N = _small_case()
    C, M = table.shape
    x[:] = 0
    dstar, t0 = 17, 40
    for c in range(C):
        x[c, t0 + table[c, dstar]] = 1.0
    out = O.dedisperse_f64(x, table, N)
    assert out[dstar, t0] == C
    # every other output counts the channels whose track crosses the impulse
    cnt = np.zeros((M, N))
    for c in range(C):
        for d in range(M):
            t = t0 + table[c, dstar] - table[c, d]
            if 0 <= t < N:
                cnt[d, t] += 1
    assert np.array_equal(out, cnt)


def test_dedisperse_vs_numpy_gather():
    x, table, N = _small_case(seed=3)
    C, M = table.shape
    idx = np.arange(N)[None, None, :] + table[:, :, None]          # [C][M][N]
    g = x[np.arange(C)[:, None, None], idx].astype(np.float64)
    ref64 = g.sum(axis=0)
    out64 = O.dedisperse_f64(x, table, N)
    assert np.allclose(out64, ref64, rtol=0, atol=1e-9)
    # FP32 in channel order: sequential float32 adds, bit for bit
    acc = np.zeros((M, N), np.float32)
    for c in range(C):
        acc += x[c][idx[c]]
    assert np.array_equal(O.dedisperse_f32seq(x, table, N), acc)
    assert np.abs(acc - out64).max() < 1e-4


def test_f16_conversion_matches_numpy():
    rng = np.random.default_rng(5)
    v = np.concatenate([
        rng.standard_normal(100000).astype(np.float32) * 4,
        (rng.standard_normal(20000) * 1e-6).astype(np.float32),        # subnormal halves
        np.array([65504, 65519.99, 65520, 1e9, -1e9, 0.0, -0.0, 2.0 ** -25, 2.0 ** -24,
                  3 * 2.0 ** -26, np.inf, -np.inf], np.float32),
        (np.arange(2048, dtype=np.float32) + 0.5) * 2.0 ** -10,          # ties
    ]).astype(np.float32)
    h = O.f32_to_f16(v)
    assert np.array_equal(h, v.astype(np.float16).view(np.uint16))
    back = O.f16_to_f32(h)
    assert np.array_equal(back, v.astype(np.float16).astype(np.float32))


def test_dedisperse_f16_and_u8():
    x, table, N = _small_case(seed=9)
    C, M = table.shape
    h = O.f32_to_f16(x)
    xf = O.f16_to_f32(h)
    assert np.array_equal(O.dedisperse_f16(h, table, N), O.dedisperse_f32seq(xf, table, N))
    u = np.random.default_rng(2).integers(0, 256, x.shape).astype(np.uint8)
    idx = np.arange(N)[None, None, :] + table[:, :, None]
    s = u[np.arange(C)[:, None, None], idx].astype(np.int64).sum(0)
    expect = (s.astype(np.float64) * (8.0 / 255.0) + C * -4.0).astype(np.float32)
    assert np.array_equal(O.dedisperse_u8(u, table, N), expect)


def _pairs_numpy(h, table, N):
    """Channel-pair sums restated with numpy: the exact pair sum in float64
    (two halves never need more than 41 significant bits), rounded straight
    to float16, then float32 adds in pair order."""
    C, M = table.shape
    idx = np.arange(N)[None, None, :] + table[:, :, None]
    g = h.view(np.float16)[np.arange(C)[:, None, None], idx].astype(np.float64)  # [C][M][N]
    acc = np.zeros((M, N), np.float32)
    for c in range(0, C, 2):
        s = g[c] + g[c + 1] if c + 1 < C else g[c]
        acc += s.astype(np.float16).astype(np.float32)
    return acc


@pytest.mark.parametrize("C", [32, 31, 1, 2])
def test_dedisperse_f16_pairs_vs_numpy(C):
    x, table, N = _small_case(C=max(C, 2), seed=11)
    x, table = x[:C], table[:C]
    h = O.f32_to_f16(x * 3)
    assert np.array_equal(O.dedisperse_f16_pairs(h, table, N), _pairs_numpy(h, table, N))


def test_dedisperse_f16_pairs_known_answers():
    # two channels, DM 0 only: ties of the FP16 pair sum go to even, where
    # the FP32 channel-order sum of the same halves keeps the low bit
    table = np.zeros((2, 1), np.int32)
    cases = [(2048.0, 1.0, 2048.0, 2049.0),               # ulp 2 at 2048: tie -> 2048
             (1.0, 2.0 ** -11, 1.0, 1.0 + 2.0 ** -11),    # half an ulp of 1: tie -> 1
             (1.0, 3 * 2.0 ** -11, 1.0 + 2.0 ** -9, 1.0 + 3 * 2.0 ** -11),  # 1.5 ulp -> 2 ulp
             (2.0 ** -24, 2.0 ** -24, 2.0 ** -23, 2.0 ** -23),  # subnormals: exact
             (65504.0, 16.0, np.inf, 65520.0)]            # tie at the overflow edge -> inf
    for a, b, pair, seq in cases:
        h = O.f32_to_f16(np.array([[a], [b]], np.float32))
        assert O.dedisperse_f16_pairs(h, table, 1)[0, 0] == np.float32(pair), (a, b)
        assert O.dedisperse_f16(h, table, 1)[0, 0] == np.float32(seq), (a, b)
    # integer-valued halves with exact pair sums: equal to the exact sum
    x, table, N = _small_case(C=33, seed=4)
    xi = np.round(x * 8).astype(np.float32)
    h = O.f32_to_f16(xi)
    idx = np.arange(N)[None, None, :] + table[:, :, None]
    exact = xi[np.arange(33)[:, None, None], idx].astype(np.int64).sum(0)
    assert np.array_equal(O.dedisperse_f16_pairs(h, table, N), exact.astype(np.float32))


# ---------------------------------------------------------------- statistics
def test_channel_stats_and_normalise():
    rng = np.random.default_rng(1)
    raw = (rng.standard_normal((8, 5000)) * 3 + 7).astype(np.float32)
    m, s = O.channel_stats(raw)
    assert np.allclose(m, raw.astype(np.float64).mean(1), rtol=1e-13)
    assert np.allclose(s, raw.astype(np.float64).std(1), rtol=1e-10)
    # integer-valued data: sums are exact, so the moments are exact
    ri = rng.integers(-50, 50, (4, 3001)).astype(np.float32)
    mi, si = O.channel_stats(ri)
    ex = ri.astype(np.float64)
    assert np.array_equal(mi, ex.sum(1) / ex.shape[1])
    assert np.array_equal(si, np.sqrt((ex * ex).sum(1) / ex.shape[1] - mi * mi))
    n32 = O.normalise(raw, m, s, "f32")
    assert np.abs(n32.mean(1)).max() < 1e-5 and np.abs(n32.std(1) - 1).max() < 1e-5
    n16 = O.normalise(raw, m, s, "f16")
    assert np.array_equal(n16, n32.astype(np.float16).view(np.uint16))
    u8 = O.normalise(raw, m, s, "u8")
    v = np.clip((raw.astype(np.float64) - m[:, None]) * (1.0 / s[:, None]), -4, 4)
    assert np.array_equal(u8, np.floor((v + 4) * 31.875 + 0.5).astype(np.uint8))
    deq = u8.astype(np.float64) * (8.0 / 255.0) - 4.0
    assert np.abs(deq - v).max() <= 4.0 / 255.0 + 1e-12
