"""Pins of the oracle's mu-law quantiser (multibeam_dedispersion_writer.cpp:
21-118): closed-form levels, monotonicity, the complex byte layout."""
import numpy as np

import oracle


def test_power_levels():
    # 0 -> the lowest level: (0*Q + Q/2) * 2^bits = 0.5 -> 0
    assert oracle.mulaw_power(np.zeros(4, np.float32), 8, 255, 1.0).tolist() == [0] * 4
    # the table's end: d ~ 1 - 2e-5 -> level 255 (255*Q + Q/2) * 256 = 255.5 -> 255
    assert oracle.mulaw_power(np.array([1.0, 5.0], np.float32), 8, 255, 1.0).tolist() == [255, 255]
    # one value by hand: x = max/2 -> k = 16384, d = log10(1 + 255/2)/log10(256)
    d = np.log10(np.float32(1 + 16384 * 255 / 32768.0)) / np.log10(256.0)
    want = int((int(d * 256) / 256 + 1 / 512) * 256)
    assert oracle.mulaw_power(np.array([0.5], np.float32), 8, 255, 1.0)[0] == want


def test_power_monotone_and_compressive():
    x = np.linspace(0, 1, 5000, dtype=np.float32)
    y = oracle.mulaw_power(x, 6, 100, 1.0).astype(int)
    assert (np.diff(y) >= 0).all()
    # logarithmic: the lower half of the input gets more than half the levels
    assert y[len(x) // 2] > 32


def test_complex_layout():
    pairs = np.array([[0, 0], [-5, 7], [32767, -32768], [100, -1], [-32000, 32000]], np.int16)
    got = oracle.mulaw_complex(pairs, 3, 255)
    assert got[0] == 0x00
    assert got[1] == 0x80  # re negative, both magnitudes in the lowest level
    assert got[2] == 0x7F  # re top level, im negative (clamped -32768) top level
    assert got[3] & 0x0F == 0x08 and got[3] & 0x80 == 0
    assert got[4] >> 4 == 0xF and got[4] & 0x0F == 0x7  # 0x80 | 0x70, 0x07
