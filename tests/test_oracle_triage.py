"""Pins of the oracle's triage-data restatements: RFI injection and event
snippets on hand-computed cases."""
import numpy as np

import oracle


def _item(chan, kind, start, length, amp, period=1.0):
    a = np.zeros(1, oracle.RFI_ITEM_DTYPE)
    a[0] = (chan, kind, start, length, amp, period)
    return a


def test_burst_and_mains():
    raw = np.zeros((4, 100), np.float32)
    gain = np.array([1, 2, 3, 4], np.float32)
    out = oracle.rfi_inject(raw, _item(2, 0, 10, 5, 1.5), gain, 0)
    assert (out[2, 10:15] == 4.5).all() and out.sum() == 4.5 * 5
    out = oracle.rfi_inject(raw, _item(1, 1, 0, 100, 2.0, 40.0), gain, 0)
    np.testing.assert_allclose(out[1, [0, 10, 20, 30]], [2.0, 4.0, 2.0, 0.0], atol=1e-6)
    # the item is clipped to the row length
    out = oracle.rfi_inject(raw, _item(0, 0, 95, 50, 1.0), gain, 0)
    assert out[0, 95:].sum() == 5


def test_telegraph_blocks():
    raw = np.zeros((1, 64 * 4000), np.float32)
    out = oracle.rfi_inject(raw, _item(0, 2, 0, raw.shape[1], 1.0, 64.0), np.ones(1, np.float32), 9)
    blocks = out[0].reshape(-1, 64)
    assert ((blocks == 0).all(1) | (blocks == 1).all(1)).all()  # whole blocks on or off
    frac = (blocks[:, 0] == 1).mean()
    assert abs(frac - 0.25) < 0.03
    again = oracle.rfi_inject(raw, _item(0, 2, 0, raw.shape[1], 1.0, 64.0), np.ones(1, np.float32), 9)
    np.testing.assert_array_equal(out, again)


def test_snippet_planes():
    ndms, nsamp, nchans, ld = 10, 50, 8, 60
    out = (np.arange(ndms)[:, None] * 100 + np.arange(nsamp)[None, :]).astype(np.float32)
    store = np.ones((nchans, ld), np.float32)
    table = np.zeros((nchans, ndms), np.int32)
    table[:, 4] = np.arange(nchans)  # event DM 4: delay c samples in channel c
    store[np.arange(nchans), 20 + np.arange(nchans)] = 5.0  # a pulse on that track at t = 20
    sn = oracle.snippet(out, store, table, 4, 20, SD=4, ST=8, SB=4, clip=1e4)
    # plane 0: rows dm 2..5, samples 16..23, stored as BF16
    want = np.vectorize(oracle.bf16_round)(out[2:6, 16:24].astype(np.float64))
    np.testing.assert_array_equal(sn[0], want)
    # plane 1: 4 subbands of 2 channels; sample 20 sees the pulse in every channel
    assert (sn[1][:, 4] == 10.0).all()
    assert (sn[1][:, [0, 1, 2, 3, 5, 6, 7]] == 2.0).all()
    # outside the arrays: zeros; clipping
    sn = oracle.snippet(out, store, table, 0, 2, SD=4, ST=8, SB=4, clip=30.0)
    assert (sn[0][:2] == 0).all() and (sn[0][:, :2] == 0).all()
    assert sn[0].max() == 30.0
