"""Triage data on the GPU: RFI injection and snippet extraction against the
oracle (bit exact), the synthetic search stream (labels, statistics), and
the learned triage against the plain threshold + DBSCAN path on held-out
chunks: at an operating threshold picked on a separate validation stream,
it keeps the plain path's recall on injected pulses (within 0.03) at no more
than half of its false positives."""
import numpy as np
import pytest

import oracle
from mdsm_amd import abi
from mdsm_amd.triage_train import THRESHOLDS, operating_threshold, recall_fp

pytestmark = pytest.mark.gpu


def test_inject_rfi_matches_oracle():
    rng = np.random.default_rng(1)
    C, n = 16, 5000
    raw = rng.standard_normal((C, n)).astype(np.float32)
    gain = rng.uniform(0.5, 2, C).astype(np.float32)
    items = [abi.RfiItem(3, abi.RFI_BURST, 100, 700, 12.5, 1.0),
             abi.RfiItem(7, abi.RFI_MAINS, 0, n, 6.0, 312.5),
             abi.RfiItem(7, abi.RFI_TELEGRAPH, 50, 4000, 9.0, 64.0),  # same channel: applied in order
             abi.RfiItem(15, abi.RFI_BURST, 4990, 100, 3.0, 1.0)]
    d_raw = abi.DeviceBuffer.from_numpy(raw)
    d_gain = abi.DeviceBuffer.from_numpy(gain)
    abi.inject_rfi(items, d_gain.ptr, 42, d_raw.ptr, n, n)
    abi.synchronize()
    got = d_raw.download(np.float32, (C, n))
    arr = np.array([(i.chan, i.kind, i.start, i.len, i.amp, i.period) for i in items], oracle.RFI_ITEM_DTYPE)
    want = oracle.rfi_inject(raw, arr, gain, 42)
    # the 50 Hz sinusoid comes from IEEE operations only on both sides
    np.testing.assert_array_equal(got, want)


def test_snippets_match_oracle():
    rng = np.random.default_rng(2)
    ndms, nsamp, nchans, ld = 100, 2000, 128, 2400
    out = (rng.standard_normal((ndms, nsamp)) * 11).astype(np.float32)
    store = rng.standard_normal((nchans, ld)).astype(np.float32)
    table = np.sort(rng.integers(0, 400, (nchans, ndms)), axis=1).astype(np.int32)
    ev_dm = np.array([0, 50, 99, 10, 70], np.int32)
    ev_t = np.array([0, 1000, 1999, 5, 1900], np.int32)
    d_out, d_store, d_tab = (abi.DeviceBuffer.from_numpy(a) for a in (out, store, table))
    d_snip = abi.DeviceBuffer(len(ev_dm) * 2 * 64 * 256 * 2)
    s0, s1 = 1 / np.sqrt(nchans), 1 / np.sqrt(nchans / 64)
    abi.triage_snippets(d_out.ptr, ndms, nsamp, d_store.ptr, ld, nchans, d_tab.ptr, ev_dm, ev_t, d_snip.ptr,
                        scale0=s0, scale1=s1, clip=3.0)
    abi.synchronize()
    got = abi.from_bf16_bits(d_snip.download(np.uint16, (len(ev_dm), 2, 64, 256)))
    for i in range(len(ev_dm)):
        want = oracle.snippet(out, store, table, int(ev_dm[i]), int(ev_t[i]), scale0=s0, scale1=s1, clip=3.0)
        np.testing.assert_array_equal(got[i], want, err_msg=f"event {i}")


def _config(npulses=20):
    C, M, N = 256, 512, 1 << 15
    s = abi.Survey.make(1500.0, -300.0 / C, C, 64e-6, 0.0, 0.2, M, N)
    return abi.TriageConfig.make(s, npulses=npulses, dm=(10.0, 100.0), width=(1.0, 32.0), snr=(6.0, 30.0),
                                 n_spikes=8, n_bursts=6, n_mains=2, n_telegraph=2, rfi=(8.0, 40.0))


def _stream(seed, nchunks, max_events):
    sim = abi.TriageSim(_config(), seed)
    d_snip = abi.DeviceBuffer(max_events * 2 * 64 * 256 * 2)
    ev, n = sim.run(nchunks, d_snip.ptr, max_events)
    stats = sim.stats.as_dict()
    sim.close()
    # every cluster of every chunk is an event (a full buffer would drop some
    # and bias the recall comparison)
    assert n < max_events and stats["clusters"] == n
    return d_snip, ev[:n], stats


def test_search_stream_statistics():
    d_snip, ev, st = _stream(11, 4, 4000)
    print(st, flush=True)
    assert st["chunks"] == 4 and st["injected"] == 80
    assert st["injected_found"] >= 0.8 * st["injected"]
    assert st["clusters"] == len(ev) and st["false_clusters"] == (ev["label"] == 0).sum()
    assert set(np.unique(ev["label"])) == {0, 1}
    x = abi.from_bf16_bits(d_snip.download(np.uint16, (len(ev), 2, 64, 256)))
    assert np.isfinite(x).all()
    # a detected pulse is brightest near the centre of its DM-time snippet
    strong = np.flatnonzero((ev["label"] == 1) & (ev["snr"] > 15))
    assert len(strong)
    centre = x[strong, 0, 24:40, 96:160].max(axis=(1, 2))
    edge = np.abs(x[strong, 0, :, :32]).max(axis=(1, 2))
    assert (centre > edge).mean() > 0.9


def _predict(net, d, n):
    d_prob = abi.DeviceBuffer(4 * n)
    abi.predict(net, d.ptr, n, d_prob.ptr)
    abi.synchronize()
    return d_prob.download(np.float32, (n,))


def test_learned_triage_beats_plain_false_positives():
    d_tr, ev_tr, st_tr = _stream(1, 40, 20000)
    d_va, ev_va, st_va = _stream(3, 10, 20000)  # validation: picks the operating threshold
    d_te, ev_te, st_te = _stream(2, 15, 20000)  # held out: evaluated
    lab = abi.DeviceBuffer.from_numpy(ev_tr["label"].astype(np.int32))
    net = abi.TriageNet(64, 256, 256)
    net.init(3)
    for ep in range(8):
        loss = abi.train_epoch(net, d_tr.ptr, lab.ptr, len(ev_tr), 128, 1e-3, 100 + ep)
    print("final training loss", loss, flush=True)
    # operating point from the validation stream: the lowest threshold whose
    # false positives stay under a quarter of the plain path's there (a lower
    # threshold can only keep more of the injected pulses)
    pv = _predict(net, d_va, len(ev_va))
    print("validation (recall, fp) per threshold:", {t: recall_fp(pv, ev_va, st_va, t) for t in THRESHOLDS},
          "plain fp", st_va["false_clusters"], flush=True)
    thr = operating_threshold(pv, ev_va, st_va)
    prob = _predict(net, d_te, len(ev_te))
    net.close()
    recall, fp = recall_fp(prob, ev_te, st_te, thr)
    plain_recall = st_te["injected_found"] / st_te["injected"]
    print(f"threshold {thr}: plain recall {plain_recall:.3f} fp {st_te['false_clusters']}; "
          f"cnn recall {recall:.3f} fp {fp}")
    assert recall >= plain_recall - 0.03
    assert fp <= 0.5 * st_te["false_clusters"]
