"""Train and evaluate the learned triage on synthetic search streams.

The plain search (bandpass normalisation, brute-force dedispersion,
per-block FP64 statistics, boxcar detection at 6 sigma, 2-D DBSCAN with
eps=3 / minPts=4, cluster peaks) runs over seeded chunks with injected
dispersed pulses (DM 10-800, width 1-64 samples, S/N 6-30) and synthetic
RFI (zero-DM broadband spikes, narrowband bursts, 50 Hz-modulated and
on/off channels).  Every cluster peak becomes an event with two BF16
snippets (64 DM trials x 256 samples; 64 subbands x 256 samples
dedispersed at the event DM) and a label (1 when it matches an injected
pulse).  The network trains on the training stream with the library's own
MFMA forward/backward kernels and is evaluated on held-out streams against
the plain path: recall on injected pulses and false-positive clusters.

    python -m mdsm_amd.triage_train --events 200000 --epochs 6 --json out.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi

SNIP = (64, 256)
ROW = 2 * SNIP[0] * SNIP[1]

# Training stream: the full single-beam survey (1024 channels over 300 MHz
# from 1500 MHz, 64 us, 4096 DM trials of 0.2), 2^17-sample chunks.
TRAIN_SURVEY = dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.2,
                    ndms=4096, nsamp=1 << 17)
# Held-out streams: beams shaped like the uneven multi-beam batch
# (512/1024 channels, DM steps 0.1-0.5, 1024-4096 trials).
EVAL_SURVEYS = [
    dict(fch1=1500.0, foff=-300.0 / 512, nchans=512, tsamp=64e-6, lowdm=0.0, dmstep=0.4, ndms=2048,
         nsamp=1 << 17),
    dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.2, ndms=4096,
         nsamp=1 << 17),
    dict(fch1=1500.0, foff=-0.29296875, nchans=1024, tsamp=64e-6, lowdm=0.0, dmstep=0.5, ndms=1700,
         nsamp=1 << 17),
]


def make_config(sv, npulses=40, n_spikes=15, n_bursts=10, n_mains=3, n_telegraph=3):
    s = abi.Survey.make(sv["fch1"], sv["foff"], sv["nchans"], sv["tsamp"], sv["lowdm"], sv["dmstep"],
                        sv["ndms"], sv["nsamp"])
    dm_hi = min(800.0, sv["lowdm"] + sv["dmstep"] * (sv["ndms"] - 1))
    return abi.TriageConfig.make(s, npulses=npulses, dm=(10.0, dm_hi), width=(1.0, 64.0), snr=(6.0, 30.0),
                                 n_spikes=n_spikes, n_bursts=n_bursts, n_mains=n_mains,
                                 n_telegraph=n_telegraph, rfi=(8.0, 40.0), snip=SNIP)


def build_stream(cfg, seed, max_events, max_chunks=100000, log=print):
    """Events + device snippets from chunks of one stream until max_events."""
    sim = abi.TriageSim(cfg, seed)
    d_snip = abi.DeviceBuffer(max_events * ROW * 2)
    events = np.zeros(max_events, abi.EVENT_DTYPE)
    n, t0 = 0, time.time()
    while n < max_events and sim.stats.chunks < max_chunks:
        _, n = sim.run(min(50, max_chunks - sim.stats.chunks), d_snip.ptr, max_events, events, n)
        log(f"  stream seed {seed}: {sim.stats.chunks} chunks, {n} events, {time.time() - t0:.0f} s")
    stats = sim.stats.as_dict()
    sim.close()
    return d_snip, events[:n], stats, time.time() - t0


THRESHOLDS = (0.9, 0.5, 0.3, 0.2, 0.1, 0.05, 0.02, 0.01, 0.005, 0.002, 0.001, 0.0005, 0.0002, 0.0001,
              5e-5, 2e-5, 1e-5, 5e-6, 2e-6, 1e-6, 5e-7, 2e-7, 1e-7, 1e-8)


def recall_fp(prob, events, stats, t):
    """Recall on injected pulses and false-positive count of the events kept at prob > t."""
    keep = prob > t
    hits = keep & (events["label"] == 1)
    found = len(set(zip(events["chunk"][hits].tolist(), events["pulse"][hits].tolist())))
    return found / max(stats["injected"], 1), int(((events["label"] == 0) & keep).sum())


def recall_fp_rule(prob, events, stats, t, snr_keep):
    """As recall_fp, for the rule that keeps an event when the network's
    probability exceeds t or the plain path's S/N reaches snr_keep."""
    keep = (prob > t) | (events["snr"] >= snr_keep)
    hits = keep & (events["label"] == 1)
    found = len(set(zip(events["chunk"][hits].tolist(), events["pulse"][hits].tolist())))
    return found / max(stats["injected"], 1), int(((events["label"] == 0) & keep).sum())


SNR_KEEP = (1e9, 40.0, 30.0, 25.0, 20.0, 16.0, 14.0, 12.0, 10.0, 9.0, 8.0)


def hybrid_rule(prob, events, stats, t, fp_fraction=0.5):
    """The lowest S/N rescue level whose false positives, together with the
    network's at threshold t, stay under fp_fraction of the plain path's on
    a validation stream (1e9: no rescue)."""
    budget = fp_fraction * stats["false_clusters"]
    ok = [x for x in SNR_KEEP if recall_fp_rule(prob, events, stats, t, x)[1] <= budget]
    return min(ok) if ok else SNR_KEEP[0]


def recall_fp_wide(prob, events, stats, t, wide_snr, wide_k=6):
    """As recall_fp, for the rule that keeps an event when the network's
    probability exceeds t or the event is found at the widest boxcar
    (width_log2 >= wide_k) with S/N >= wide_snr."""
    keep = (prob > t) | ((events["width_log2"] >= wide_k) & (events["snr"] >= wide_snr))
    hits = keep & (events["label"] == 1)
    found = len(set(zip(events["chunk"][hits].tolist(), events["pulse"][hits].tolist())))
    return found / max(stats["injected"], 1), int(((events["label"] == 0) & keep).sum())


WIDE_SNR_KEEP = (1e9, 12.0, 10.0, 9.0, 8.0, 7.5, 7.0, 6.5, 6.0)


def wide_rule(prob, events, stats, fp_fraction=0.5):
    """(threshold, widest-boxcar S/N rescue level) of the highest recall on a
    validation stream whose false positives stay under fp_fraction of the
    plain path's (ties: fewer false positives)."""
    budget = fp_fraction * stats["false_clusters"]
    best = None
    for t in THRESHOLDS:
        for w in WIDE_SNR_KEEP:
            r, f = recall_fp_wide(prob, events, stats, t, w)
            if f <= budget and (best is None or (r, -f) > (best[2], -best[3])):
                best = (t, w, r, f)
    return (best[0], best[1]) if best else (THRESHOLDS[0], WIDE_SNR_KEEP[0])


def missed_pulses(prob, events, t):
    """Features of the best event (highest S/N) of every injected pulse that
    the plain path found but no event of which the network keeps at t."""
    lab = events["label"] == 1
    keys = set(zip(events["chunk"][lab].tolist(), events["pulse"][lab].tolist()))
    kept = set(zip(events["chunk"][lab & (prob > t)].tolist(), events["pulse"][lab & (prob > t)].tolist()))
    out = []
    for ch, pu in sorted(keys - kept):
        sel = np.flatnonzero(lab & (events["chunk"] == ch) & (events["pulse"] == pu))
        b = sel[np.argmax(events["snr"][sel])]
        out.append({"chunk": ch, "pulse": pu, "snr": float(events["snr"][b]),
                    "width_log2": int(events["width_log2"][b]), "dm_idx": int(events["dm_idx"][b]),
                    "sample": int(events["sample"][b]), "cluster_size": int(events["cluster_size"][b]),
                    "prob": float(prob[b]), "events": int(len(sel))})
    return out


def operating_threshold(prob, events, stats, fp_fraction=0.25):
    """Lowest threshold whose false positives stay under fp_fraction of the
    plain path's on a validation stream (lower thresholds keep more pulses)."""
    budget = fp_fraction * stats["false_clusters"]
    ok = [t for t in THRESHOLDS if recall_fp(prob, events, stats, t)[1] <= budget]
    return min(ok) if ok else max(THRESHOLDS)


def build_mixed(cfgs, seeds, max_events, log=print):
    """One training stream drawn from several surveys: equal shares of
    max_events from each (cfg, seed), appended into one snippet buffer."""
    d_snip = abi.DeviceBuffer(max_events * ROW * 2)
    events = np.zeros(max_events, abi.EVENT_DTYPE)
    n, t0, stats = 0, time.time(), []
    for i, (cfg, seed) in enumerate(zip(cfgs, seeds)):
        quota = max_events * (i + 1) // len(cfgs)
        sim = abi.TriageSim(cfg, seed)
        while n < quota:
            _, n = sim.run(50, d_snip.ptr, quota, events, n)
            log(f"  stream seed {seed}: {sim.stats.chunks} chunks, {n} events, {time.time() - t0:.0f} s")
        stats.append(sim.stats.as_dict())
        sim.close()
    total = {k: sum(st[k] for st in stats) for k in stats[0]}
    return d_snip, events[:n], total, time.time() - t0


def oversample_positives(d_snip, events, k, k_wide=1, log=print):
    """A training buffer in which every labelled-pulse event appears k times
    (k * k_wide times for the widest boxcars, 2^5 and 2^6 samples, whose
    weak members the network otherwise rates as RFI), rows copied on the
    device, and every other event once: weights the pulse class without
    touching the loss kernel.  Returns (buffer, labels, n)."""
    lab = events["label"] == 1
    reps = np.where(lab, k * np.where(events["width_log2"] >= 5, k_wide, 1), 1).astype(np.int64)
    extra = np.repeat(np.arange(len(events)), reps - 1)
    n0 = len(events)
    n = n0 + len(extra)
    d = abi.DeviceBuffer(n * ROW * 2)
    row = ROW * 2
    abi.check(abi.lib().mdsm_memcpy_dtod(d.ptr, d_snip.ptr, n0 * row, None))
    for j, i in enumerate(extra.tolist()):
        abi.check(abi.lib().mdsm_memcpy_dtod(d.ptr + (n0 + j) * row, d_snip.ptr + i * row, row, None))
    abi.synchronize()
    labels = np.concatenate([events["label"].astype(np.int32), events["label"][extra].astype(np.int32)])
    log(f"  oversampled {int(lab.sum())} pulse events (x{k}, widest boxcars x{k * k_wide}): {n} training rows")
    return d, labels, n


def batch_beam_surveys(seed=2026):
    """The beam shapes of the 32-beam uneven batch (nchans 512/1024, ndms
    1024-4096, dmstep 0.1-0.5, 1500 MHz top over 300 MHz, 64 us), drawn as
    the benchmark draws them, each searched in 2^17-sample chunks."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(32):
        C = int(rng.choice([512, 1024]))
        N = int(rng.integers(1 << 17, (1 << 20) + 1))  # (the batch's length; chunks are 2^17 here)
        M = int(rng.integers(1024, 4097))
        step = float(rng.uniform(0.1, 0.5))
        for _ in range(20):  # (the batch's pulse draws, to keep the generator in step)
            rng.uniform(10.0, 800.0), rng.uniform(0, np.log(64)), rng.uniform(8, 30), rng.integers(0, N - 64)
        out.append(dict(fch1=1500.0, foff=-300.0 / C, nchans=C, tsamp=64e-6, lowdm=0.0, dmstep=step,
                        ndms=M, nsamp=1 << 17))
    return out


def score(prob, events, stats, thresholds):
    """Plain vs learned triage on one held-out stream."""
    out = {"plain": {"recall": stats["injected_found"] / stats["injected"], "false_positives":
                     stats["false_clusters"], "clusters": stats["clusters"]}}
    for t in thresholds:
        recall, fp = recall_fp(prob, events, stats, t)
        out[f"cnn@{t}"] = {"recall": recall, "false_positives": fp, "kept": int((prob > t).sum())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--events", type=int, default=200000)
    ap.add_argument("--eval-chunks", type=int, default=40)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--threshold", type=float, default=None,
                    help="operating threshold (default: picked on a validation stream)")
    ap.add_argument("--single-survey", action="store_true",
                    help="train on the single-beam survey only (default: three beam shapes)")
    ap.add_argument("--pos-weight", type=int, default=1,
                    help="pulse events repeated this many times in the training set (class weight)")
    ap.add_argument("--wide-weight", type=int, default=1,
                    help="extra repetition of pulse events found at the two widest boxcars")
    ap.add_argument("--batch-beam-chunks", type=int, default=4,
                    help="chunks per beam of the held-out stream over the 32 batch beam shapes (0: skip)")
    ap.add_argument("--save", default=None, help="write trained parameters (.npy, canonical layout)")
    ap.add_argument("--json", default=None, help="write the results as JSON")
    args = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    res = {"config": vars(args)}
    if args.single_survey:
        log(f"building the training stream ({args.events} events, one survey)")
        d_tr, ev_tr, st_tr, t_sim = build_stream(make_config(TRAIN_SURVEY), args.seed, args.events, log=log)
    else:
        # equal shares from the single-beam survey and the two other beam
        # shapes of the held-out streams (other seeds than the held-out ones)
        shapes = [TRAIN_SURVEY, EVAL_SURVEYS[0], EVAL_SURVEYS[2]]
        log(f"building the training stream ({args.events} events from {len(shapes)} beam shapes)")
        d_tr, ev_tr, st_tr, t_sim = build_mixed([make_config(sv) for sv in shapes],
                                                [args.seed + k for k in range(len(shapes))],
                                                args.events, log=log)
    res["train"] = {"events": len(ev_tr), "pulses": int(ev_tr["label"].sum()), "stats": st_tr,
                    "sim_seconds": t_sim}
    n_tr = len(ev_tr)
    if args.pos_weight > 1:
        d_tr, labels, n_tr = oversample_positives(d_tr, ev_tr, args.pos_weight, args.wide_weight, log=log)
    else:
        labels = ev_tr["label"].astype(np.int32)
    res["train"]["rows"] = n_tr
    res["train"]["pos_weight"] = args.pos_weight
    res["train"]["wide_weight"] = args.wide_weight
    lab = abi.DeviceBuffer.from_numpy(labels)
    net = abi.TriageNet(SNIP[0], SNIP[1], args.batch)
    net.init(args.seed)
    res["epochs"] = []
    for ep in range(args.epochs):
        t = time.time()
        lr = args.lr * (0.5 ** (ep // 2))
        loss = abi.train_epoch(net, d_tr.ptr, lab.ptr, n_tr, args.batch, lr, args.seed * 1000 + ep)
        abi.synchronize()
        dt = time.time() - t
        res["epochs"].append({"loss": loss, "seconds": dt, "events_per_s": n_tr / dt})
        log(f"epoch {ep}: loss {loss:.4f}, {n_tr / dt:.0f} events/s")
    del d_tr
    # operating point from a validation stream of the single-beam survey
    d_va, ev_va, st_va, _ = build_stream(make_config(TRAIN_SURVEY), args.seed + 50, 2000 * args.eval_chunks,
                                         max_chunks=max(args.eval_chunks // 2, 1), log=log)
    if len(ev_va) >= 2000 * args.eval_chunks:
        raise RuntimeError("validation stream filled its event buffer; raise the capacity")
    d_prob = abi.DeviceBuffer(4 * max(len(ev_va), 1))
    abi.predict(net, d_va.ptr, len(ev_va), d_prob.ptr)
    abi.synchronize()
    pv = d_prob.download(np.float32, (len(ev_va),))
    thr = args.threshold if args.threshold is not None else operating_threshold(pv, ev_va, st_va)
    # a second operating point at the target bound itself (half the
    # plain path's false positives on the validation stream)
    thr_half = operating_threshold(pv, ev_va, st_va, fp_fraction=0.5)
    res["validation"] = score(pv, ev_va, st_va, THRESHOLDS)
    res["threshold"] = thr
    res["threshold_half"] = thr_half
    # the hybrid rule at the half budget: the network at thr_half, or the
    # plain path's S/N at the lowest rescue level the budget allows
    snr_keep = hybrid_rule(pv, ev_va, st_va, thr_half)
    res["hybrid_snr_keep"] = snr_keep
    log(f"hybrid rule: prob > {thr_half} or S/N >= {snr_keep}")
    # the widest-boxcar rescue at the half budget: threshold and rescue level
    # chosen together on the validation stream
    thr_wide, wide_snr = wide_rule(pv, ev_va, st_va)
    res["wide_rule"] = {"threshold": thr_wide, "wide_snr_keep": wide_snr}
    log(f"wide rule: prob > {thr_wide} or (widest boxcar and S/N >= {wide_snr})")
    log(f"operating thresholds {thr} (quarter of the plain false positives), {thr_half} (half)")
    del d_va, d_prob
    thresholds = THRESHOLDS
    res["eval"] = []
    for i, sv in enumerate(EVAL_SURVEYS):
        # room for every event of every chunk: a stream cut inside a chunk
        # would count that chunk's pulses for the plain path but not for the
        # network
        cap = 2000 * args.eval_chunks
        d_te, ev_te, st_te, _ = build_stream(make_config(sv), args.seed + 100 + i, cap,
                                             max_chunks=args.eval_chunks, log=log)
        if len(ev_te) >= cap:
            raise RuntimeError("held-out stream filled its event buffer; raise the capacity")
        n = len(ev_te)
        d_prob = abi.DeviceBuffer(4 * max(n, 1))
        t = time.time()
        abi.predict(net, d_te.ptr, n, d_prob.ptr)
        abi.synchronize()
        t_inf = time.time() - t
        prob = d_prob.download(np.float32, (n,))
        r = score(prob, ev_te, st_te, thresholds)
        r["survey"] = sv
        r["operating_point"] = r[f"cnn@{thr}"]
        r["operating_point_half"] = r[f"cnn@{thr_half}"]
        hr, hf = recall_fp_rule(prob, ev_te, st_te, thr_half, snr_keep)
        r["hybrid_half"] = {"recall": hr, "false_positives": hf, "snr_keep": snr_keep}
        wr, wf = recall_fp_wide(prob, ev_te, st_te, thr_wide, wide_snr)
        r["wide_half"] = {"recall": wr, "false_positives": wf, "threshold": thr_wide, "wide_snr_keep": wide_snr}
        r["missed_at_half"] = missed_pulses(prob, ev_te, thr_half)
        r["inference_events_per_s"] = n / max(t_inf, 1e-9)
        res["eval"].append(r)
        log(json.dumps({k: r[k] for k in ("plain", "operating_point", "operating_point_half", "hybrid_half",
                                          "wide_half", "missed_at_half")}))
    if args.batch_beam_chunks > 0:
        # held-out streams over the 32 beam shapes of the batch, aggregated
        # (configs: inference on the batch's streams, the plain path judging)
        agg = {"plain_found": 0, "injected": 0, "plain_fp": 0, "clusters": 0, "events": 0,
               "cnn_found": {}, "cnn_fp": {}, "hyb_found": 0, "hyb_fp": 0, "wide_found": 0, "wide_fp": 0}
        t_all = time.time()
        for b, sv in enumerate(batch_beam_surveys()):
            cap = 2000 * args.batch_beam_chunks
            d_te, ev_te, st_te, _ = build_stream(make_config(sv), args.seed + 1000 + b, cap,
                                                 max_chunks=args.batch_beam_chunks, log=lambda m: None)
            if len(ev_te) >= cap:
                raise RuntimeError("batch-shape stream filled its event buffer; raise the capacity")
            n = len(ev_te)
            d_prob = abi.DeviceBuffer(4 * max(n, 1))
            abi.predict(net, d_te.ptr, n, d_prob.ptr)
            abi.synchronize()
            prob = d_prob.download(np.float32, (n,))
            agg["plain_found"] += st_te["injected_found"]
            agg["injected"] += st_te["injected"]
            agg["plain_fp"] += st_te["false_clusters"]
            agg["clusters"] += st_te["clusters"]
            agg["events"] += n
            hr, hf = recall_fp_rule(prob, ev_te, st_te, thr_half, snr_keep)
            agg["hyb_found"] += round(hr * st_te["injected"])
            agg["hyb_fp"] += hf
            wr, wf = recall_fp_wide(prob, ev_te, st_te, thr_wide, wide_snr)
            agg["wide_found"] += round(wr * st_te["injected"])
            agg["wide_fp"] += wf
            for t in sorted(set((thr, thr_half) + THRESHOLDS)):
                keep = prob > t
                hits = keep & (ev_te["label"] == 1)
                found = len(set(zip(ev_te["chunk"][hits].tolist(), ev_te["pulse"][hits].tolist())))
                agg["cnn_found"][t] = agg["cnn_found"].get(t, 0) + found
                agg["cnn_fp"][t] = agg["cnn_fp"].get(t, 0) + int(((ev_te["label"] == 0) & keep).sum())
            del d_te, d_prob
        c2 = {"beams": 32, "chunks_per_beam": args.batch_beam_chunks, "seconds": time.time() - t_all,
              "plain": {"recall": agg["plain_found"] / max(agg["injected"], 1),
                        "false_positives": agg["plain_fp"], "clusters": agg["clusters"]},
              "injected": agg["injected"]}
        for t in sorted(set((thr, thr_half) + THRESHOLDS)):
            c2[f"cnn@{t}"] = {"recall": agg["cnn_found"][t] / max(agg["injected"], 1),
                              "false_positives": agg["cnn_fp"][t]}
        c2["operating_point"] = c2[f"cnn@{thr}"]
        c2["operating_point_half"] = c2[f"cnn@{thr_half}"]
        c2["hybrid_half"] = {"recall": agg["hyb_found"] / max(agg["injected"], 1),
                             "false_positives": agg["hyb_fp"], "snr_keep": snr_keep}
        c2["wide_half"] = {"recall": agg["wide_found"] / max(agg["injected"], 1),
                           "false_positives": agg["wide_fp"], "threshold": thr_wide, "wide_snr_keep": wide_snr}
        res["eval_batch_shapes"] = c2
        log(json.dumps({"batch_shapes": {k: c2[k] for k in ("plain", "operating_point",
                                                           "operating_point_half", "hybrid_half",
                                                           "wide_half", "injected")}}))
    if args.save:
        np.save(args.save, net.params())
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    net.close()
    return res


if __name__ == "__main__":
    main()
