"""Throughput benchmark of the MI355X LST engine on uneven batches.

One step = one launch set of the batched engine over the 64 events of an
uneven batch (pileup drawn U{50..250} per event, 35 charged helices per
interaction with pT = 0.4 + Exp(0.5) GeV, 200 hard tracks, 1% noise hits,
OT800_IT615_pt0.8 geometry).  The events are prepared on the host once
(pinned; the reference preloads its inputs the same way, bin/sdl.cc:352-384);
inside the timed region each step copies its events' inputs to the device
(as the reference's timed loop does, code/core/trkCore.cc:877-879), runs the
whole chain, and copies the track-candidate lists back to the host.

The same events also run through the plain per-event path (one event handle,
one event at a time, the same host-to-device copy and output copy per event),
and the two track-candidate sets are compared event by event.  Prints ONE
JSON line with the roofline of the dominant kernel of the batched launch set
and the oracle's CPU throughput on all host cores and on one.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PU200-like events/sec on 1 MI355X, batched vs per-event, at matched TC set"
# the kernel each per-event timer slot brackets (lst_event_kernel_times)
KERNEL_NAMES = ["k_hits", "k_pixel_segments", "k_md_eval", "k_ls_eval", "k_t3_select", "k_t5_final",
                "k_pls_pairs<false>", "k_pix_stage_a<PT5Kind>", "k_pix_stage_a<PT3Kind>", "k_t5_tc_pairs",
                "k_pt5_pairs", "k_pt3_pairs", "k_cross_pls_pairs", "k_pls_pairs<true>", "k_t5_dup_after_build",
                "k_t3_enum", "k_t5_dnn"]
HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_c3_batch.md")


def event_counts(lst, ev) -> dict:
    L = lst.lib()
    c = lambda k: int(L.lst_event_count(ev.h, k.encode()))
    return {"hits": c("hits"), "mds": int(ev.field("mds.nMDs").sum()), "segs": int(ev.field("segs.nSegments").sum()),
            "t3": int(ev.field("t3.nTriplets").sum()), "t5": int(ev.field("t5.nQuintuplets").sum()),
            "pls": c("pixelSegments"), "pt5": c("pixelQuintuplets"), "pt3": c("pixelTriplets"),
            "tc": c("trackCandidates"), "t3surv": c("tripletSharedMDCandidates"), "mdcand": c("miniDoubletCandidates"),
            "lscand": c("segmentCandidates"), "pt5surv": c("pixelQuintupletSurvivors"),
            "pt3surv": c("pixelTripletSurvivors"), "arena_mib": c("arenaMiB")}


def algorithmic_bytes(n: dict) -> dict:
    """Minimal HBM bytes of each kernel for one event: inputs read once, outputs written once.

    Record sizes are the fields each kernel reads/writes:
      k_hits             : read x,y,z,detId (16 B/hit), write rt,phi,eta,module,4 edges (30 B/hit)
      k_pixel_segments   : read 21 floats and 4 hit indices (100 B/pLS), write 2 MDs + 1 segment + pLS (400 B/pLS)
      k_md_eval          : read x,y,z,rt of the module-pair hits (16 B/hit), write one pass bit per
                           candidate pair and 36 B of parked results per mini-doublet
      k_ls_eval          : read 52 B of each mini-doublet (anchor point and phi, dphi, dz, dphichange, edge
                           phis, module, slope pair), write one pass bit per candidate and 24 B of parked
                           results per segment
      k_t3_enum          : per inner segment 20 B (outer mini-doublet, its segment range, prefix, outer module),
                           per shared-MD candidate 16 B
      k_t3_select        : read 22 B/segment and 20 B/mini-doublet, write 60 B/triplet
      k_t5_final         : read 53 B/triplet, 28 B/segment, 52 B/mini-doublet, write 119 B/quintuplet
      k_pls_pairs        : read eta, phi, score, isQuad, flag and 4 hit indices (30 B/pLS), write 1 B/pLS
      k_pix_stage_a<PT5Kind>: read 64 B/pLS and the 128 B record of each quintuplet, write 40 B per survivor
      k_pix_stage_a<PT3Kind>: read 64 B/pLS and the 64 B record of each triplet, write 40 B per survivor
      k_pt5_pairs        : read score and 14 hit indices (60 B/pT5), write 1 B/pT5
      k_pt3_pairs        : read score, layer and 10 hit indices (45 B/pT3), write 1 B/pT3
      k_t5_tc_pairs,
      k_t5_dup_after_build: read eta, phi, score, flags and 10 hit indices (54 B/quintuplet), write 1 B
      k_cross_pls_pairs  : read 50 B/pLS (direction, flags, hit indices through the MDs), 25 B/track candidate
      k_out_copy         : read and write the 68 B output record of each track candidate
      k_md_cand_fill     : write the 16 B record of each candidate hit pair
      k_md_place         : per mini-doublet read its candidate record (16 B) and parked results (36 B) and the
                           derived fields of its two hits (84 B), write the 134 B mini-doublet
      k_ls_cand_fill     : write the 8 B record of each candidate mini-doublet pair, read 8 B per mini-doublet
      k_ls_place         : per segment read its candidate record (8 B), parked results (24 B), the two
                           mini-doublets' anchor hits and modules (12 B), write the 44 B segment
      k_t3_place         : per triplet read its list entry (8 B) and results (24 B), write the 60 B triplet
      k_t5_place         : per quintuplet read its list entry (12 B) and results (48 B), write the 119 B quintuplet
      k_pix_write<PT5Kind>: per stage-A survivor its pLS, mask position and object (12 B) and its mask word
                           (8 B); per pixel quintuplet its parked result (40 B) and the 150 B record written
    """
    h, md, ls, t3, t5 = n["hits"], n["mds"], n["segs"], n["t3"], n["t5"]
    pls, pt5, pt3, tc = n["pls"], n["pt5"], n["pt3"], n["tc"]
    return {
        "k_hits": 46 * h,
        "k_pixel_segments": 500 * pls,
        "k_md_eval": 16 * h + n["mdcand"] / 8 + 36 * md,
        "k_ls_eval": 52 * md + n["lscand"] / 8 + 24 * ls,
        "k_t3_enum": 20 * ls + 16 * n["t3surv"],
        "k_t3_select": 22 * ls + 20 * md + 60 * t3,
        "k_t5_final": 53 * t3 + 28 * ls + 52 * md + 119 * t5,
        "k_pls_pairs": 31 * pls,
        "k_pix_stage_a<PT5Kind>": 64 * pls + 128 * t5 + 40 * n["pt5surv"],
        "k_pix_stage_a<PT3Kind>": 64 * pls + 64 * t3 + 40 * n["pt3surv"],
        "k_t5_tc_pairs": 55 * t5,
        "k_pt5_pairs": 61 * pt5,
        "k_pt3_pairs (y shares)": 46 * pt3,
        "k_cross_pls_pairs": 50 * pls + 25 * tc,
        "k_t5_dup_after_build": 55 * t5,
        "k_out_copy": 136 * tc,
        "k_md_cand_fill": 16 * n["mdcand"],
        "k_md_place": 270 * md,
        "k_ls_cand_fill": 8 * n["lscand"] + 8 * md,
        "k_ls_place": 88 * ls,
        "k_t3_place": 92 * t3,
        "k_t5_place": 179 * t5,
        "k_pix_write<PT5Kind>": 20 * n["pt5surv"] + 190 * pt5,
    }


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this
    workload (calibrated; see the file), or None when the kernel is not in it."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    for line in text.splitlines():
        m = re.match(r"\|\s*`?([^|`]+?)`?\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|\s*([0-9.]+)\s*\|", line)
        if m and m.group(1).strip() == kernel:
            return float(m.group(4)) * 1e6  # total MB per dispatch
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(events, budget_s: float, threads: int) -> dict:
    """The oracle (serial C++ restatement, one event per thread as the reference runs one event per OpenMP
    thread, bin/sdl.cc:405-419) over the same events for about budget_s seconds."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import orc
    from tracklooper_amd import synth
    g = orc.Geometry(synth.geometry_dir())
    t0 = time.perf_counter()
    done = [0] * threads

    def worker(w):
        i = w
        while time.perf_counter() - t0 < budget_s:
            orc.Event(g).run_lst(events[i % len(events)], "tc")  # ctypes releases the GIL during the call
            done[w] += 1
            i += threads

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(worker, range(threads)))
    dt = time.perf_counter() - t0
    return {"value": sum(done) / dt, "events": sum(done), "seconds": dt, "threads": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed launch sets (64 events each)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-events", type=int, default=64, help="events per launch set")
    ap.add_argument("--batches", type=int, default=2, help="distinct uneven batches cycled through")
    ap.add_argument("--per-event-steps", type=int, default=2, help="batches the per-event path is timed on")
    ap.add_argument("--capacity", type=int, default=8192, help="track candidates copied back per event")
    ap.add_argument("--sets-in-flight", type=int, default=2, choices=(1, 2),
                    help="launch sets the batch keeps in flight (2: each with its own arenas and stream)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds per CPU-baseline configuration")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host cores available (at most 16)")
    ap.add_argument("--workers", type=int, default=8, help="processes generating the events")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (reference for FETCH_SIZE/WRITE_SIZE scaling)")
    ap.add_argument("--profile-only", action="store_true",
                    help="stage, tune and run --steps launch sets, nothing else (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (each box has one MI355X); --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tracklooper_amd import synth

    if not args.copy_calibration:  # events are generated before the process touches the GPU (fork pool)
        t_gen = time.perf_counter()
        E = args.batch_events
        events = []
        for b in range(args.batches):
            events += synth.c3_batch(b, E, workers=args.workers)
        t_gen = time.perf_counter() - t_gen
        print(f"generated {len(events)} events in {t_gen:.1f} s", file=sys.stderr, flush=True)

    import torch
    if args.copy_calibration:
        x = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        x.fill_(1.0)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print("copied 1 GiB device-to-device 3 times", flush=True)
        return
    import tracklooper_amd as lst
    L = lst.lib()
    es = lst.LSTESData(synth.geometry_dir())
    n_ev = len(events)

    # ---- batched engine: launch sets of E events
    batch = lst.Batch(es, slots=E)
    batch.stage(events)
    tune = {"default_ms": None, "tuned_ms": None}
    if not args.no_autotune:
        d_ms, t_ms, _ = batch.autotune(0, repeats=2)
        tune = {"default_ms": d_ms, "tuned_ms": t_ms}
        print(f"autotune: summed kernel time {d_ms:.2f} ms at the default work-group counts, {t_ms:.2f} ms tuned",
              file=sys.stderr, flush=True)
    batch.set_sets_in_flight(args.sets_in_flight)  # the tuner timed one set alone
    # output descriptors of a run of K launch sets: event i of the run writes into buffer i % E (the sets are
    # consumed one after the other, so the buffers hold the last set's lists at the end)
    outs1, bufs = lst.Batch._outputs(E, args.capacity)

    def outputs_for(n_sets):
        outs = (lst._Output * (n_sets * E))()
        for i in range(n_sets * E):
            outs[i] = outs1[i % E]
        return outs

    warm_outs, timed_outs = outputs_for(args.warmup), outputs_for(args.steps)
    n_staged = args.batches * E
    # one call runs all the steps: launch set k + 1's inputs cross to the device while set k computes
    if args.warmup:
        lst._check(L.lst_batch_run_staged(batch.h, 0, args.warmup * E, warm_outs))
    torch.cuda.synchronize()
    if args.profile_only:
        lst._check(L.lst_batch_run_staged(batch.h, 0, args.steps * E, timed_outs))
        torch.cuda.synchronize()
        print(json.dumps({"profile_only": True, "steps": args.steps}), flush=True)
        return
    t0 = time.perf_counter()
    lst._check(L.lst_batch_run_staged(batch.h, (args.warmup * E) % n_staged, args.steps * E, timed_outs))
    torch.cuda.synchronize()
    wall_batch = time.perf_counter() - t0
    # outputs of every event for the comparison with the per-event path
    batched = []
    for b in range(args.batches):
        batched += batch.run_staged(b * E, E, args.capacity)
    # per-kernel device time of one launch set (no graph), for the roofline
    kms = np.zeros(L.lst_tune_count(), np.float32)
    L.lst_batch_kernel_times.argtypes = [lst.C.c_void_p, lst.C.c_uint32, lst.C.c_void_p, lst.C.c_uint32]
    lst._check(L.lst_batch_kernel_times(batch.h, 0, kms.ctypes.data_as(lst.C.c_void_p), len(kms)))
    batch_kernel_ms = {L.lst_tune_name(k).decode(): float(kms[k]) for k in range(len(kms)) if kms[k] > 0}
    tuning = batch.tuning()
    batch.close()

    # ---- plain per-event path: one handle, one event at a time
    import ctypes as C
    ev = lst.Event(es)
    L.lst_event_stage_input.argtypes = [C.c_void_p, C.POINTER(lst._Input), C.c_uint32]
    L.lst_event_run_staged.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
    L.lst_event_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.lst_event_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    for i, e in enumerate(events):
        inp, keep = lst._make_input(e)
        lst._check(L.lst_event_stage_input(ev.h, C.byref(inp), i))
    n_tc = C.c_uint32(0)
    cap = args.capacity
    hits = np.zeros((cap, 14), np.uint32)
    ln = np.zeros(cap, np.uint32)
    sd = np.zeros(cap, np.int32)
    ty = np.zeros(cap, np.int16)

    def one_event(i):  # inputs to the device, the chain, the output lists to the host
        lst._check(L.lst_event_run_staged(ev.h, i, 9))
        lst._check(L.lst_event_get_output(ev.h, C.byref(n_tc), lst._p(hits), lst._p(ln), lst._p(sd), lst._p(ty),
                                          cap))

    for i in range(min(n_ev, 8)):
        one_event(i)
    torch.cuda.synchronize()
    n_per = args.per_event_steps * E
    t0 = time.perf_counter()
    for i in range(n_per):
        one_event(i % n_ev)
    torch.cuda.synchronize()
    wall_per = time.perf_counter() - t0

    # outputs and object counts of every event; per-kernel times of the per-event path
    per_event_out, counts = [], []
    lst._check(L.lst_event_set_timing(ev.h, 1))
    ktimes = np.zeros((n_ev, len(KERNEL_NAMES)), np.float32)
    for i in range(n_ev):
        lst._check(L.lst_event_run_staged(ev.h, i, 9))
        lst._check(L.lst_event_kernel_times(ev.h, ktimes[i].ctypes.data_as(C.c_void_p), len(KERNEL_NAMES)))
        per_event_out.append(ev.getOutput())
        counts.append(event_counts(lst, ev))
    lst._check(L.lst_event_set_timing(ev.h, 0))

    tc_match = all(b["n_tc"] == len(p["tc_hitIdxs"]) and b["tc_hitIdxs"] == p["tc_hitIdxs"] and
                   np.array_equal(b["tc_trackCandidateType"], p["tc_trackCandidateType"]) and
                   np.array_equal(b["tc_seedIdx"], p["tc_seedIdx"])
                   for b, p in zip(batched, per_event_out))
    if not tc_match:
        raise SystemExit("batched track candidates differ from the per-event ones")
    from tracklooper_amd import metrics
    physics = metrics.summarize([metrics.event_metrics(o, e) for o, e in zip(per_event_out, events)])

    # roofline of the dominant kernel of the batched launch set (first batch)
    first = counts[:E]
    abytes = {}
    for c in first:
        for k, b in algorithmic_bytes(c).items():
            abytes[k] = abytes.get(k, 0) + b
    # per launch set, averaged over the cycled batches (what a profile of the timed steps sees)
    mean_bytes = {}
    for c in counts:
        for k, b in algorithmic_bytes(c).items():
            mean_bytes[k] = mean_bytes.get(k, 0) + b * E / len(counts)
    dom = max(batch_kernel_ms, key=batch_kernel_ms.get)
    dom_bytes = abytes.get(dom)
    achieved = dom_bytes / (batch_kernel_ms[dom] * 1e-3) / 1e9 if dom_bytes else None
    traffic = measured_traffic(dom)

    value = args.steps * E / wall_batch
    per_event_rate = n_per / wall_per
    pus = np.concatenate([synth.c3_pileups(b, E) for b in range(args.batches)])
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_batch * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"uneven batches of {E} synthetic events, pileup U{{50..250}} per event "
                        f"(mean {pus.mean():.0f}), OT800_IT615_pt0.8, plain chi2 T5 selection; one step = one "
                        f"launch set of {E} events: inputs host-to-device, hits..track candidates, output lists "
                        f"device-to-host; {args.batches} batches cycled; the {args.steps} steps run in one call, "
                        f"each set's input copies overlapping the previous set's kernels"
                        + (", two launch sets in flight (each on its own arenas and stream)"
                           if args.sets_in_flight == 2 else ""),
            "batched_events_per_s": value,
            "per_event_events_per_s": per_event_rate,
            "per_event_ms_per_event": wall_per * 1e3 / n_per,
            "batched_over_per_event": value / per_event_rate,
            "tc_sets_match": tc_match,
            "events_compared": n_ev,
            "physics": {k: physics[k] for k in ("sim_tracks", "efficiency", "fake_rate", "duplicate_rate")},
            "slots": E,
            "sets_in_flight": args.sets_in_flight,
            "autotune_summed_kernel_ms": tune,
            "tuned_grid": {k: v for k, v in tuning.items() if v},
            "device_mib_per_slot": counts[0]["arena_mib"],
            "geometry": "OT800_IT615_pt0.8",
            "mean_hits": float(np.mean([c["hits"] for c in counts])),
            "min_hits": int(np.min([c["hits"] for c in counts])),
            "max_hits": int(np.max([c["hits"] for c in counts])),
            "mean_track_candidates": float(np.mean([c["tc"] for c in counts])),
            "batch_kernel_ms": batch_kernel_ms,
            "batch_algorithmic_mb": {k: b / 1e6 for k, b in mean_bytes.items()},
            "batch_kernel_ms_sum": float(sum(batch_kernel_ms.values())),
            "per_event_kernel_ms": {KERNEL_NAMES[k]: float(ktimes[:, k].mean()) for k in range(len(KERNEL_NAMES))},
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "algorithmic_bytes": dom_bytes, "kernel_ms": batch_kernel_ms[dom],
                     "events_per_launch": E},
    }
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(len(os.sched_getaffinity(0)), 16)
        one = cpu_baseline(events[:E], args.cpu_budget, 1)
        many = cpu_baseline(events[:E], args.cpu_budget, threads) if threads > 1 else one
        result["cpu_baseline"] = {
            "value": many["value"], "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"{many['events']} events of the first uneven batch in {many['seconds']:.1f} s, one event per "
                      f"thread, {threads} threads ({cpu_model()}); 1 thread: {one['value']:.3f} events/s "
                      f"({one['events']} events in {one['seconds']:.1f} s)",
            "one_thread_value": one["value"], "cpu_model": cpu_model()}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
