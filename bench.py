"""Throughput benchmark of the MI355X LST engine.

One step = the LST chain on one synthetic PU200-like event (about 7000
helices with pT = 0.4 + Exp(0.5) GeV plus 200 hard tracks, 1% noise hits,
OT800_IT615_pt0.8 geometry), inputs already resident in HBM (prepared once on
the host and staged; generation and staging are outside the timed region).
Prints ONE JSON line (see the `result` dict) with the roofline of the
dominant kernel and the oracle's CPU throughput on a bounded sample.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The batched engine runs one HIP stream per event context.  HIP gives a
# process 4 hardware queues unless told otherwise and the default stream holds
# one of them, so a fourth context would share a queue (and serialize) with
# another; 8 queues give every context its own.  Read when HIP initializes.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

METRIC = "PU200-like events/sec on 1 MI355X, batched vs per-event, at matched TC set"
STAGE_NAMES = ["hits", "pixel_segments", "mini_doublets", "segments", "triplets", "quintuplets",
               "pixel_segment_cleaning", "pixel_quintuplets", "pixel_triplets", "track_candidates"]
# the kernel each timer slot brackets (lst_event_kernel_times)
KERNEL_NAMES = ["k_hits", "k_pixel_segments", "k_md_eval", "k_ls_eval", "k_t3_select", "k_t5_final",
                "k_pls_pairs<false>", "k_pix_stage_a<PT5Kind>", "k_pix_stage_a<PT3Kind>", "k_t5_tc_pairs",
                "k_pt5_pairs", "k_pt3_pairs", "k_cross_pls_pairs", "k_pls_pairs<true>", "k_t5_dup_after_build",
                "k_t3_enum", "k_t5_dnn"]
ORACLE_STAGES = ["hits", "pls", "md", "ls", "t3", "t5", "plsclean", "pt5", "pt3", "tc"]
LAST_STAGE = 9  # the full chain up to track candidates
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch from rocprofv3 FETCH_SIZE + WRITE_SIZE runs of this
# benchmark, scaled by a 1 GiB copy calibration (profiles/pmc_traffic_pu200_chain.md).
# Re-measured whenever a kernel changes; None for kernels not yet profiled.
MEASURED_TRAFFIC = {
    "k_hits": 9291438,
    "k_pixel_segments": 2275347,
    "k_md_eval": 11193552,
    "k_ls_eval": 59016072,
    "k_t3_select": 29454083,
    "k_t5_final": 30587625,
    "k_pls_pairs<false>": 1040702,
    "k_pix_stage_a<PT5Kind>": 13754338,
    "k_pix_stage_a<PT3Kind>": 9036923,
    "k_t5_tc_pairs": 3106630,
    "k_pt5_pairs": 3444484,
    "k_pt3_pairs": 130123,
    "k_cross_pls_pairs": 1850537,
    "k_pls_pairs<true>": 1066195,
    "k_t5_dup_after_build": 2989959,
    "k_t3_enum": 11423061,
}


def _input_struct(lst, ev):
    import numpy as np
    a = {k: np.ascontiguousarray(v) for k, v in ev.items() if k.startswith(("see_", "ph2_"))}
    p = lambda x: x.ctypes.data_as(C.c_void_p)
    inp = lst._Input(
        *[p(a[k]) for k in ("see_px", "see_py", "see_pz", "see_dxy", "see_dz", "see_ptErr", "see_etaErr",
                            "see_stateTrajGlbX", "see_stateTrajGlbY", "see_stateTrajGlbZ", "see_stateTrajGlbPx",
                            "see_stateTrajGlbPy", "see_stateTrajGlbPz", "see_q", "see_hitIdx", "see_hitIdx_off")],
        len(a["see_px"]), *[p(a[k]) for k in ("ph2_detId", "ph2_x", "ph2_y", "ph2_z")], len(a["ph2_x"]))
    return inp, a


def algorithmic_bytes(counts: dict) -> dict:
    """Minimal HBM bytes per launch of each timed kernel: inputs read once, outputs written once.

    Record sizes are the fields each kernel reads/writes:
      k_hits             : read x,y,z,detId (16 B/hit), write rt,phi,eta,module,4 edges (30 B/hit)
      k_pixel_segments   : read 21 floats and 4 hit indices (100 B/pLS), write 2 MDs + 1 segment + pLS (400 B/pLS)
      k_md_eval          : read x,y,z,rt of the module-pair hits (16 B/hit), write a 4 B result index per
                           candidate pair and 36 B of parked results per mini-doublet
      k_ls_eval          : read 44 B of each mini-doublet, write a 4 B result index per candidate and
                           24 B of parked results per segment
      k_t3_enum          : per inner segment 20 B (outer mini-doublet, its segment range, prefix, outer module),
                           per shared-MD candidate 4 B read (listed segment) and 12 B written
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
    """
    h, md, ls, t3, t5 = counts["hits"], counts["mds"], counts["segs"], counts["t3"], counts["t5"]
    pls, pt5, pt3, tc = counts["pls"], counts["pt5"], counts["pt3"], counts["tc"]
    return {
        "k_hits": 46 * h,
        "k_pixel_segments": 500 * pls,
        "k_md_eval": 16 * h + 4 * counts.get("mdcand", 0) + 36 * md,
        "k_ls_eval": 44 * md + 4 * counts.get("lscand", 0) + 24 * ls,
        "k_t3_enum": 20 * ls + 16 * counts.get("t3surv", 0),
        "k_t3_select": 22 * ls + 20 * md + 60 * t3,
        "k_t5_final": 53 * t3 + 28 * ls + 52 * md + 119 * t5,
        "k_pls_pairs<false>": 31 * pls,
        "k_pls_pairs<true>": 31 * pls,
        "k_pix_stage_a<PT5Kind>": 64 * pls + 128 * t5 + 40 * counts.get("pt5surv", 0),
        "k_pix_stage_a<PT3Kind>": 64 * pls + 64 * t3 + 40 * counts.get("pt3surv", 0),
        "k_t5_tc_pairs": 55 * t5,
        "k_pt5_pairs": 61 * pt5,
        "k_pt3_pairs": 46 * pt3,
        "k_cross_pls_pairs": 50 * pls + 25 * tc,
        "k_t5_dup_after_build": 55 * t5,
    }


def cpu_baseline(events, budget_s: float = 20.0) -> dict:
    """The oracle (serial C++ restatement, 1 thread) on a bounded sample of the same events."""
    from oracle import orc
    from tracklooper_amd import synth
    g = orc.Geometry(synth.geometry_dir())
    n, t0 = 0, time.perf_counter()
    stage = ORACLE_STAGES[LAST_STAGE]
    while time.perf_counter() - t0 < budget_s and n < 4 * len(events):
        orc.Event(g).run_lst(events[n % len(events)], stage)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "events/s", "cores": 1, "kind": "port",
            "sample": f"{n} PU200-like events through the same stages, serial oracle, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--events", type=int, default=4, help="distinct seeded events cycled through")
    ap.add_argument("--n-pu", type=int, default=200)
    ap.add_argument("--contexts", type=int, default=4, help="event contexts of the batched engine")
    ap.add_argument("--capacity", type=int, default=8192, help="track candidates copied back per event")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device-to-device 3 times (reference for FETCH_SIZE/WRITE_SIZE scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (each box has one MI355X); --gpus {args.gpus} is not supported.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
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
    from tracklooper_amd import synth

    t_gen = time.perf_counter()
    events = [synth.generate_event("pu", i, n_pu=args.n_pu) for i in range(args.events)]
    t_gen = time.perf_counter() - t_gen
    es = lst.LSTESData(synth.geometry_dir())
    L = lst.lib()

    # batched engine (measured first, while its contexts' streams are the only ones besides the default
    # stream): contexts run the events round robin, outputs copied to the host
    batch = lst.Batch(es, n_contexts=args.contexts)
    batch.stage(events)
    outs, bufs = lst.Batch._outputs(args.steps, args.capacity)
    for b in bufs:  # host output buffers mapped before the timed region
        for a in b.values():
            a.fill(0)
    # untimed: graph capture in every context and the clocks' ramp after the host-side work above
    batch_warmup = max(args.warmup, len(events), 8 * args.contexts)
    warm_outs, warm_bufs = lst.Batch._outputs(batch_warmup, args.capacity)
    lst._check(L.lst_batch_run_staged(batch.h, 0, batch_warmup, warm_outs))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lst._check(L.lst_batch_run_staged(batch.h, 0, args.steps, outs))
    torch.cuda.synchronize()
    wall_batch = time.perf_counter() - t0
    batch.close()

    # per-event engine: one context, every kernel launched for each event
    ev = lst.Event(es)
    L.lst_event_stage_input.argtypes = [C.c_void_p, C.POINTER(lst._Input), C.c_uint32]
    L.lst_event_run_staged.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
    L.lst_event_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.lst_event_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    L.lst_event_stream.argtypes = [C.c_void_p]
    L.lst_event_stream.restype = C.c_void_p
    for i, e in enumerate(events):
        inp, keep = _input_struct(lst, e)
        lst._check(L.lst_event_stage_input(ev.h, C.byref(inp), i))
    stream = torch.cuda.ExternalStream(L.lst_event_stream(ev.h))
    torch.cuda.synchronize()

    def step(i):
        lst._check(L.lst_event_run_staged(ev.h, i % len(events), LAST_STAGE))

    for i in range(args.warmup):
        step(i)
    ev.synchronize()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for i in range(args.steps):
        step(i)
    stop.record(stream)
    ev.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = start.elapsed_time(stop)
    ms_per_step = wall * 1e3 / args.steps

    # per-kernel device time of the main kernel of every stage, and object counts per event
    lst._check(L.lst_event_set_timing(ev.h, 1))
    ktimes = np.zeros((len(events), len(KERNEL_NAMES)), np.float32)
    counts = []
    for i in range(len(events)):
        step(i)
        lst._check(L.lst_event_kernel_times(ev.h, ktimes[i].ctypes.data_as(C.c_void_p), len(KERNEL_NAMES)))
        counts.append({"hits": int(lst.lib().lst_event_count(ev.h, b"hits")),
                       "mds": int(ev.field("mds.nMDs").sum()),
                       "segs": int(ev.field("segs.nSegments").sum()),
                       "t3": int(ev.field("t3.nTriplets").sum()),
                       "t5": int(ev.field("t5.nQuintuplets").sum()),
                       "pls": int(lst.lib().lst_event_count(ev.h, b"pixelSegments")),
                       "pt5": int(lst.lib().lst_event_count(ev.h, b"pixelQuintuplets")),
                       "pt3": int(lst.lib().lst_event_count(ev.h, b"pixelTriplets")),
                       "tc": int(lst.lib().lst_event_count(ev.h, b"trackCandidates")),
                       "t3cand": int(lst.lib().lst_event_count(ev.h, b"tripletCandidates")),
                       "t3surv": int(lst.lib().lst_event_count(ev.h, b"tripletSharedMDCandidates")),
                       "mdcand": int(lst.lib().lst_event_count(ev.h, b"miniDoubletCandidates")),
                       "lscand": int(lst.lib().lst_event_count(ev.h, b"segmentCandidates")),
                       "pt5surv": int(lst.lib().lst_event_count(ev.h, b"pixelQuintupletSurvivors")),
                       "pt3surv": int(lst.lib().lst_event_count(ev.h, b"pixelTripletSurvivors")),
                       "arena_mib": int(lst.lib().lst_event_count(ev.h, b"arenaMiB"))})
    lst._check(L.lst_event_set_timing(ev.h, 0))
    per_event_out = []
    for i in range(len(events)):
        step(i)
        per_event_out.append(ev.getOutput())

    batched = lst.Batch._collect(outs, bufs)
    tc_match = all(b["n_tc"] == len(per_event_out[i % len(events)]["tc_hitIdxs"]) and
                   b["tc_hitIdxs"] == per_event_out[i % len(events)]["tc_hitIdxs"] and
                   np.array_equal(b["tc_trackCandidateType"], per_event_out[i % len(events)]["tc_trackCandidateType"])
                   for i, b in enumerate(batched))
    if not tc_match:
        raise SystemExit("batched track candidates differ from the per-event ones")
    from tracklooper_amd import metrics
    physics = metrics.summarize([metrics.event_metrics(o, e) for o, e in zip(per_event_out, events)])

    avg_ms = ktimes.mean(0)[: len(KERNEL_NAMES)]
    per_kernel = {KERNEL_NAMES[k]: float(avg_ms[k]) for k in range(len(KERNEL_NAMES))}
    dom = max(per_kernel, key=per_kernel.get)
    mean_counts = {k: float(np.mean([c[k] for c in counts])) for k in counts[0]}
    abytes = algorithmic_bytes(mean_counts)
    dom_bytes = abytes.get(dom, 0.0)
    achieved = dom_bytes / (per_kernel[dom] * 1e-3) / 1e9 if per_kernel[dom] > 0 else 0.0

    result = {
        "metric": METRIC,
        "value": args.steps / wall_batch,
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
            "workload": f"batched engine, {args.contexts} event contexts round robin, one PU{args.n_pu}-like "
                        f"synthetic event per step ({args.events} seeded events cycled, inputs resident in HBM, "
                        f"output lists copied to the host), stages hits..{STAGE_NAMES[LAST_STAGE]}",
            "batched_events_per_s": args.steps / wall_batch,
            "per_event_events_per_s": args.steps / wall,
            "per_event_ms_per_step": ms_per_step,
            "batched_over_per_event": (args.steps / wall_batch) / (args.steps / wall),
            "tc_sets_match": tc_match,
            "physics": {k: physics[k] for k in ("sim_tracks", "efficiency", "fake_rate", "duplicate_rate")},
            "contexts": args.contexts,
            "batch_warmup_events": batch_warmup,
            "gpu_max_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
            "device_mib_per_context": counts[0]["arena_mib"],
            "geometry": "OT800_IT615_pt0.8",
            "mean_hits": mean_counts["hits"], "mean_mini_doublets": mean_counts["mds"],
            "mean_segments": mean_counts["segs"], "mean_triplets": mean_counts["t3"],
            "mean_quintuplets": mean_counts["t5"], "mean_pixel_quintuplets": mean_counts["pt5"],
            "mean_pixel_triplets": mean_counts["pt3"], "mean_track_candidates": mean_counts["tc"],
            "per_event_gpu_ms_per_step": gpu_ms / args.steps,
            "kernel_ms": per_kernel,
        },
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": MEASURED_TRAFFIC.get(dom),
                     "algorithmic_bytes": dom_bytes},
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(events, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
