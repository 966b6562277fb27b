"""Physics performance of track candidates on events with truth.

Restates the reference's analysis definitions:

- truth matching (``matchedSimTrkIdxs``, code/core/trkCore.cc:231-403): the
  hits of a track candidate are made unique by (index, type); a candidate is
  matched to a simulated track when more than 75% of those hits belong to it.
  Hits of the synthetic events belong to at most one track (noise: none), so
  the reference's permutations over a hit's several sim tracks reduce to one;
- a candidate with no matched track is fake, and one matched to a track that
  more than one candidate matches is a duplicate
  (code/core/write_sdl_ntuple.cc:206-308);
- efficiency: simulated tracks of the signal vertex with pT > 0.9, |eta| < 2.4,
  |vz| < 30 cm and transverse vertex < 2.5 cm (efficiency/src/performance.cc:
  33-37,600-680, pt_cut default efficiency/src/helper.cc:27) that at least one
  candidate matches, over all such tracks.

The hit lists are the engine's output lists (``LST::getOutput``,
SDL/LST.cc:366-420): pixel hits first for pT5/pT3 (4, or 3 when the triplet
seed's repeated fourth hit was dropped), pixel hits only for pLS candidates and
outer-tracker hits only for T5 candidates.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

PIXEL, OUTER = 0, 4  # hit types of the reference ntuple (pixel, Phase-2 outer tracker)
TYPE_NAMES = {4: "T5", 5: "pT3", 7: "pT5", 8: "pLS"}


def hit_types(tc_type: int, length: int) -> list[int]:
    """Hit type of every entry of a candidate's output hit list."""
    if tc_type == 8:
        return [PIXEL] * length
    if tc_type == 4:
        return [OUTER] * length
    n_outer = 10 if tc_type == 7 else 6  # pT5: the T5's 10 hits; pT3: the T3's 6
    return [PIXEL] * (length - n_outer) + [OUTER] * n_outer


def match_candidate(hits, types, ph2_trk, pix_trk, threshold: float = 0.75) -> list[int]:
    """Simulated tracks matched by one candidate (trkCore.cc:231-403, one track per hit)."""
    unique = list(dict.fromkeys(zip((int(h) for h in hits), types)))
    n = len(unique)
    if n == 0:
        return []
    trk = [int(pix_trk[h]) if t == PIXEL else int(ph2_trk[h]) for h, t in unique]
    # most common track id, first one in order of appearance on ties (std::max_element)
    order = list(dict.fromkeys(trk))
    counts = [trk.count(t) for t in order]
    best = int(np.argmax(counts))
    if order[best] < 0:
        return []
    return [order[best]] if counts[best] > n * threshold else []


def event_metrics(out: dict, ev: dict, pt_cut: float = 0.9, eta_cut: float = 2.4) -> dict:
    """Per-candidate fake/duplicate flags and per-track match counts of one event.

    out: ``{"tc_hitIdxs", "tc_trackCandidateType"}`` as returned by ``Event.getOutput``.
    ev: the event dict with the ``truth_*`` arrays of ``synth.generate_event``.
    """
    types = np.asarray(out["tc_trackCandidateType"])
    matched = [match_candidate(h, hit_types(int(t), len(h)), ev["truth_hit_trk"], ev["truth_pix_trk"])
               for h, t in zip(out["tc_hitIdxs"], types)]
    n_trk = len(ev["truth_pt"])
    per_trk = np.zeros(n_trk, np.int64)
    per_trk_mask = np.zeros(n_trk, np.int64)
    for m, t in zip(matched, types):
        for i in m:
            per_trk[i] += 1
            per_trk_mask[i] |= 1 << int(t)
    is_fake = np.array([len(m) == 0 for m in matched], bool)
    is_dup = np.array([any(per_trk[i] > 1 for i in m) for m in matched], bool)
    vxy = np.zeros(n_trk)  # synthetic helices start on the beam line
    denom = (ev["truth_is_hard"] & (np.asarray(ev["truth_q"]) != 0) & (ev["truth_pt"] > pt_cut)
             & (np.abs(ev["truth_eta"]) < eta_cut) & (np.abs(ev["truth_vz"]) < 30.0) & (vxy < 2.5))
    return {"n_tc": len(types), "tc_type": types, "tc_is_fake": is_fake, "tc_is_duplicate": is_dup,
            "tc_matched": matched, "sim_tc_matched": per_trk, "sim_tc_matched_mask": per_trk_mask,
            "sim_in_denominator": denom}


def summarize(per_event: list[dict]) -> dict:
    """Efficiency, fake rate and duplicate rate over events, overall and per candidate type."""
    denom = sum(int(m["sim_in_denominator"].sum()) for m in per_event)
    numer = sum(int((m["sim_in_denominator"] & (m["sim_tc_matched"] > 0)).sum()) for m in per_event)
    n_tc = sum(m["n_tc"] for m in per_event)
    fake = sum(int(m["tc_is_fake"].sum()) for m in per_event)
    dup = sum(int(m["tc_is_duplicate"].sum()) for m in per_event)
    res = {"events": len(per_event), "sim_tracks": denom, "efficiency": numer / denom if denom else None,
           "track_candidates": n_tc, "fake_rate": fake / n_tc if n_tc else None,
           "duplicate_rate": dup / n_tc if n_tc else None, "by_type": {}}
    for code, name in TYPE_NAMES.items():
        k = sum(int((m["tc_type"] == code).sum()) for m in per_event)
        f = sum(int((m["tc_is_fake"] & (m["tc_type"] == code)).sum()) for m in per_event)
        d = sum(int((m["tc_is_duplicate"] & (m["tc_type"] == code)).sum()) for m in per_event)
        e = sum(int((m["sim_in_denominator"] & (((m["sim_tc_matched_mask"] >> code) & 1) > 0)).sum())
                for m in per_event)
        res["by_type"][name] = {"track_candidates": k, "fake_rate": f / k if k else None,
                                "duplicate_rate": d / k if k else None,
                                "efficiency": e / denom if denom else None}
    return res


def main(argv=None):
    """Efficiency / fake / duplicate rates of the engine on seeded synthetic events."""
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--kind", default="pu", choices=["pu", "muon"])
    ap.add_argument("--n-pu", type=int, default=200)
    ap.add_argument("--events", type=int, default=4)
    ap.add_argument("--t5-mode", type=int, default=0, help="0: plain chi2 selection, 1: T5 DNN")
    args = ap.parse_args(argv)
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    es = lst.LSTESData(synth.geometry_dir())
    per_event = []
    for i in range(args.events):
        ev = (synth.generate_event("pu", i, n_pu=args.n_pu) if args.kind == "pu"
              else synth.generate_event("muon", i))
        e = lst.Event(es, t5_mode=args.t5_mode).run_lst(ev, "tc")
        per_event.append(event_metrics(e.getOutput(), ev))
        e.close()
    print(json.dumps(summarize(per_event)), flush=True)


if __name__ == "__main__":
    sys.exit(main())
