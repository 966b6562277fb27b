"""Truth-matched combinations of an event, stage by stage.

For every simulated track above a pT threshold, the combinations its own hits
allow at each stage are enumerated from the truth table, and those the chain
built are counted:

* mini-doublets: a lower-sensor hit and an upper-sensor hit of the track on
  one module (the builder tries every such pair, SDL/MiniDoublet.h:896-992);
* segments: two of the track's built mini-doublets on modules the module map
  connects (SDL/Segment.h:798-910 loops over exactly those pairs);
* triplets: two of its built segments sharing the middle mini-doublet
  (SDL/Triplet.h:874-1005), except the layer combinations the reference
  rejects outright (sdl layers 12-13-14, 13-14-15, 14-15-16,
  SDL/Triplet.h:241-298);
* quintuplets: two of its built triplets sharing a mini-doublet
  (SDL/Quintuplet.h:2988-3115).

`run` is anything with ``field(name)`` returning the stage arrays under the
names of the engine's field table (the oracle's Event or the engine's Event).
"""
from __future__ import annotations

import numpy as np

EXCLUDED_T3_LAYERS = {(12, 13, 14), (13, 14, 15), (14, 15, 16)}


class Connectivity:
    """Module map and partner modules of a geometry (oracle.orc.Geometry)."""

    def __init__(self, geom):
        self.partner = geom.field("partnerModuleIndices").astype(np.int64)
        self.is_lower = geom.field("isLower").astype(bool)
        self.map = geom.field("moduleMap").astype(np.int64)
        self.n_conn = geom.field("nConnectedModules").astype(np.int64)
        self.stride = int(geom.scalar("stride"))
        self.n_lower = int(geom.scalar("nLowerModules"))
        self.sdl_layers = geom.field("sdlLayers").astype(np.int64)

    def connected(self, m: int) -> set:
        return set(self.map[m * self.stride: m * self.stride + self.n_conn[m]].tolist())


def _objects(ranges, counts, n_lower):
    for m in range(n_lower):
        if ranges[m] < 0:
            continue
        yield from range(int(ranges[m]), int(ranges[m]) + int(counts[m]))


def truth_combinations(run, ev: dict, conn: Connectivity, pt_min: float = 1.0) -> dict:
    """{"md"|"ls"|"t3"|"t5": [built, allowed]} and "t5_true": [(quintuplet slot, track)]."""
    nL = conn.n_lower
    n_hits = len(ev["ph2_x"])
    trk_of_input, pt = ev["truth_hit_trk"], ev["truth_pt"]
    htrk = trk_of_input[run.field("hits.idxs")[:n_hits]]
    hmod = run.field("hits.moduleIndices")[:n_hits].astype(np.int64)
    good = (htrk >= 0) & (pt[np.maximum(htrk, 0)] > pt_min)
    out = {k: [0, 0] for k in ("md", "ls", "t3", "t5")}

    A, O = run.field("mds.anchorHitIndices"), run.field("mds.outerHitIndices")
    md_mod = run.field("mds.moduleIndices").astype(np.int64)
    md_of = {frozenset((int(A[k]), int(O[k]))): k
             for k in _objects(run.field("ranges.miniDoubletModuleIndices"), run.field("mds.nMDs"), nL)}
    hits_of = {}
    for h in np.nonzero(good)[0]:
        hits_of.setdefault((int(htrk[h]), int(hmod[h])), []).append(int(h))
    true_md = {}
    for (t, m), lower in hits_of.items():
        if m >= nL or not conn.is_lower[m]:
            continue
        for a in lower:
            for b in hits_of.get((t, int(conn.partner[m])), []):
                out["md"][1] += 1
                k = md_of.get(frozenset((a, b)))
                if k is not None:
                    out["md"][0] += 1
                    true_md.setdefault(t, []).append(k)

    smd = run.field("segs.mdIndices")
    seg_of = {(int(smd[2 * k]), int(smd[2 * k + 1])): k
              for k in _objects(run.field("ranges.segmentModuleIndices"), run.field("segs.nSegments"), nL)}
    true_seg = {}
    for t, mds in true_md.items():
        for k1 in mds:
            targets = conn.connected(int(md_mod[k1]))
            for k2 in mds:
                if int(md_mod[k2]) in targets:
                    out["ls"][1] += 1
                    s = seg_of.get((k1, k2))
                    if s is not None:
                        out["ls"][0] += 1
                        true_seg.setdefault(t, []).append(s)

    tsi = run.field("t3.segmentIndices")
    t3_of = {(int(tsi[2 * k]), int(tsi[2 * k + 1])): k
             for k in _objects(run.field("ranges.tripletModuleIndices"), run.field("t3.nTriplets"), nL)}
    true_t3 = {}
    for t, segs in true_seg.items():
        for s1 in segs:
            for s2 in segs:
                if smd[2 * s1 + 1] != smd[2 * s2]:
                    continue
                layers = tuple(int(conn.sdl_layers[md_mod[k]]) for k in (smd[2 * s1], smd[2 * s1 + 1], smd[2 * s2 + 1]))
                if layers in EXCLUDED_T3_LAYERS:
                    continue
                out["t3"][1] += 1
                k = t3_of.get((s1, s2))
                if k is not None:
                    out["t3"][0] += 1
                    true_t3.setdefault(t, []).append(k)

    qti = run.field("t5.tripletIndices")
    t5_of = {(int(qti[2 * k]), int(qti[2 * k + 1])): k
             for k in _objects(run.field("ranges.quintupletModuleIndices"), run.field("t5.nQuintuplets"), nL)}
    out["t5_true"] = []
    for t, t3s in true_t3.items():
        for a in t3s:
            for b in t3s:
                # the outer mini-doublet of the inner triplet is the inner one of the outer triplet
                if smd[2 * tsi[2 * a + 1] + 1] != smd[2 * tsi[2 * b]]:
                    continue
                out["t5"][1] += 1
                k = t5_of.get((a, b))
                if k is not None:
                    out["t5"][0] += 1
                    out["t5_true"].append((k, t))
    return out


def add_counts(total: dict, one: dict) -> dict:
    for k in ("md", "ls", "t3", "t5"):
        total.setdefault(k, [0, 0])
        total[k][0] += one[k][0]
        total[k][1] += one[k][1]
    return total
