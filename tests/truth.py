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
  (SDL/Quintuplet.h:2988-3115);
* pixel quintuplets and triplets (`pixel_truth`): a pLS of the track's own
  seed with one of the track's built quintuplets (triplets) on a module the
  pLS's superbin connects, under the builders' own preconditions (pLS not a
  duplicate, object not a duplicate / not already part of a pT5, PS module,
  SDL/PixelTriplet.h:2679-2788 and :1026-1164).

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
    out["t3_true"] = [(k, t) for t, ks in true_t3.items() for k in ks]

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


class PixelMap:
    """The superbin -> connected lower modules lists of a geometry, per pixel type (0 all, 1 pos, 2 neg),
    laid out as the builders read them (SDL/Event.cc:813-816)."""

    SIZE_SUPERBINS = 45000

    def __init__(self, geom):
        self.modules = geom.field("connectedPixels").astype(np.int64)
        idx = [geom.field(n).astype(np.int64) for n in
               ("connectedPixelsIndex", "connectedPixelsIndexPos", "connectedPixelsIndexNeg")]
        size = [geom.field(n).astype(np.int64) for n in
                ("connectedPixelsSizes", "connectedPixelsSizesPos", "connectedPixelsSizesNeg")]
        last = self.SIZE_SUPERBINS - 1
        off_pos = idx[0][last] + size[0][last]
        off_neg = idx[1][last] + size[1][last] + off_pos
        self.index = [idx[0], idx[1] + off_pos, idx[2] + off_neg]
        self.size = size
        self.module_type = geom.field("moduleType").astype(np.int64)
        self.layers = geom.field("layers").astype(np.int64)

    def connected(self, superbin: int, pixel_type: int) -> np.ndarray:
        if superbin < 0 or superbin >= self.SIZE_SUPERBINS or not 0 <= pixel_type <= 2:
            return np.zeros(0, np.int64)
        b = self.index[pixel_type][superbin]
        return self.modules[b:b + self.size[pixel_type][superbin]]


PS_MODULE = 0  # moduleType of a PS module (SDL/Module.h)


def pixel_truth(run, ev: dict, conn: Connectivity, pmap: PixelMap, comb: dict, pt_min: float = 1.0) -> dict:
    """{"pt5"|"pt3": [built, allowed]}: run after the pixel triplets (its pLS duplicate flags are those of the
    first cleaning pass); comb = truth_combinations of the same run."""
    nL = conn.n_lower
    n_pls = int(run.scalar("nPLS"))
    pls_base = int(run.field("ranges.segmentModuleIndices")[nL])
    seed_of = run.field("segs.seedIdx")[:n_pls].astype(np.int64)
    trk_of_seed = ev["truth_seed_trk"]
    pt = ev["truth_pt"]
    is_dup = run.field("segs.isDup")[:n_pls]
    part_pt5 = run.field("segs.partOfPT5")[:n_pls]
    superbin = run.field("segs.superbin")[:n_pls].astype(np.int64)
    ptype = run.field("segs.pixelType")[:n_pls].astype(np.int64)
    t5_mod = run.field("t5.lowerModuleIndices").astype(np.int64)
    t5_dup = run.field("t5.isDup")
    t3_mod = run.field("t3.lowerModuleIndices").astype(np.int64)
    t3_part5 = run.field("t3.partOfPT5")
    n5, n3 = int(run.scalar("nPixelQuintuplets")), int(run.scalar("nPixelTriplets"))
    pt5 = set(zip(run.field("pt5.pixelIndices")[:n5].tolist(), run.field("pt5.T5Indices")[:n5].tolist()))
    pt3 = set(zip(run.field("pt3.pixelSegmentIndices")[:n3].tolist(), run.field("pt3.tripletIndices")[:n3].tolist()))
    t5_by_trk, t3_by_trk = {}, {}
    for k, t in comb["t5_true"]:
        t5_by_trk.setdefault(t, []).append(k)
    for k, t in comb["t3_true"]:
        t3_by_trk.setdefault(t, []).append(k)
    out = {"pt5": [0, 0], "pt3": [0, 0]}
    for i in range(n_pls):
        t = int(trk_of_seed[seed_of[i]])
        if t < 0 or pt[t] <= pt_min or is_dup[i]:
            continue
        mods = set(int(m) for m in pmap.connected(int(superbin[i]), int(ptype[i]))
                   if m < nL and pmap.module_type[m] == PS_MODULE)
        for k in t5_by_trk.get(t, []):  # SDL/PixelTriplet.h:2679-2788
            if int(t5_mod[5 * k]) in mods and not t5_dup[k]:
                out["pt5"][1] += 1
                out["pt5"][0] += (pls_base + i, k) in pt5
        if part_pt5[i]:
            continue
        for k in t3_by_trk.get(t, []):  # SDL/PixelTriplet.h:1026-1164
            m = int(t3_mod[3 * k])
            if (m in mods and pmap.layers[m] in (1, 2) and pmap.module_type[int(t3_mod[3 * k + 1])] == PS_MODULE
                    and not t3_part5[k]):
                out["pt3"][1] += 1
                out["pt3"][0] += (pls_base + i, k) in pt3
    return out
