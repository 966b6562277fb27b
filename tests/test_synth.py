"""Synthetic-event generator invariants (CPU only)."""
import numpy as np

from tracklooper_amd import synth


def test_hits_grouped_and_bounded():
    ev = synth.generate_event("pu", 0, n_pu=30)
    det = ev["ph2_detId"]
    # hits of one module are contiguous (SDL/Hit.h:261-267 takes min/max indices)
    change = np.nonzero(np.diff(det.astype(np.int64)))[0]
    assert len(np.unique(det)) == len(change) + 1
    _, counts = np.unique(det, return_counts=True)
    assert counts.max() <= 127  # int8 hit counts (SDL/Hit.h:31-32)


def test_deterministic_seeding():
    a = synth.generate_event("muon", 5)
    b = synth.generate_event("muon", 5)
    c = synth.generate_event("muon", 6)
    assert all(np.array_equal(a[k], b[k]) for k in a)
    assert not np.array_equal(a["ph2_x"], c["ph2_x"])


def test_hits_on_module_planes_and_pairs():
    geo = synth._geometry_cache("OT800_IT615_pt0.8")
    ev = synth.generate_event("muon", 2)
    row_of = {int(d): i for i, d in enumerate(geo.detid)}
    lower_of = {int(u): k for k, u in enumerate(geo.up)}
    lower_of.update({int(l): k for k, l in enumerate(geo.lo)})
    trk = ev["truth_hit_trk"]
    for d, x, y, z in zip(ev["ph2_detId"], ev["ph2_x"], ev["ph2_y"], ev["ph2_z"]):
        r = row_of[int(d)]
        k = lower_of[r]
        dist = np.dot(np.array([x, y, z], np.float64) - geo.cent[r], geo.n[k])
        assert abs(dist) < 1e-3  # on the sensor plane
    # muons leave hits in lower/upper pairs
    assert len(ev["ph2_x"]) % 2 == 0 and np.all(trk >= 0)


def test_seed_consistency():
    ev = synth.generate_event("muon", 1)
    n = len(ev["see_px"])
    assert n == 10 or n == len(np.unique(ev["truth_seed_trk"]))
    assert ev["see_hitIdx_off"][-1] == len(ev["see_hitIdx"])
    pt = np.hypot(ev["see_px"], ev["see_py"])
    tpt = ev["truth_pt"][ev["truth_seed_trk"]]
    assert np.allclose(pt, tpt, rtol=0.06)
