"""Truth invariants of the oracle's cut logic on noise-free muon events (CPU).

The reference's cuts are built so that the combinations a real track's own
hits form pass them: on perfect helices (hits at the exact crossing points)
at least 99 % of the truth-allowed mini-doublets, segments, triplets and
quintuplets of tracks above 1 GeV must be built, stage by stage (observed:
all but one triplet of 1327).  At the synthetic detector's resolution (20 um
r-phi smearing, strips reported at their centres) mini-doublets and segments
still pass at 99 % or more; triplets and quintuplets lose a few percent to
the r-z windows (observed 95.9 % and 98.3 %).  A cut with a wrong sign, a
swapped operand or a mistyped threshold loses whole classes of true
combinations and fails these bounds.

Pixel quintuplets (pLS of the track's seed with one of its true quintuplets on
a module the pLS's superbin connects): 98.5 % built on exact helices and 98.3 %
at resolution (asserted >= 97 %; the misses are forward, |eta| 1.5-2.3, where
the synthetic pixel disks are a rough model).  Pixel triplets are built only
for what the pT5 builder left (pLS not in a pT5, triplets not in one), a few
dozen combinations biased to those the pT5 cuts rejected: 8 of 12 and 31 of 45
(asserted >= 50 %, which a broken cut would miss by far).

On exact helices the fitted quantities of the true quintuplets have closed
forms: the three-hit circle radii (inner, bridge, outer) and the regression
radius equal the helix radius pT / (0.3 B), and the r-phi and r-z chi-squared
vanish.  They are asserted for 1 < pT < 5 GeV, where the float32 circle fits
are well conditioned (above that the sagitta over a 30 cm lever arm falls to
micrometres and the fits lose digits: 4 % on the regression radius at 40 GeV).
"""
import numpy as np
import pytest

from tests import truth

N_EVENTS = 40  # 10 muons each, pT log-uniform in [0.5, 50] GeV, |eta| < 2.4


@pytest.fixture(scope="module")
def conn(oracle_geometry):
    return truth.Connectivity(oracle_geometry)


def _run(oracle_geometry, conn, exact):
    from oracle import orc
    from tracklooper_amd import synth
    pmap = truth.PixelMap(oracle_geometry)
    tot, rows = {"pt5": [0, 0], "pt3": [0, 0]}, []
    for i in range(N_EVENTS):
        ev = synth.generate_event("muon", 100 + i, exact=exact)
        o = orc.Event(oracle_geometry).run_lst(ev, "pt3")
        c = truth.truth_combinations(o, ev, conn)
        truth.add_counts(tot, c)
        p = truth.pixel_truth(o, ev, conn, pmap, c)
        for k in ("pt5", "pt3"):
            tot[k][0] += p[k][0]
            tot[k][1] += p[k][1]
        if c["t5_true"]:
            k = np.array([s for s, _ in c["t5_true"]])
            t = np.array([t for _, t in c["t5_true"]])
            rows.append(np.stack([ev["truth_pt"][t], o.field("t5.innerRadius")[k], o.field("t5.bridgeRadius")[k],
                                  o.field("t5.outerRadius")[k], o.field("t5.regressionRadius")[k],
                                  o.field("t5.chiSquared")[k], o.field("t5.rzChiSquared")[k]], 1))
    return tot, np.concatenate(rows)


@pytest.fixture(scope="module")
def exact_run(oracle_geometry, conn):
    return _run(oracle_geometry, conn, True)


def test_true_combinations_built_on_exact_helices(exact_run):
    tot, _ = exact_run
    print({k: f"{a}/{b}" for k, (a, b) in tot.items()})
    floor = {"md": 0.99, "ls": 0.99, "t3": 0.99, "t5": 0.99, "pt5": 0.97, "pt3": 0.5}
    for stage, (built, allowed) in tot.items():
        assert allowed > (300 if stage != "pt3" else 5), stage
        assert built / allowed >= floor[stage], f"{stage}: {built} of {allowed} true combinations built"


def test_true_combinations_built_at_detector_resolution(oracle_geometry, conn):
    tot, _ = _run(oracle_geometry, conn, False)
    print({k: f"{a}/{b}" for k, (a, b) in tot.items()})
    floor = {"md": 0.99, "ls": 0.99, "t3": 0.94, "t5": 0.96, "pt5": 0.97, "pt3": 0.5}
    for stage, (built, allowed) in tot.items():
        assert built / allowed >= floor[stage], f"{stage}: {built} of {allowed} true combinations built"


def test_quintuplet_fits_on_exact_helices(exact_run):
    from tracklooper_amd import synth
    _, r = exact_run
    sel = (r[:, 0] > 1.0) & (r[:, 0] < 5.0)
    assert sel.sum() > 100
    r = r[sel]
    R = r[:, 0] / synth.K_CURV  # helix radius in cm
    for col, name in ((1, "innerRadius"), (2, "bridgeRadius"), (3, "outerRadius"), (4, "regressionRadius")):
        rel = np.abs(r[:, col] / R - 1)
        print(f"{name}: max relative deviation from the helix radius {rel.max():.2e}")
        assert rel.max() < 1e-3, name
    print(f"chiSquared max {r[:, 5].max():.2e}, rzChiSquared max {r[:, 6].max():.2e}")
    assert r[:, 5].max() < 1e-4
    assert r[:, 6].max() < 1e-2
