"""Bad inputs the reference cannot represent (SDL/Hit.h:31-32,245-246,261-267)
are reported as LST_E_INPUT: by the host checks of lst_run before any upload,
and by the device-side checks for hits added with addHitToEvent (GPU)."""
import numpy as np
import pytest

from tests.parity import compare
from tests.test_input_checks import crowded_event

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def es(geom_dir):
    import tracklooper_amd as lst
    return lst.LSTESData(geom_dir)


def _muon():
    from tracklooper_amd import synth
    return synth.generate_event("muon", 4)


def _unknown_det(ev):
    ev = dict(ev)
    det = ev["ph2_detId"].copy()
    det[det == det[0]] = 7  # no module has detId 7
    ev["ph2_detId"] = det
    return ev


def _non_contiguous(ev):
    ev = dict(ev)
    det = ev["ph2_detId"].copy()
    det[-1] = det[0]  # the first module's detId once more at the end
    ev["ph2_detId"] = det
    return ev


@pytest.mark.parametrize("make,msg", [(_unknown_det, "no module"), (_non_contiguous, "contiguous"),
                                      (lambda ev: crowded_event()[0], "127")])
def test_lst_run_rejects_bad_hits(es, make, msg):
    import tracklooper_amd as lst
    e = lst.Event(es)
    with pytest.raises(lst.LSTError, match=msg):
        e.run_lst(make(_muon()), "tc")
    e.run_lst(_muon(), "tc")  # the handle stays usable
    assert e.getNumberOfTrackCandidates() > 0


@pytest.mark.parametrize("make,msg", [(_unknown_det, "no module"), (_non_contiguous, "contiguous")])
def test_device_checks_report_bad_hits(es, oracle_geometry, make, msg):
    """Hits added with addHitToEvent are checked on the device; the status comes with the next
    synchronisation."""
    import tracklooper_amd as lst
    from oracle import orc
    good = orc.Event(oracle_geometry).run_lst(_muon(), "md")
    f = good.field
    x, y, z, det, idx = f("in_x"), f("in_y"), f("in_z"), f("in_detId").copy(), f("in_idx")
    n_ot = len(_muon()["ph2_x"])
    if make is _unknown_det:
        det[det == det[0]] = 7
    else:
        det[n_ot - 1] = det[0]
    e = lst.Event(es)
    e.addHitToEvent(x, y, z, det, idx)
    with pytest.raises(lst.LSTError, match=msg):
        e.synchronize()


def test_per_module_builder_on_wrapped_counts_matches_oracle(es, oracle_geometry, monkeypatch):
    """A module pair with more than 127 hits on each side, added without the host checks and built by the
    per-module mini-doublet builder: no out-of-range reads, no mini-doublets on that module, the other
    modules' objects equal the oracle's, and the device checks report LST_E_INPUT."""
    import tracklooper_amd as lst
    from oracle import orc
    monkeypatch.setenv("LST_PER_MODULE_BUILDERS", "1")
    ev, d_lo, _ = crowded_event()
    ref = orc.Event(oracle_geometry).run_lst(ev, "md")
    f = ref.field
    e = lst.Event(es)
    e.addHitToEvent(f("in_x"), f("in_y"), f("in_z"), f("in_detId"), f("in_idx"))
    e.addPixelSegmentToEvent(f("in_hitIdx0"), f("in_hitIdx1"), f("in_hitIdx2"), f("in_hitIdx3"), f("in_dPhiChange"),
                             f("in_ptIn"), f("in_ptErr"), f("in_px"), f("in_py"), f("in_pz"), f("in_eta"),
                             f("in_etaErr"), f("in_phi"), f("in_charge"), f("in_seedIdx"), f("in_superbin"),
                             f("in_pixelType"), f("in_isQuad"))
    e.createMiniDoublets()
    with pytest.raises(lst.LSTError, match="127"):
        e.synchronize()
    compare(e, ref, "md")
    lo = int(f("hits.moduleIndices")[:len(ev["ph2_detId"])][ev["ph2_detId"] == d_lo][0])
    assert e.field("mds.nMDs")[lo] == 0
