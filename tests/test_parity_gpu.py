"""HIP engine vs oracle, stage by stage, on seeded synthetic events (GPU)."""
import numpy as np
import pytest

from tests.parity import DEEPEST, Report, compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(geom_dir):
    import tracklooper_amd as lst
    es = lst.LSTESData(geom_dir)
    return lst, es


def _run_both(engine, oracle_geometry, ev, upto, t5_mode=0):
    from oracle import orc
    lst, es = engine
    prod = lst.Event(es, t5_mode=t5_mode).run_lst(ev, upto)
    prod.synchronize()
    ref = orc.Event(oracle_geometry, t5_mode=t5_mode).run_lst(ev, upto)
    return prod, ref


@pytest.mark.parametrize("index", range(5))
def test_muon_gun_all_stages(engine, oracle_geometry, index):
    from tracklooper_amd import synth
    ev = synth.generate_event("muon", index)
    prod, ref = _run_both(engine, oracle_geometry, ev, DEEPEST)
    rep = compare(prod, ref, DEEPEST)
    assert ref.field("t3.nTriplets").sum() > 0
    assert ref.field("t5.nQuintuplets").sum() > 0
    assert ref.scalar("nTrackCandidates") > 0
    print("\n".join(rep.lines) or "bit-identical")


@pytest.mark.parametrize("n_pu", [50, 200])
def test_pileup_all_stages(engine, oracle_geometry, n_pu):
    from tracklooper_amd import synth
    ev = synth.generate_event("pu", 0, n_pu=n_pu)
    prod, ref = _run_both(engine, oracle_geometry, ev, DEEPEST)
    rep = compare(prod, ref, DEEPEST)
    n_t3 = ref.field("t3.nTriplets").sum()
    n_t5 = ref.field("t5.nQuintuplets").sum()
    n_tc = ref.scalar("nTrackCandidates")
    assert n_t3 > 1000 and n_t5 > 100 and n_tc > 100
    print(f"T3={n_t3} T5={n_t5} pT5={ref.scalar('nPixelQuintuplets')} pT3={ref.scalar('nPixelTriplets')} "
          f"TC={n_tc}; floats compared {rep.float_compared}, not bit-identical {rep.float_inexact}")
    print("\n".join(rep.lines))


def test_empty_event(engine, oracle_geometry):
    """No hits and no seeds: every stage runs and produces nothing."""
    from tracklooper_amd import synth
    ev = synth.generate_event("muon", 0)
    ev = {k: (v[:0] if k.startswith(("ph2_", "see_")) else v) for k, v in ev.items()}
    ev["see_hitIdx_off"] = np.zeros(1, np.uint32)
    prod, ref = _run_both(engine, oracle_geometry, ev, DEEPEST)
    compare(prod, ref, DEEPEST)
    assert prod.getNumberOfMiniDoublets() == 0 and prod.getNumberOfTriplets() == 0
    assert prod.getNumberOfQuintuplets() == 0
    assert prod.getNumberOfTrackCandidates() == 0 and len(prod.getOutput()["tc_len"]) == 0


def test_repeat_runs_identical(engine, oracle_geometry):
    """Reusing one event handle gives identical results (no state leaks between events)."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    _, es = engine
    e = lst.Event(es)
    ev0 = synth.generate_event("pu", 1, n_pu=50)
    ev1 = synth.generate_event("muon", 7)
    from tests.parity import slots

    def hits():
        sl = slots(e.field("ranges.tripletModuleIndices"), e.field("t3.nTriplets"), 6)
        q = slots(np.maximum(e.field("ranges.quintupletModuleIndices"), 0), e.field("t5.nQuintuplets"), 10)
        return e.field("t3.hitIndices")[sl], e.field("t5.hitIndices")[q], e.field("t5.isDup")[q[::10] // 10]

    e.run_lst(ev0, DEEPEST)
    a = hits()
    e.run_lst(ev1, DEEPEST)
    e.run_lst(ev0, DEEPEST)
    b = hits()
    assert len(a[0]) > 0 and len(a[1]) > 0
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_staged_device_path_matches(engine, oracle_geometry):
    """The throughput path (prepared once into pinned memory, copied to the device at each run)
    gives exactly the oracle's results, for a slot reused with another event."""
    import ctypes as C
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    from oracle import orc
    _, es = engine
    e = lst.Event(es)
    L = lst.lib()
    L.lst_event_stage_input.argtypes = [C.c_void_p, C.POINTER(lst._Input), C.c_uint32]
    L.lst_event_run_staged.argtypes = [C.c_void_p, C.c_uint32, C.c_int]
    for ev in (synth.generate_event("muon", 3), synth.generate_event("pu", 2, n_pu=100)):
        inp, keep = lst._make_input(ev)
        lst._check(L.lst_event_stage_input(e.h, C.byref(inp), 3))
        lst._check(L.lst_event_run_staged(e.h, 3, lst.STAGES[DEEPEST]))
        e.synchronize()
        ref = orc.Event(oracle_geometry).run_lst(ev, DEEPEST)
        compare(e, ref, DEEPEST)


def test_stage_by_stage_event_api(engine, oracle_geometry):
    """SDL::Event driven method by method (addHitToEvent ... createTrackCandidates) on the prepared
    inputs of a PU200 event gives the oracle's objects (SDL/Event.h:115-150)."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    from oracle import orc
    _, es = engine
    ev = synth.generate_event("pu", 9, n_pu=200)
    ref = orc.Event(oracle_geometry).run_lst(ev, DEEPEST)
    f = ref.field
    e = lst.Event(es)
    e.addHitToEvent(f("in_x"), f("in_y"), f("in_z"), f("in_detId"), f("in_idx"))
    e.addPixelSegmentToEvent(f("in_hitIdx0"), f("in_hitIdx1"), f("in_hitIdx2"), f("in_hitIdx3"), f("in_dPhiChange"),
                             f("in_ptIn"), f("in_ptErr"), f("in_px"), f("in_py"), f("in_pz"), f("in_eta"),
                             f("in_etaErr"), f("in_phi"), f("in_charge"), f("in_seedIdx"), f("in_superbin"),
                             f("in_pixelType"), f("in_isQuad"))
    with pytest.raises(lst.LSTError):  # out of order: triplets before segments
        e.createTriplets()
    e.createMiniDoublets()
    e.createSegmentsWithModuleMap()
    e.createTriplets()
    e.createQuintuplets()
    e.pixelLineSegmentCleaning()
    e.createPixelQuintuplets()
    e.createPixelTriplets()
    e.createTrackCandidates()
    e.synchronize()
    rep = compare(e, ref, DEEPEST)
    assert e.getNumberOfTrackCandidates() == ref.scalar("nTrackCandidates") > 1000
    print(f"stage by stage: TC={e.getNumberOfTrackCandidates()}; floats compared {rep.float_compared}")


def test_kernel_timing_does_not_disturb_runs(engine):
    """Per-kernel timing returns positive durations and leaves the event usable."""
    import ctypes as C
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    _, es = engine
    e = lst.Event(es)
    L = lst.lib()
    L.lst_event_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.lst_event_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    ev = synth.generate_event("pu", 4, n_pu=50)
    lst._check(L.lst_event_set_timing(e.h, 1))
    e.run_lst(ev, DEEPEST)
    ms = np.zeros(16, np.float32)
    lst._check(L.lst_event_kernel_times(e.h, ms.ctypes.data_as(C.c_void_p), 16))
    assert np.all(ms > 0)  # the full chain brackets all 16 timed kernels
    e2 = lst.Event(es)
    lst._check(L.lst_event_set_timing(e2.h, 1))
    e2.run_lst(ev, "md")  # hits, pixel segments and mini-doublets only
    lst._check(L.lst_event_kernel_times(e2.h, ms.ctypes.data_as(C.c_void_p), 16))
    assert np.all(ms[:3] > 0) and np.all(ms[3:] == 0)
    e.run_lst(ev, DEEPEST)  # must not report a stale runtime error
    lst._check(L.lst_event_set_timing(e.h, 0))
    e.run_lst(ev, DEEPEST)


def test_output_lists_match(engine, oracle_geometry):
    """lst_event_get_output returns the reference's four output lists (SDL/LST.cc:366-391)."""
    from tracklooper_amd import synth
    ev = synth.generate_event("pu", 3, n_pu=50)
    prod, ref = _run_both(engine, oracle_geometry, ev, "tc")
    out = prod.getOutput()
    n = ref.scalar("nTrackCandidates")
    assert len(out["tc_len"]) == n and n > 0
    off = ref.field("out.hitOffsets")
    flat = ref.field("out.hitIdxs")
    for i in range(n):
        assert list(out["tc_hitIdxs"][i]) == list(flat[off[i]:off[i + 1]])
    assert np.array_equal(out["tc_len"], ref.field("out.len")[:n])
    assert np.array_equal(out["tc_seedIdx"], ref.field("out.seedIdx")[:n])
    assert np.array_equal(out["tc_trackCandidateType"], ref.field("out.trackCandidateType")[:n])


def test_per_module_builders_match(oracle_geometry, geom_dir, monkeypatch):
    """The per-module T3/T5 builders (taken when an event overflows the
    flattened path's lists) give the same objects as the oracle."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    from oracle import orc
    monkeypatch.setenv("LST_PER_MODULE_BUILDERS", "1")
    es = lst.LSTESData(geom_dir)
    ev = synth.generate_event("pu", 7, n_pu=100)
    prod = lst.Event(es).run_lst(ev, DEEPEST)
    prod.synchronize()
    ref = orc.Event(oracle_geometry).run_lst(ev, DEEPEST)
    compare(prod, ref, DEEPEST)


def test_large_paths_match(oracle_geometry, geom_dir, monkeypatch):
    """The per-mini-doublet segment and T3 lists built by direct counting (the
    path of modules too large for the LDS counting sort) and the eta binnings
    that load their items twice (the path of counts too large for registers)
    give the same objects as the oracle."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    from oracle import orc
    monkeypatch.setenv("LST_LARGE_PATHS", "1")
    es = lst.LSTESData(geom_dir)
    ev = synth.generate_event("pu", 8, n_pu=200)
    prod = lst.Event(es).run_lst(ev, DEEPEST)
    prod.synchronize()
    ref = orc.Event(oracle_geometry).run_lst(ev, DEEPEST)
    compare(prod, ref, DEEPEST)


@pytest.mark.parametrize("case", [("muon", 0, 0), ("muon", 1, 0), ("pu", 0, 50), ("pu", 1, 200)])
def test_t5_dnn_mode_all_stages(engine, oracle_geometry, case):
    """DNN selection of the quintuplets (USE_RZCHI2 + USE_T5_DNN) through the whole chain."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    kind, index, n_pu = case
    ev = synth.generate_event(kind, index, n_pu=n_pu) if kind == "pu" else synth.generate_event(kind, index)
    prod, ref = _run_both(engine, oracle_geometry, ev, DEEPEST, t5_mode=lst.LST_T5_DNN)
    rep = compare(prod, ref, DEEPEST)
    assert ref.field("t5.nQuintuplets").sum() > 0
    print(f"DNN mode T5={ref.field('t5.nQuintuplets').sum()} TC={ref.scalar('nTrackCandidates')}; "
          f"floats compared {rep.float_compared}, not bit-identical {rep.float_inexact}")


def test_t5_dnn_mode_per_module_builder(engine, oracle_geometry, monkeypatch):
    """The per-module quintuplet builder evaluates the DNN inline: same results."""
    import tracklooper_amd as lst
    from tracklooper_amd import synth
    monkeypatch.setenv("LST_PER_MODULE_BUILDERS", "1")
    ev = synth.generate_event("pu", 2, n_pu=50)
    prod, ref = _run_both(engine, oracle_geometry, ev, DEEPEST, t5_mode=lst.LST_T5_DNN)
    compare(prod, ref, DEEPEST)


def test_physics_metrics_match_oracle(engine, oracle_geometry):
    """Efficiency / fake / duplicate flags computed on the engine's output equal the oracle's."""
    from tracklooper_amd import metrics, synth
    ev = synth.generate_event("pu", 5, n_pu=50)
    prod, ref = _run_both(engine, oracle_geometry, ev, "tc")
    flat, off, typ = ref.field("out.hitIdxs"), ref.field("out.hitOffsets"), ref.field("out.trackCandidateType")
    n = ref.scalar("nTrackCandidates")
    ref_out = {"tc_hitIdxs": [list(flat[off[i]:off[i + 1]]) for i in range(n)], "tc_trackCandidateType": typ[:n]}
    a = metrics.event_metrics(prod.getOutput(), ev)
    b = metrics.event_metrics(ref_out, ev)
    np.testing.assert_array_equal(a["tc_is_fake"], b["tc_is_fake"])
    np.testing.assert_array_equal(a["tc_is_duplicate"], b["tc_is_duplicate"])
    np.testing.assert_array_equal(a["sim_tc_matched"], b["sim_tc_matched"])
    s = metrics.summarize([a])
    assert s["efficiency"] > 0.9 and s["fake_rate"] < 0.05
    print(s)


@pytest.fixture(scope="module")
def pt06(geom_dir):
    """The pt0.6 geometry (module-map fan-out 68) with pT cut 0.6: the DNN configuration."""
    import tracklooper_amd as lst
    from oracle import orc
    from tracklooper_amd import synth
    d = synth.geometry_dir("OT800_IT615_pt0.6")
    return lst.LSTESData(d, pt_cut=0.6), orc.Geometry(d, stride=68)


@pytest.mark.parametrize("case", [("muon", 0, 1, 0), ("pu", 0, 0, 50), ("pu", 1, 1, 50), ("pu", 2, 0, 200),
                                  ("pu", 2, 1, 200), ("pu", 3, 1, 250)])
def test_pt06_geometry_all_stages(pt06, case):
    """The pt0.6 configuration (denser combinatorics; at PU200-250 the per-module capacities and the
    flattened builders' work arrays are under the most pressure), plain and DNN T5 selection."""
    import tracklooper_amd as lst
    from oracle import orc
    from tracklooper_amd import synth
    es, geom = pt06
    kind, index, mode, n_pu = case
    g = "OT800_IT615_pt0.6"
    ev = synth.generate_event(kind, index, geom_name=g) if kind == "muon" else \
        synth.generate_event(kind, index, n_pu=n_pu, geom_name=g)
    prod = lst.Event(es, t5_mode=mode).run_lst(ev, DEEPEST)
    prod.synchronize()
    ref = orc.Event(geom, pt_cut=0.6, t5_mode=mode).run_lst(ev, DEEPEST)
    rep = compare(prod, ref, DEEPEST)
    assert ref.field("t5.nQuintuplets").sum() > 0 and ref.scalar("nTrackCandidates") > 0
    print(f"pt0.6 mode {mode}: T5={ref.field('t5.nQuintuplets').sum()} TC={ref.scalar('nTrackCandidates')}; "
          f"floats compared {rep.float_compared}, not bit-identical {rep.float_inexact}")


def test_t5dnn_record_matches_oracle(engine, oracle_geometry):
    """Training rows of the T5 DNN (candidates at its call site): inputs, score, triplet pair and
    hit indices equal the oracle's, row for row (rows sorted by triplet pair)."""
    import tracklooper_amd as lst
    from oracle import orc
    from tracklooper_amd import synth
    _, es = engine
    ev = synth.generate_event("pu", 4, n_pu=50)
    e = lst.Event(es, t5_mode=lst.LST_T5_DNN).record_t5dnn().run_lst(ev, "t5")
    rec = e.t5dnn_record()
    ref = orc.Event(oracle_geometry, t5_mode=1).run_lst(ev, "t5")
    rpair = ref.field("dnn.t3Pair").reshape(-1, 2)
    assert len(rpair) > 500 and len(rec["t3_pair"]) == len(rpair)
    a = np.lexsort((rec["t3_pair"][:, 1], rec["t3_pair"][:, 0]))
    b = np.lexsort((rpair[:, 1], rpair[:, 0]))
    np.testing.assert_array_equal(rec["t3_pair"][a], rpair[b])
    np.testing.assert_array_equal(rec["features"][a].view(np.uint32),
                                  ref.field("dnn.features").reshape(-1, 38)[b].view(np.uint32))
    np.testing.assert_array_equal(rec["score"][a].view(np.uint32), ref.field("dnn.score")[b].view(np.uint32))
    np.testing.assert_array_equal(rec["hit_idxs"][a], ref.field("dnn.hitIdxs").reshape(-1, 10)[b])
