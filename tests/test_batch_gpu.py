"""Batched engine (lst_batch_*): per event the same output lists as the
per-event path (SDL/LST.cc:366-420) and as the oracle, for host inputs,
staged inputs, partial launch sets and an uneven 64-event batch (GPU)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def es(geom_dir):
    import tracklooper_amd as lst
    return lst.LSTESData(geom_dir)


@pytest.fixture(scope="module")
def events():
    from tracklooper_amd import synth
    evs = [synth.generate_event("muon", i) for i in range(3)]
    evs += [synth.generate_event("pu", i, n_pu=50) for i in range(3)]
    evs.append(synth.generate_event("pu", 7, n_pu=200))
    return evs


def _per_event(es, events):
    import tracklooper_amd as lst
    outs = []
    e = lst.Event(es)
    for ev in events:
        e.run_lst(ev, "tc")
        outs.append(e.getOutput())
    e.close()
    return outs


@pytest.fixture(scope="module")
def per_event(es, events):
    return _per_event(es, events)


def _same(a, b):
    assert len(a["tc_hitIdxs"]) == len(b["tc_hitIdxs"])
    assert a["tc_hitIdxs"] == b["tc_hitIdxs"]
    np.testing.assert_array_equal(a["tc_len"], b["tc_len"])
    np.testing.assert_array_equal(a["tc_seedIdx"], b["tc_seedIdx"])
    np.testing.assert_array_equal(a["tc_trackCandidateType"], b["tc_trackCandidateType"])


@pytest.mark.parametrize("slots", [1, 3, 8])
def test_batch_host_inputs_match_per_event(es, events, per_event, slots):
    """7 events in launch sets of 1, 3 (the last one partial) or 8 (one partial set) events."""
    import tracklooper_amd as lst
    b = lst.Batch(es, slots=slots)
    res = b.run(events)
    for r, p in zip(res, per_event):
        assert r["status"] == 0
        assert r["n_tc"] == len(p["tc_hitIdxs"])
        _same(r, p)
    assert sum(r["n_tc"] for r in res) > 1000
    b.close()


def test_batch_staged_inputs_match_per_event_twice(es, events, per_event):
    import tracklooper_amd as lst
    b = lst.Batch(es, slots=4)
    b.stage(events)
    for _ in range(2):  # arenas are reused; results must not depend on what ran before
        res = b.run_staged(0, len(events))
        for r, p in zip(res, per_event):
            _same(r, p)
    res = b.run_staged(5, 4)  # staged events 5, 6, 0, 1
    for r, i in zip(res, [5, 6, 0, 1]):
        _same(r, per_event[i])
    b.close()


def test_batch_two_sets_in_flight_match_per_event(es, events, per_event):
    """Two launch sets in flight (each with its own arenas and stream): 7 staged events in sets of 2 (both
    sets used three times and more), host inputs, then back to one set in flight: same lists."""
    import tracklooper_amd as lst
    b = lst.Batch(es, slots=2)
    b.set_sets_in_flight(2)
    b.stage(events)
    for _ in range(2):
        res = b.run_staged(0, 3 * len(events))  # staged event i % 7, 21 events in 11 launch sets
        for k, r in enumerate(res):
            assert r["status"] == 0
            _same(r, per_event[k % len(events)])
    for r, p in zip(b.run(events), per_event):
        _same(r, p)
    b.set_sets_in_flight(1)
    for r, p in zip(b.run(events), per_event):
        _same(r, p)
    b.close()


def test_batch_without_graphs_matches(es, events, per_event, monkeypatch):
    """LST_BATCH_GRAPHS=0 launches the kernels one by one: same lists."""
    import tracklooper_amd as lst
    monkeypatch.setenv("LST_BATCH_GRAPHS", "0")
    b = lst.Batch(es, slots=4)
    for r, p in zip(b.run(events), per_event):
        _same(r, p)
    b.close()


def test_batch_overflow_rerun_matches(es, events, per_event, monkeypatch):
    """With LST_PER_MODULE_BUILDERS=1 every event reports an overflow of the flattened builders and is run
    again by the batch's per-event handle (the per-module builders): same lists."""
    import tracklooper_amd as lst
    monkeypatch.setenv("LST_PER_MODULE_BUILDERS", "1")
    b = lst.Batch(es, slots=4)
    for r, p in zip(b.run(events[3:]), per_event[3:]):
        _same(r, p)
    b.close()


def test_batch_tuning_keeps_results(es, events, per_event):
    """The autotuner's work-group counts (and extreme hand-set ones) change no output."""
    import tracklooper_amd as lst
    b = lst.Batch(es, slots=len(events))
    b.stage(events)
    d_ms, t_ms, table = b.autotune(0, repeats=1)
    assert t_ms <= d_ms * 1.05 and any(table.values())
    for r, p in zip(b.run_staged(0, len(events)), per_event):
        _same(r, p)
    b.set_tuning({k: 1 for k in table})  # one work-group per event for every grid-stride kernel
    for r, p in zip(b.run_staged(0, len(events)), per_event):
        _same(r, p)
    b.close()


def test_batch_bad_event_reported_alone(es, events, per_event):
    """An event with a detId absent from the geometry gets LST_E_INPUT in its status; the other events of
    its launch set are unaffected and the call reports the error."""
    import tracklooper_amd as lst
    bad = dict(events[3])
    det = bad["ph2_detId"].copy()
    det[det == det[0]] = 7  # no module has detId 7; the hits stay contiguous
    bad["ph2_detId"] = det
    b = lst.Batch(es, slots=4)
    evs = [events[0], bad, events[4], events[5]]
    with pytest.raises(lst.LSTError):
        b.run(evs)
    arr, keep = lst.Batch._inputs(evs)
    outs, bufs = lst.Batch._outputs(len(evs), 8192)
    status = lst.lib().lst_run_batch(b.h, arr, len(evs), outs)
    assert status & 0xFF == 3
    res = lst.Batch._collect(outs, bufs)
    assert res[1]["status"] & 0xFF == 3
    for i, j in ((0, 0), (2, 4), (3, 5)):
        assert res[i]["status"] == 0
        _same(res[i], per_event[j])
    b.close()


@pytest.fixture(scope="module")
def c3_events():
    """The first uneven batch: 64 events, pileup U{50..250} each."""
    from tracklooper_amd import synth
    return synth.c3_batch(0, 64, workers=8)


@pytest.mark.timeout(600)
def test_c3_uneven_batch_matches_oracle(es, c3_events, oracle_geometry):
    """64 events of uneven occupancy in one launch set: every event's track-candidate lists equal the
    per-event path's and the oracle's (oracle events run on 16 host threads)."""
    import tracklooper_amd as lst
    from oracle import orc
    from tracklooper_amd import synth
    pus = synth.c3_pileups(0, 64)
    assert pus.min() < 70 and pus.max() > 230
    b = lst.Batch(es, slots=64)
    b.stage(c3_events)
    b.autotune(0, repeats=1)
    res = b.run_staged(0, 64)
    b.close()
    per = _per_event(es, c3_events)

    def oracle_out(ev):
        r = orc.Event(oracle_geometry).run_lst(ev, "tc")
        n = r.scalar("nTrackCandidates")
        flat, off = r.field("out.hitIdxs"), r.field("out.hitOffsets")
        return n, [list(flat[off[i]:off[i + 1]]) for i in range(n)], r.field("out.trackCandidateType")[:n].copy(), \
            r.field("out.seedIdx")[:n].copy()

    with ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(oracle_out, c3_events))
    hits = [len(ev["ph2_x"]) for ev in c3_events]
    truncated = 0
    for i, (r, p, (n, ref_hits, ref_type, ref_seed)) in enumerate(zip(res, per, refs)):
        assert r["status"] & 0xFF == 0, i  # LST_W_TRUNCATED (0x100) is allowed: capacities truncate as in the reference
        truncated += bool(r["status"] & 0x100)
        _same(r, p)
        assert r["n_tc"] == n, i
        assert r["tc_hitIdxs"] == ref_hits, i
        np.testing.assert_array_equal(r["tc_trackCandidateType"], ref_type)
        np.testing.assert_array_equal(r["tc_seedIdx"], ref_seed)
    print(f"64 events, {min(hits)}..{max(hits)} hits, {sum(r['n_tc'] for r in res)} track candidates, "
          f"{truncated} events truncated at a capacity, identical to the per-event path and the oracle")
