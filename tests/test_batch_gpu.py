"""Batched engine (lst_batch_*): per event the same output lists as the
per-event path (SDL/LST.cc:366-420), for host inputs and staged inputs (GPU)."""
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


@pytest.fixture(scope="module")
def per_event(es, events):
    import tracklooper_amd as lst
    outs = []
    for ev in events:
        e = lst.Event(es).run_lst(ev, "tc")
        outs.append(e.getOutput())
        e.close()
    return outs


def _same(a, b):
    assert len(a["tc_hitIdxs"]) == len(b["tc_hitIdxs"])
    assert a["tc_hitIdxs"] == b["tc_hitIdxs"]
    np.testing.assert_array_equal(a["tc_len"], b["tc_len"])
    np.testing.assert_array_equal(a["tc_seedIdx"], b["tc_seedIdx"])
    np.testing.assert_array_equal(a["tc_trackCandidateType"], b["tc_trackCandidateType"])


@pytest.mark.parametrize("n_contexts", [1, 3])
def test_batch_host_inputs_match_per_event(es, events, per_event, n_contexts):
    import tracklooper_amd as lst
    b = lst.Batch(es, n_contexts=n_contexts)
    res = b.run(events)
    for r, p in zip(res, per_event):
        assert r["n_tc"] == len(p["tc_hitIdxs"])
        _same(r, p)
    assert sum(r["n_tc"] for r in res) > 1000


def test_batch_staged_inputs_match_per_event_twice(es, events, per_event):
    import tracklooper_amd as lst
    b = lst.Batch(es, n_contexts=4)
    b.stage(events)
    for _ in range(2):  # contexts are reused; results must not depend on what ran before
        res = b.run_staged(0, len(events))
        for r, p in zip(res, per_event):
            _same(r, p)
    # a sub-range of the slots
    res = b.run_staged(3, 2)
    for r, p in zip(res, per_event[3:5]):
        _same(r, p)


def test_batch_output_truncation_reports_full_count(es, events, per_event):
    import tracklooper_amd as lst
    b = lst.Batch(es, n_contexts=2)
    cap = 5
    res = b.run(events[3:5], capacity=cap)
    for r, p in zip(res, per_event[3:5]):
        n = len(p["tc_hitIdxs"])
        assert n > cap and r["n_tc"] == n
        assert r["tc_hitIdxs"] == p["tc_hitIdxs"][:cap]
        np.testing.assert_array_equal(r["tc_trackCandidateType"], p["tc_trackCandidateType"][:cap])


def test_batch_dnn_mode_matches_per_event(es, events):
    import tracklooper_amd as lst
    sub = events[3:6]
    want = []
    for ev in sub:
        e = lst.Event(es, t5_mode=lst.LST_T5_DNN).run_lst(ev, "tc")
        want.append(e.getOutput())
        e.close()
    res = lst.Batch(es, n_contexts=2, t5_mode=lst.LST_T5_DNN).run(sub)
    for r, p in zip(res, want):
        _same(r, p)


@pytest.mark.parametrize("graphs", ["1", "0"])
def test_batch_one_context_replays_match(es, events, per_event, monkeypatch, graphs):
    """One context runs every staged event in turn (its captured chain replayed per event, or, with
    LST_BATCH_GRAPHS=0, every kernel launched); each replay equals the per-event result."""
    import tracklooper_amd as lst
    monkeypatch.setenv("LST_BATCH_GRAPHS", graphs)
    b = lst.Batch(es, n_contexts=1)
    b.stage(events)
    res = b.run_staged(0, 2 * len(events))
    for i, r in enumerate(res):
        _same(r, per_event[i % len(events)])
