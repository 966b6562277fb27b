"""The truth invariants on the HIP engine (GPU): on exact-helix muon events the
engine builds the same truth-matched combinations as the oracle, at least
99 % of those the tracks' hits allow at every stage (see
tests/test_truth_invariants.py for the definitions)."""
import pytest

from tests import truth

pytestmark = pytest.mark.gpu


def test_engine_builds_the_true_combinations(geom_dir, oracle_geometry):
    import tracklooper_amd as lst
    from oracle import orc
    from tracklooper_amd import synth
    es = lst.LSTESData(geom_dir)
    conn = truth.Connectivity(oracle_geometry)
    prod_tot, ref_tot = {}, {}
    e = lst.Event(es)
    for i in range(20):
        ev = synth.generate_event("muon", 100 + i, exact=True)
        truth.add_counts(prod_tot, truth.truth_combinations(e.run_lst(ev, "t5"), ev, conn))
        truth.add_counts(ref_tot, truth.truth_combinations(orc.Event(oracle_geometry).run_lst(ev, "t5"), ev, conn))
    e.close()
    print({k: f"{a}/{b}" for k, (a, b) in prod_tot.items()})
    assert prod_tot == ref_tot
    for stage, (built, allowed) in prod_tot.items():
        assert allowed > 100 and built / allowed >= 0.99, stage
