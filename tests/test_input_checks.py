"""Input handling on the CPU: the ctypes input builder converts every array to
the C type of lst_input, and the oracle follows the reference's index checks
for modules whose int8 hit counts wrap (SDL/Hit.h:31-32, SDL/MiniDoublet.h:919-923)."""
import numpy as np
import pytest


def _event():
    from tracklooper_amd import synth
    return synth.generate_event("muon", 2)


def test_make_input_converts_types():
    import tracklooper_amd as lst
    ev = _event()
    wide = {k: (v.astype(np.float64) if v.dtype == np.float32 else v.astype(np.int64))
            for k, v in ev.items() if k.startswith(("see_", "ph2_"))}
    inp, keep = lst._make_input(wide)
    ref, _ = lst._make_input(ev)
    got = dict(zip([k for k in wide], keep))
    for k, a in got.items():
        want = lst._INPUT_TYPES.get(k, np.float32)
        assert a.dtype == want and a.flags.c_contiguous, k
        np.testing.assert_array_equal(a, np.asarray(ev[k], want))
    assert inp.n_seeds == ref.n_seeds and inp.n_hits == ref.n_hits


def test_make_input_rejects_bad_offsets():
    import tracklooper_amd as lst
    ev = dict(_event())
    ev["see_hitIdx_off"] = ev["see_hitIdx_off"][:-1]
    with pytest.raises(lst.LSTError):
        lst._make_input(ev)
    ev = dict(_event())
    off = ev["see_hitIdx_off"].copy()
    off[-1] = len(ev["see_hitIdx"]) + 5
    ev["see_hitIdx_off"] = off
    with pytest.raises(lst.LSTError):
        lst._make_input(ev)


def crowded_event(extra=130):
    """A muon event with `extra` more hits on both modules of one module pair that already holds a
    mini-doublet (hits kept grouped by module), so that both int8 hit counts exceed 127; returns the
    event and the two detIds."""
    from oracle import orc
    from tracklooper_amd import synth
    ev = _event()
    g = orc.Geometry(synth.geometry_dir())
    base = orc.Event(g).run_lst(ev, "md")
    n_md = base.field("mds.nMDs")
    det = ev["ph2_detId"]
    mods = base.field("hits.moduleIndices")[:len(det)]
    partner = g.field("partnerModuleIndices")
    for lo in np.unique(mods):
        if lo >= len(n_md) - 1 or n_md[lo] == 0:
            continue
        up = partner[lo]
        d_lo, d_up = det[mods == lo][0], det[mods == up][0]
        break
    rng = np.random.default_rng(5)
    out = {k: v for k, v in ev.items()}
    x, y, z, d = [ev[k].copy() for k in ("ph2_x", "ph2_y", "ph2_z", "ph2_detId")]
    for dd in (d_lo, d_up):
        at = int(np.nonzero(d == dd)[0][-1]) + 1
        k = int(np.nonzero(d == dd)[0][0])
        jit = rng.normal(0, 0.01, (extra, 3)).astype(np.float32)
        x = np.insert(x, at, x[k] + jit[:, 0])
        y = np.insert(y, at, y[k] + jit[:, 1])
        z = np.insert(z, at, z[k] + jit[:, 2])
        d = np.insert(d, at, np.full(extra, dd, np.uint32))
    # the pixel seeds' hit indices address the pixel hits appended after the OT hits: unchanged
    out.update(ph2_x=x, ph2_y=y, ph2_z=z, ph2_detId=d)
    return out, int(d_lo), int(d_up)


def test_oracle_builds_no_mini_doublet_on_wrapped_counts():
    """More than 127 hits on both modules of a pair: the int8 counts wrap negative and the reference's
    index checks skip every hit pair of the module, so it holds no mini-doublet; other modules are
    unaffected."""
    from oracle import orc
    from tracklooper_amd import synth
    g = orc.Geometry(synth.geometry_dir())
    ev, d_lo, d_up = crowded_event()
    r = orc.Event(g).run_lst(ev, "md")
    det = ev["ph2_detId"]
    mods = r.field("hits.moduleIndices")[:len(det)]
    lo = int(mods[det == d_lo][0])
    assert r.field("hits.hitRangesnLower")[lo] < 0 and r.field("hits.hitRangesnUpper")[lo] < 0
    assert r.field("mds.nMDs")[lo] == 0
    assert r.field("mds.nMDs").sum() > 0
    # with 127 hits on each the counts are valid and the module builds mini-doublets again
    orig = _event()["ph2_detId"]
    ev2, _, _ = crowded_event(extra=127 - max(int((orig == d_lo).sum()), int((orig == d_up).sum())))
    r2 = orc.Event(g).run_lst(ev2, "md")
    lo2 = int(r2.field("hits.moduleIndices")[:len(ev2["ph2_detId"])][ev2["ph2_detId"] == d_lo][0])
    assert 0 < r2.field("hits.hitRangesnLower")[lo2] <= 127
    assert r2.field("mds.nMDs")[lo2] > 0
