"""Pins the oracle (CPU restatement) to what the mathematics and the shipped
geometry fix, so that a plausible transcription mistake fails here (CPU only).

The reference ships no golden vectors or fixtures for this path (it has no
tests); what pins the oracle is: closed forms (circumcircle, phi wrapping,
pseudorapidity, binary search, the superbin formula), the geometry tables'
own structure, and invariants of the stage outputs on seeded events.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import orc


@pytest.fixture(scope="module")
def L():
    orc.build()
    lib = orc.lib()
    for n in ("orc_phi_mpi_pi",):
        getattr(lib, n).restype = C.c_float
        getattr(lib, n).argtypes = [C.c_float]
    lib.orc_phi.restype = C.c_float
    lib.orc_phi.argtypes = [C.c_float] * 2
    lib.orc_delta_phi.restype = C.c_float
    lib.orc_delta_phi.argtypes = [C.c_float] * 4
    lib.orc_eta.restype = C.c_float
    lib.orc_eta.argtypes = [C.c_float] * 3
    lib.orc_binary_search.restype = C.c_int
    lib.orc_binary_search.argtypes = [C.c_void_p, C.c_uint, C.c_uint]
    lib.orc_circle_radius.restype = C.c_float
    lib.orc_circle_radius.argtypes = [C.c_float] * 6 + [C.POINTER(C.c_float)] * 2
    return lib


def test_phi_wrapping(L):
    rng = np.random.default_rng(1)
    for x in rng.uniform(-40, 40, 500):
        w = L.orc_phi_mpi_pi(float(x))
        assert -math.pi - 1e-5 <= w <= math.pi + 1e-5
        assert abs(math.remainder(w - np.float32(x), 2 * math.pi)) < 1e-4
    # special cases: inside (-pi, pi] returned unchanged
    for x in (0.0, 1.0, -3.0, 3.14159):
        assert L.orc_phi_mpi_pi(x) == np.float32(x)


def test_phi_matches_atan2(L):
    rng = np.random.default_rng(2)
    for x, y in rng.normal(0, 50, (500, 2)):
        p = L.orc_phi(float(x), float(y))
        assert abs(math.remainder(p - math.atan2(np.float32(y), np.float32(x)), 2 * math.pi)) < 2e-6
    # delta phi is antisymmetric and wraps
    assert abs(L.orc_delta_phi(1, 0, 0, 1) - math.pi / 2) < 1e-6
    assert abs(L.orc_delta_phi(0, 1, 1, 0) + math.pi / 2) < 1e-6
    assert abs(abs(L.orc_delta_phi(-1, 1e-3, -1, -1e-3)) - 2e-3) < 1e-6


def test_eta_closed_form(L):
    rng = np.random.default_rng(3)
    for x, y, z in rng.normal(0, 80, (500, 3)):
        e = L.orc_eta(float(x), float(y), float(z))
        exact = math.asinh(np.float32(z) / math.hypot(np.float32(x), np.float32(y)))
        # the reference form sign(z) acosh(r3/rt) is ill-conditioned at eta ~ 0: a one-ulp
        # error in r3/rt moves acosh by up to sqrt(2 * 2^-23)
        tol = max(1e-5 * max(1.0, abs(exact)), math.sqrt(2 * 2.0 ** -22))
        assert abs(e - exact) < tol
    assert L.orc_eta(10.0, 0.0, 0.0) == 0.0


def test_circumcircle_closed_form(L):
    rng = np.random.default_rng(4)
    g, f = C.c_float(), C.c_float()
    for _ in range(300):
        # well-conditioned triangles: arcs of 0.5-1.5 rad between the points
        cx, cy = rng.normal(0, 30, 2)
        R = rng.uniform(30, 300)
        a = rng.uniform(0, 2 * np.pi) + np.cumsum(rng.uniform(0.5, 1.5, 3))
        xs, ys = cx + R * np.cos(a), cy + R * np.sin(a)
        r = L.orc_circle_radius(*map(float, (xs[0], ys[0], xs[1], ys[1], xs[2], ys[2])), C.byref(g), C.byref(f))
        # float32 algebra on squared coordinates: ~1e-4 relative is what the formula offers
        assert abs(r - R) / R < 1e-3
        assert math.hypot(g.value - cx, f.value - cy) / R < 1e-3
    # collinear points -> -1 (SDL/Triplet.h:785-790)
    assert L.orc_circle_radius(0, 0, 1, 1, 2, 2, C.byref(g), C.byref(f)) == -1.0


def test_binary_search(L):
    data = np.array([1, 5, 9, 100, 4000000000], np.uint32)
    p = data.ctypes.data_as(C.c_void_p)
    for i, v in enumerate(data):
        assert L.orc_binary_search(p, int(v), len(data)) == i
    assert L.orc_binary_search(p, 7, len(data)) == -1


def test_geometry_structure(oracle_geometry):
    g = oracle_geometry
    n, nl = g.scalar("nModules"), g.scalar("nLowerModules")
    assert (n, nl) == (26401, 13200)
    det = g.field("detIds")
    partner = g.field("partnerModuleIndices")
    lower = g.field("isLower").astype(bool)
    # index layout: lower modules, then the pixel module, then upper modules
    assert lower[:nl].all() and not lower[nl:].any() and det[nl] == 1
    # partner relation is an involution between lower and upper modules
    m = np.arange(n) != nl
    assert np.all(partner[partner[m]] == np.arange(n)[m])
    assert np.all(partner[:nl] > nl)
    # detId bit fields
    d = det.astype(np.int64)
    sub = (d >> 25) & 7
    assert np.array_equal(g.field("subdets")[m], sub[m])
    # sdl layers: barrel 1-6, endcap PS 7-11, endcap 2S 12-16 (SDL/ModuleMethods.h:363-364)
    sdl = g.field("sdlLayers")
    lay = g.field("layers")
    mt = g.field("moduleType")
    exp = lay + 6 * (sub == 4) + 5 * ((sub == 4) & (mt == 1))
    assert np.array_equal(sdl[m], exp[m])
    # module map: fan-out and targets are lower modules
    ncon = g.field("nConnectedModules")
    assert ncon.max() == 34 and (ncon[:nl] > 0).sum() == 9656
    mm = g.field("moduleMap").reshape(n, -1)
    for i in np.nonzero(ncon)[0][:2000]:
        assert np.all(mm[i, : ncon[i]] < nl)
    # pixel map sizes add up (SDL/ModuleMethods.h:87-88)
    tot = sum(int(g.field(k).sum()) for k in ("connectedPixelsSizes", "connectedPixelsSizesPos",
                                             "connectedPixelsSizesNeg"))
    assert tot == g.scalar("nPixels") == 1795336


def test_stage_invariants_on_seeded_events(oracle_geometry):
    from tracklooper_amd import synth
    from tests.parity import slots
    for idx in range(3):
        ev = synth.generate_event("muon", idx)
        e = orc.Event(oracle_geometry).run_lst(ev, "t3")
        nl = 13200
        # MD: both hits on the module pair, anchor/outer split as the module type says
        n = e.field("mds.nMDs")[:nl]
        start = e.field("ranges.miniDoubletModuleIndices")[:nl]
        sl = slots(start, n)
        hm = e.field("hits.moduleIndices")
        mod = e.field("mds.moduleIndices")[sl]
        a, o = e.field("mds.anchorHitIndices")[sl], e.field("mds.outerHitIndices")[sl]
        partner = oracle_geometry.field("partnerModuleIndices")
        pair = np.sort(np.stack([hm[a], hm[o]], 1), 1)
        exp = np.sort(np.stack([mod, partner[mod]], 1), 1)
        assert np.array_equal(pair, exp)
        # LS: outer module connected to inner module; MDs from those modules
        ns = e.field("segs.nSegments")[:nl]
        ss = slots(e.field("ranges.segmentModuleIndices")[:nl], ns)
        md = e.field("segs.mdIndices").reshape(-1, 2)[ss]
        mdmod = e.field("mds.moduleIndices")
        inner, outer = e.field("segs.innerLowerModuleIndices")[ss], e.field("segs.outerLowerModuleIndices")[ss]
        assert np.array_equal(mdmod[md[:, 0]], inner) and np.array_equal(mdmod[md[:, 1]], outer)
        mm = oracle_geometry.field("moduleMap").reshape(26401, -1)
        ncon = oracle_geometry.field("nConnectedModules")
        for i, j in zip(inner, outer):
            assert j in mm[i, : ncon[i]]
        # T3: consecutive segments share the middle mini-doublet
        nt = e.field("t3.nTriplets")
        ts = slots(e.field("ranges.tripletModuleIndices"), nt)
        seg = e.field("t3.segmentIndices").reshape(-1, 2)[ts]
        allmd = e.field("segs.mdIndices").reshape(-1, 2)
        assert np.array_equal(allmd[seg[:, 0], 1], allmd[seg[:, 1], 0])
        # occupancy bookkeeping
        assert np.all(e.field("t3.totOccupancyTriplets") >= nt)


def test_prepare_input_closed_forms(oracle_geometry):
    from tracklooper_amd import synth
    ev = synth.generate_event("pu", 3, n_pu=20)
    e = orc.Event(oracle_geometry).run_lst(ev, "hits")
    pt = np.hypot(ev["see_stateTrajGlbPx"].astype(np.float64), ev["see_stateTrajGlbPy"])
    keep = pt > 0.8 - 2 * ev["see_ptErr"].astype(np.float32)
    seed = e.field("in_seedIdx")
    assert np.array_equal(seed, np.nonzero(keep)[0])
    # pLS eta = asinh(pz/pt) of the last-hit momentum (ROOT Eta_FromRhoZ closed form)
    exp_eta = np.arcsinh(ev["see_stateTrajGlbPz"][seed].astype(np.float64) / pt[seed]).astype(np.float32)
    assert np.allclose(e.field("in_eta"), exp_eta, rtol=1e-6, atol=1e-6)
    # superbin = 25*72*etabin + 25*phibin + dzbin (SDL/LST.cc:326-335)
    p = np.stack([ev["see_px"], ev["see_py"], ev["see_pz"]], 1)[seed].astype(np.float64)
    eta = np.arcsinh(p[:, 2] / np.hypot(p[:, 0], p[:, 1])).astype(np.float32).astype(np.float64)
    phi = np.arctan2(p[:, 1], p[:, 0]).astype(np.float32).astype(np.float64)
    etab = np.floor((eta + 2.6) / (5.2 / 25)).astype(int)
    phib = np.floor((phi + np.pi) / (2 * np.pi / 72)).astype(int)
    dzb = np.floor((ev["see_dz"][seed] + 30) / np.float32(2.4)).astype(int)
    sb = e.field("in_superbin")
    ok = np.abs(sb - (1800 * etab + 25 * phib + dzb)) == 0
    assert ok.mean() > 0.99  # bin edges may round differently in the last ulp
    # 3 or 4 pseudo-hits per pLS, detId 1, appended after the OT hits
    n_ot = len(ev["ph2_x"])
    quad = e.field("in_isQuad").astype(int)
    assert len(e.field("in_x")) == n_ot + (3 * len(seed) + quad.sum())
    assert np.all(e.field("in_detId")[n_ot:] == 1)
