"""Pins the oracle's later stages (T5 circle fit, pixel objects, track
candidates) to closed forms and to order-free restatements of their rules
(CPU only).

- The T5 regression is the weighted algebraic (Kasa) circle fit: on points
  that lie on a circle it returns that circle with chi2 ~ 0, and on noisy
  points it equals a float64 weighted least-squares solve.
- Every duplicate-removal rule of the reference is a predicate on pairs; the
  oracle's serial loops must give the same flags as the predicate evaluated
  over all pairs at once (the documented tie-breaks included).
- Track candidates come out in kernel order (pT5, pT3, T5, pLS), each block
  made of the non-duplicate objects of its kind in index order, and every
  output hit list is the candidate's hits with the repeated fourth pixel hit
  dropped.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import orc


@pytest.fixture(scope="module")
def L():
    orc.build()
    lib = orc.lib()
    fp = C.POINTER(C.c_float)
    lib.orc_regression_circle.restype = C.c_float
    lib.orc_regression_circle.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, fp, fp, fp]
    lib.orc_circle_chi2.restype = C.c_float
    lib.orc_circle_chi2.argtypes = [C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_float, C.c_float, C.c_float]
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _fit(L, x, y, d1, d2, sl, flat):
    g, f, chi2 = C.c_float(), C.c_float(), C.c_float()
    r = L.orc_regression_circle(len(x), _p(x), _p(y), _p(d1), _p(d2), _p(sl), _p(flat), C.byref(g), C.byref(f),
                                C.byref(chi2))
    return r, g.value, f.value, chi2.value


def _arc(rng, n=5):
    """Points on a well-conditioned arc (span >= 1.5 rad): the reference's
    float32 normal equations lose all precision on short arcs of large
    circles, where no float32 method can recover the circle."""
    R = rng.uniform(20.0, 80.0)
    phi0 = rng.uniform(-np.pi, np.pi)
    d = rng.uniform(0.2, 1.0) * R
    cx, cy = d * np.cos(phi0), d * np.sin(phi0)
    ang = rng.uniform(-np.pi, np.pi) + np.sort(rng.uniform(0.0, 1.0, n)) * rng.uniform(1.5, 3.0)
    return cx, cy, R, cx + R * np.cos(ang), cy + R * np.sin(ang)


def _unit_weights(x, y):
    """Uncertainties with sigma_i = sqrt(5) for every point, whatever the module
    orientation: d1 = d2 makes sigma = 2 d |p|, so the weights 1/sigma^2 sum to 1."""
    d = (np.sqrt(5.0) / (2.0 * np.hypot(x, y))).astype(np.float32)
    return d, d.copy()


def test_regression_recovers_exact_circle(L):
    """Five points on a circle, weights summing to 1: the fit is the weighted
    algebraic (Kasa) fit and returns that circle with chi2 ~ 0.

    The reference's normal equations (SDL/Quintuplet.h:1256-1305) centre the
    sums without dividing by the weight total, so they equal the Kasa fit only
    when the weights sum to 1; with unscaled weights the result moves away from
    the circle, which the last check pins."""
    rng = np.random.default_rng(11)
    err = []
    for _ in range(300):
        cx, cy, R, x, y = _arc(rng)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        d1, d2 = _unit_weights(x32.astype(np.float64), y32.astype(np.float64))
        sl = rng.uniform(-5, 5, 5).astype(np.float32)
        flat = rng.integers(0, 2, 5).astype(np.uint8)
        r, g, f, chi2 = _fit(L, x32, y32, d1, d2, sl, flat)
        err.append(max(abs(r - R), abs(g - cx), abs(f - cy)) / R)
    # float32 normal equations: ~1e-6 typical, up to a few % when the points bunch up
    err = np.array(err)
    assert np.median(err) < 1e-4 and np.percentile(err, 90) < 1e-3 and err.max() < 0.1
    # weights scaled by 0.37: no longer the circle
    cx, cy, R, x, y = _arc(np.random.default_rng(99))
    x32, y32 = x.astype(np.float32), y.astype(np.float32)
    d1, d2 = _unit_weights(x32.astype(np.float64), y32.astype(np.float64))
    scale = np.float32(1 / np.sqrt(0.37))
    r, *_ = _fit(L, x32, y32, d1 * scale, d2 * scale, np.zeros(5, np.float32), np.ones(5, np.uint8))
    assert abs(r - R) / R > 1e-2


def test_regression_equals_weighted_least_squares(L):
    """Noisy points, weights summing to 1: the fit equals a float64 weighted
    algebraic least-squares solve of x^2 + y^2 = 2g x + 2f y - c."""
    rng = np.random.default_rng(12)
    err, chi_err = [], []
    for _ in range(300):
        cx, cy, R, x, y = _arc(rng)
        x = x + rng.normal(0, 0.05, 5)
        y = y + rng.normal(0, 0.05, 5)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        X, Y = x32.astype(np.float64), y32.astype(np.float64)
        d1, d2 = _unit_weights(X, Y)
        sl = np.full(5, -999.0, np.float32)
        flat = np.ones(5, np.uint8)
        r, g, f, chi2 = _fit(L, x32, y32, d1, d2, sl, flat)
        sigma = 2 * np.sqrt((X * d1) ** 2 + (Y * d2) ** 2)
        w = 1.0 / sigma
        A = np.stack([X, Y, -np.ones(5)], 1) * w[:, None]
        b = (X * X + Y * Y) * w
        (tg, tf, c), *_ = np.linalg.lstsq(A, b, rcond=None)
        gg, ff = tg / 2, tf / 2
        rr = np.sqrt(gg * gg + ff * ff - c)
        err.append(max(abs(r - rr), abs(g - gg), abs(f - ff)) / rr)
        res = (X * X + Y * Y - tg * X - tf * Y + c) / sigma
        chi_err.append(abs(chi2 - float(np.sum(res * res))) / max(float(np.sum(res * res)), 1e-3))
    err, chi_err = np.array(err), np.array(chi_err)
    assert np.median(err) < 1e-4 and np.percentile(err, 90) < 1e-3 and err.max() < 0.1
    assert np.median(chi_err) < 1e-2


def test_circle_chi2_vanishes_on_circle(L):
    rng = np.random.default_rng(13)
    for _ in range(100):
        cx, cy, R, x, y = _arc(rng)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        d1 = np.ones(5, np.float32)
        d2 = np.ones(5, np.float32)
        sl = rng.uniform(-5, 5, 5).astype(np.float32)
        flat = rng.integers(0, 2, 5).astype(np.uint8)
        on = L.orc_circle_chi2(5, _p(x32), _p(y32), _p(d1), _p(d2), _p(sl), _p(flat), cx, cy, R)
        off = L.orc_circle_chi2(5, _p(x32), _p(y32), _p(d1), _p(d2), _p(sl), _p(flat), cx, cy, R + 1.0)
        assert on < 1e-3 * off


@pytest.fixture(scope="module")
def tc_event(oracle_geometry):
    from tracklooper_amd import synth
    ev = synth.generate_event("pu", 5, n_pu=50)
    return ev, orc.Event(oracle_geometry).run_lst(ev, "tc")


def _n_matched(a, b):
    """For each row pair: how many entries of a's row occur in b's row."""
    return (a[:, None, :, None] == b[None, :, None, :]).any(-1).sum(-1)


def test_pt5_duplicate_rule(tc_event):
    _, e = tc_event
    n = e.scalar("nPixelQuintuplets")
    assert n > 50
    hits = e.field("pt5.hitIndices")[:14 * n].reshape(n, 14)
    s = e.field("pt5.score")[:n]
    m = _n_matched(hits, hits)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    drop = (m >= 7) & (i != j) & ((s[:, None] > s[None, :]) | ((s[:, None] == s[None, :]) & (i > j)))
    assert np.array_equal(e.field("pt5.isDup")[:n].astype(bool), drop.any(1))


def test_pt3_duplicate_rule(tc_event):
    _, e = tc_event
    n = e.scalar("nPixelTriplets")
    assert n > 10
    hits = e.field("pt3.hitIndices")[:10 * n].reshape(n, 10).astype(np.int64)
    m = _n_matched(hits[:, :4], hits[:, :4]) + _n_matched(hits[:, 4:], hits[:, 4:])
    lay = e.field("pt3.logicalLayers")[:5 * n].reshape(n, 5)[:, 2].astype(int)
    s = e.field("pt3.score")[:n]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    li, lj = lay[:, None], lay[None, :]
    si, sj = s[:, None], s[None, :]
    drop = (m >= 5) & (i != j) & ((lj < li) | ((li == lj) & (si > sj)) | ((li == lj) & (si == sj) & (i < j)))
    # the final flags also include crossCleanpT3; rebuild those from the pT5 pLS directions
    n5 = e.scalar("nPixelQuintuplets")
    base = e.field("ranges.segmentModuleIndices")[-1]
    pj = e.field("pt5.pixelIndices")[:n5] - base
    eta, phi = e.field("segs.eta"), e.field("segs.phi")
    d_eta = np.abs(e.field("pt3.eta_pix")[:n][:, None] - eta[pj][None, :])
    d_phi = e.field("pt3.phi_pix")[:n][:, None] - phi[pj][None, :]
    d_phi = np.where(d_phi > np.float32(np.pi), d_phi - np.float32(2 * np.pi),
                     np.where(d_phi < -np.float32(np.pi), d_phi + np.float32(2 * np.pi), d_phi))
    cross = ((d_eta * d_eta + d_phi * d_phi) < np.float32(1e-5)).any(1) if n5 else np.zeros(n, bool)
    assert np.array_equal(e.field("pt3.isDup")[:n].astype(bool), drop.any(1) | cross)


def test_track_candidate_blocks_and_outputs(tc_event):
    ev, e = tc_event
    S = e.scalar
    ntc = S("nTrackCandidates")
    typ = e.field("tc.trackCandidateType")[:ntc]
    n7, n5, n4, n8 = S("nTrackCandidatespT5"), S("nTrackCandidatespT3"), S("nTrackCandidatesT5"), S("nTrackCandidatespLS")
    assert n7 + n5 + n4 + n8 == ntc and min(n7, n4, n8) > 0
    assert np.array_equal(typ, np.repeat(np.array([7, 5, 4, 8], np.int16), [n7, n5, n4, n8]))
    direct = e.field("tc.directObjectIndices")[:ntc]
    # pT5 block: the non-duplicate pT5s in index order
    npt5 = S("nPixelQuintuplets")
    assert np.array_equal(direct[:n7], np.nonzero(e.field("pt5.isDup")[:npt5] == 0)[0])
    # pT3 block
    npt3 = S("nPixelTriplets")
    assert np.array_equal(direct[n7:n7 + n5], np.nonzero(e.field("pt3.isDup")[:npt3] == 0)[0])
    # T5 block: tight, not duplicated, not in a pT5, in slot order, at most 1000
    t5 = direct[n7 + n5:n7 + n5 + n4]
    assert np.all(np.diff(t5.astype(np.int64)) > 0) and n4 <= 1000
    assert np.all(e.field("t5.TightCutFlag")[t5] == 1) and np.all(e.field("t5.isDup")[t5] == 0)
    assert np.all(e.field("t5.partOfPT5")[t5] == 0)
    # pLS block: quadruplet seeds never flagged, in index order
    pls = direct[n7 + n5 + n4:]
    assert np.all(np.diff(pls.astype(np.int64)) > 0)
    assert np.all(e.field("segs.isQuad")[pls] == 1) and np.all(e.field("segs.isDup")[pls] == 0)
    # output lists
    ln = e.field("out.len")[:ntc]
    expect = {7: (13, 14), 5: (9, 10), 4: (10,), 8: (3, 4)}
    for t, l in zip(typ, ln):
        assert l in expect[int(t)]
    flat = e.field("out.hitIdxs")
    n_in = len(ev["ph2_x"])
    tcode = typ != 8
    off = e.field("out.hitOffsets")
    for k in np.nonzero(tcode)[0][:200]:
        assert np.all(flat[off[k]:off[k + 1]] < n_in + 4 * len(ev["see_px"]))
    seeds = e.field("out.seedIdx")[:ntc]
    assert np.all(seeds[typ == 4] == -1) and np.all(seeds[typ != 4] >= 0)


def test_pls_first_pass_rule(oracle_geometry):
    """First checkHitspLS pass: a pLS is flagged iff a partner within 0.1 in eta,
    sharing >= 3 of the lower index's hits, makes it the removal target."""
    from tracklooper_amd import synth
    ev = synth.generate_event("pu", 6, n_pu=50)
    e = orc.Event(oracle_geometry).run_lst(ev, "plsclean")
    n = e.scalar("nPLS")
    assert n > 100
    eta = e.field("segs.eta")[:n]
    q = e.field("segs.isQuad")[:n].astype(int)
    s = e.field("segs.score")[:n]
    h = e.field("segs.pLSHitsIdxs")[:4 * n].reshape(n, 4)
    lo, hi = np.triu_indices(n, 1)
    close = np.abs(eta[hi] - eta[lo]) <= np.float32(0.1)
    lo, hi = lo[close], hi[close]
    m = (h[lo][:, :, None] == h[hi][:, None, :]).any(-1).sum(-1)
    qd = q[lo] - q[hi]
    sd = s[lo] - s[hi]
    rm = np.where(qd > 0, hi, np.where(qd < 0, lo, np.where(sd < 0, hi, lo)))
    flag = np.zeros(n, bool)
    flag[rm[m >= 3]] = True
    assert np.array_equal(e.field("segs.isDup")[:n] & 1 == 1, flag)
