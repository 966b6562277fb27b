"""Seeded synthetic events for the Line Segment Tracking chain.

Produces exactly the arguments of ``SDL::LST::run`` (SDL/LST.h:29-48): outer
tracker hits ``ph2_{detId,x,y,z}`` and pixel seeds ``see_*``, plus truth
tables (hit -> simulated track) used by the physics metrics.

Model (the reference ships geometry only, no sensor outlines):
  * each OT module is a planar rectangle through its sensor centroid
    (``sensor_centroids.bin``); the plane normal is the vector from the lower
    to the upper sensor centroid of the module pair (flat barrel: radial,
    tilted barrel: tilted in r-z, endcap: along z);
  * helices in a uniform 3.8 T field along +z; a helix leaves one hit on the
    lower and one on the upper sensor of every module whose outline it crosses
    on its outgoing half turn;
  * the precise coordinate (r-phi) is smeared by 20 um; the coarse one is
    quantised to the strip-row centre (PS strip 2.5 cm, 2S strip 5 cm; PS
    macro-pixel 1.5 mm), as a strip hit is reported at its strip centre;
  * pixel seeds are built from the first 3 or 4 crossings of an inner pixel
    tracker model (barrel radii 3.0/6.8/10.9/16.0 cm, disks at
    |z| = 25..250 cm with 3 < r < 16.5 cm): momentum and impact parameters at
    the point of closest approach, and state at the last seed hit.
Everything is driven by a numpy Generator keyed by (config id, event index).
"""
from __future__ import annotations

import os
import numpy as np

B_FIELD = 3.8
K_CURV = 2.99792458e-3 * B_FIELD  # 1/R [1/cm] per 1/pt [1/GeV]
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

PIX_BARREL_R = np.array([3.0, 6.8, 10.9, 16.0])
PIX_BARREL_ZMAX = 27.0
PIX_DISK_Z = np.array([25.0, 31.6, 40.0, 50.0, 63.0, 79.0, 100.0, 125.0, 155.0, 190.0, 230.0, 250.0])
PIX_DISK_RMIN, PIX_DISK_RMAX = 3.0, 16.5


def geometry_dir(name: str = "OT800_IT615_pt0.8") -> str:
    return os.path.join(DATA_DIR, name)


class SynthGeometry:
    """Planar module model derived from the sensor centroid table."""

    def __init__(self, geom_dir: str):
        from scipy.spatial import cKDTree

        rec = np.fromfile(os.path.join(geom_dir, "sensor_centroids.bin"),
                          dtype=np.dtype([("detid", "<u4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("t", "<i4")]))
        det = rec["detid"].astype(np.int64)
        subdet = (det >> 25) & 7
        side = np.where(subdet == 4, (det >> 23) & 3, (det >> 18) & 3)
        layer = np.where(subdet == 4, (det >> 18) & 7, (det >> 20) & 7)
        module = (det >> 2) & 127
        # lower/upper assignment and partner detId: SDL/Module.h:196-227
        inv = np.zeros(len(det), bool)
        ec = subdet == 4
        inv[ec & (side == 1)] = module[ec & (side == 1)] % 2 == 1
        inv[ec & (side == 2)] = module[ec & (side == 2)] % 2 == 0
        br = subdet == 5
        c = br & (side == 3)
        inv[c & (layer <= 3)] = module[c & (layer <= 3)] % 2 == 1
        inv[c & (layer >= 4)] = module[c & (layer >= 4)] % 2 == 0
        t = br & ((side == 1) | (side == 2))
        inv[t & (layer <= 2)] = module[t & (layer <= 2)] % 2 == 1
        inv[t & (layer == 3)] = module[t & (layer == 3)] % 2 == 0
        is_lower = np.where(inv, (det & 1) == 0, (det & 1) == 1)
        partner = np.where(is_lower, np.where(inv, det - 1, det + 1), np.where(inv, det + 1, det - 1))
        order = np.argsort(det)
        pidx = order[np.searchsorted(det[order], partner)]
        assert np.all(det[pidx] == partner)

        self.detid = det.astype(np.uint32)
        self.cent = np.stack([rec["x"], rec["y"], rec["z"]], 1).astype(np.float64)
        self.mtype = rec["t"]
        self.subdet, self.side, self.layer = subdet, side, layer
        lo = np.nonzero(is_lower)[0]
        up = pidx[lo]
        self.lo, self.up = lo, up
        n = self.cent[up] - self.cent[lo]
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        phic = np.arctan2(self.cent[lo, 1], self.cent[lo, 0])
        tang = np.stack([-np.sin(phic), np.cos(phic), np.zeros_like(phic)], 1)
        radial = np.stack([np.cos(phic), np.sin(phic), np.zeros_like(phic)], 1)
        is_ec = subdet[lo] == 4
        u = np.where(is_ec[:, None], radial, np.cross(n, tang))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        v = np.cross(n, u)
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        self.n, self.u, self.v = n, u, v
        is2s = self.mtype[lo] == 25
        # half-sizes along the coarse (u) and precise (v) axes [cm]
        self.hu = np.where(is2s, 5.0, 2.5)
        self.hv = np.where(is2s, 4.6, 4.8)
        # surfaces: barrel layers and endcap disks
        self.groups = []
        for L in range(1, 7):
            m = np.nonzero((subdet[lo] == 5) & (layer[lo] == L))[0]
            if len(m):
                r = np.hypot(self.cent[lo[m], 0], self.cent[lo[m], 1])
                self.groups.append(("B", float(np.median(r)), m, cKDTree(self.cent[lo[m]])))
        for s in (1, 2):
            for L in range(1, 6):
                m = np.nonzero((subdet[lo] == 4) & (side[lo] == s) & (layer[lo] == L))[0]
                if len(m):
                    self.groups.append(("E", float(np.median(self.cent[lo[m], 2])), m, cKDTree(self.cent[lo[m]])))


class Helices:
    """Vectorised helices: charge q, pt, eta, phi0 at vertex (0, 0, vz)."""

    def __init__(self, q, pt, eta, phi0, vz):
        self.q = q.astype(np.float64)
        self.pt = pt.astype(np.float64)
        self.eta = eta.astype(np.float64)
        self.phi0 = phi0.astype(np.float64)
        self.vz = vz.astype(np.float64)
        self.R = self.pt / K_CURV
        self.sh = np.sinh(self.eta)
        self.cx = self.q * self.R * np.sin(self.phi0)
        self.cy = -self.q * self.R * np.cos(self.phi0)

    def pos(self, a, i=slice(None)):
        ph = self.phi0[i] - self.q[i] * a
        return np.stack([self.cx[i] - self.q[i] * self.R[i] * np.sin(ph),
                         self.cy[i] + self.q[i] * self.R[i] * np.cos(ph),
                         self.vz[i] + self.R[i] * a * self.sh[i]], -1)

    def dpos(self, a, i=slice(None)):
        ph = self.phi0[i] - self.q[i] * a
        return np.stack([self.R[i] * np.cos(ph), self.R[i] * np.sin(ph), self.R[i] * self.sh[i]], -1)

    def mom(self, a, i=slice(None)):
        ph = self.phi0[i] - self.q[i] * a
        return np.stack([self.pt[i] * np.cos(ph), self.pt[i] * np.sin(ph), self.pt[i] * self.sh[i]], -1)

    def alpha_cyl(self, r):
        """Turning angle where the transverse radius first reaches r (nan if never)."""
        x = r / (2.0 * self.R)
        with np.errstate(invalid="ignore"):
            return np.where(x < 1.0, 2.0 * np.arcsin(np.minimum(x, 1.0)), np.nan)

    def alpha_z(self, z):
        with np.errstate(divide="ignore", invalid="ignore"):
            a = (z - self.vz) / (self.R * self.sh)
        return np.where(a > 0, a, np.nan)


def _plane_alpha(h: Helices, idx, a0, c, n, iters=6):
    a = a0.copy()
    for _ in range(iters):
        f = np.einsum("ij,ij->i", h.pos(a, idx) - c, n)
        d = np.einsum("ij,ij->i", h.dpos(a, idx), n)
        with np.errstate(divide="ignore", invalid="ignore"):
            a = a - f / d
    return a


def _ot_hits(geo: SynthGeometry, h: Helices, rng, exact: bool = False):
    """Return hit arrays (sensor row, position, track id); exact: the helix crossing points themselves
    (no smearing, no strip-centre quantisation)."""
    rows, pos, trk = [], [], []
    ntrk = len(h.pt)
    for kind, ref, members, tree in geo.groups:
        a0 = h.alpha_cyl(ref) if kind == "B" else h.alpha_z(ref)
        ok = np.isfinite(a0) & (a0 <= np.pi)
        ti = np.nonzero(ok)[0]
        if len(ti) == 0:
            continue
        p0 = h.pos(a0[ti], ti)
        if kind == "E":
            rt = np.hypot(p0[:, 0], p0[:, 1])
            keep = (rt > 20.0) & (rt < 115.0)
            ti, p0 = ti[keep], p0[keep]
            if len(ti) == 0:
                continue
        k = min(10, len(members))
        _, nb = tree.query(p0, k=k)
        nb = nb.reshape(len(ti), k)
        best = np.full(len(ti), np.inf)
        best_m = np.full(len(ti), -1)
        best_a = np.zeros(len(ti))
        for j in range(k):
            m = members[nb[:, j]]
            c = geo.cent[geo.lo[m]]
            a = _plane_alpha(h, ti, a0[ti], c, geo.n[m])
            p = h.pos(a, ti)
            d = p - c
            du = np.abs(np.einsum("ij,ij->i", d, geo.u[m])) / geo.hu[m]
            dv = np.abs(np.einsum("ij,ij->i", d, geo.v[m])) / geo.hv[m]
            score = np.maximum(du, dv)
            good = np.isfinite(score) & (a > 0) & (a <= np.pi) & (score < best)
            best[good], best_m[good], best_a[good] = score[good], m[good], a[good]
        sel = best <= 1.0
        ti, m, a = ti[sel], best_m[sel], best_a[sel]
        for sensor_rows in (geo.lo[m], geo.up[m]):
            c = geo.cent[sensor_rows]
            aa = _plane_alpha(h, ti, a, c, geo.n[m])
            p = h.pos(aa, ti)
            d = p - c
            du = np.einsum("ij,ij->i", d, geo.u[m])
            dv = np.einsum("ij,ij->i", d, geo.v[m])
            t = geo.mtype[sensor_rows]
            # coarse coordinate: strip-row centre (strips) or 1.5 mm macro-pixels
            strip_len = np.where(t == 25, 5.0, 2.5)
            du_q = np.where(t == 23, np.round(du / 0.15) * 0.15, np.sign(du) * strip_len / 2.0)
            dv_s = dv + rng.normal(0.0, 20e-4, len(dv))
            if exact:
                du_q, dv_s = du, dv
            p = c + du_q[:, None] * geo.u[m] + dv_s[:, None] * geo.v[m]
            rows.append(sensor_rows)
            pos.append(p)
            trk.append(ti)
    if not rows:
        return np.zeros(0, np.int64), np.zeros((0, 3)), np.zeros(0, np.int64)
    return np.concatenate(rows), np.concatenate(pos), np.concatenate(trk)


def _noise_hits(geo: SynthGeometry, n, rng):
    m = rng.integers(0, len(geo.lo), n)
    which = rng.integers(0, 2, n)
    rows = np.where(which == 0, geo.lo[m], geo.up[m])
    du = rng.uniform(-1, 1, n) * geo.hu[m]
    dv = rng.uniform(-1, 1, n) * geo.hv[m]
    t = geo.mtype[rows]
    strip_len = np.where(t == 25, 5.0, 2.5)
    du = np.where(t == 23, du, np.sign(du) * strip_len / 2.0)
    p = geo.cent[rows] + du[:, None] * geo.u[m] + dv[:, None] * geo.v[m]
    return rows, p


def _pixel_crossings(h: Helices):
    """Turning angles of the inner-pixel crossings, sorted, per track (nan padded)."""
    cands = []
    for r in PIX_BARREL_R:
        a = h.alpha_cyl(r)
        with np.errstate(invalid="ignore"):
            z = h.vz + h.R * a * h.sh
            cands.append(np.where(np.abs(z) < PIX_BARREL_ZMAX, a, np.nan))
    for zd in PIX_DISK_Z:
        for s in (1.0, -1.0):
            a = h.alpha_z(s * zd)
            with np.errstate(invalid="ignore"):
                rt = 2.0 * h.R * np.abs(np.sin(a / 2.0))
                cands.append(np.where((rt > PIX_DISK_RMIN) & (rt < PIX_DISK_RMAX) & (a <= np.pi), a, np.nan))
    A = np.stack(cands, 1)
    A = np.where(np.isfinite(A), A, np.inf)
    A.sort(axis=1)
    return A


def generate_event(kind: str, index: int, *, n_pu: int | None = None, config_id: int | None = None,
                   geom: SynthGeometry | None = None, geom_name: str = "OT800_IT615_pt0.8",
                   max_module_hits: int = 127, exact: bool = False) -> dict:
    """Generate one event.

    kind: "muon" (10 muons, pT log-U[0.5,50] GeV) or "pu" (n_pu x 35 charged
    helices with pT = 0.4 + Exp(0.5) GeV plus 200 hard log-U[1,100] GeV, 1%
    noise hits).  Returns a dict with the ``LST::run`` arguments and truth.
    max_module_hits: hits kept per sensor (127 keeps the input valid for the
    reference's int8 counts; a larger value makes inputs the engine rejects).
    exact: hits at the exact helix crossings and unsmeared seeds (for checks
    that hold on perfect helices).
    """
    geo = geom or _geometry_cache(geom_name)
    cid = config_id if config_id is not None else {"muon": 1, "pu": 2}[kind]
    rng = np.random.default_rng([cid, index, 0x15A])
    if kind == "muon":
        nt = 10
        pt = np.exp(rng.uniform(np.log(0.5), np.log(50.0), nt))
        vz = rng.normal(0.0, 5.0, nt)
        noise_frac = 0.0
        hard = np.ones(nt, bool)
    elif kind == "pu":
        npu = 200 if n_pu is None else int(n_pu)
        nsoft = 35 * npu
        pt = np.concatenate([0.4 + rng.exponential(0.5, nsoft), np.exp(rng.uniform(0.0, np.log(100.0), 200))])
        nt = len(pt)
        vtx = rng.normal(0.0, 5.0, npu + 1)
        vz = np.concatenate([vtx[rng.integers(0, npu, nsoft)], np.full(200, vtx[npu])])
        noise_frac = 0.01
        hard = np.concatenate([np.zeros(nsoft, bool), np.ones(200, bool)])  # the 200 tracks of the signal vertex
    else:
        raise ValueError(kind)
    eta = rng.uniform(-2.4, 2.4, nt)
    phi0 = rng.uniform(-np.pi, np.pi, nt)
    q = rng.choice(np.array([-1.0, 1.0]), nt)
    h = Helices(q, pt, eta, phi0, vz)

    rows, pos, trk = _ot_hits(geo, h, rng, exact)
    if noise_frac > 0:
        nrow, npos = _noise_hits(geo, int(round(noise_frac * len(rows))), rng)
        rows = np.concatenate([rows, nrow])
        pos = np.concatenate([pos, npos])
        trk = np.concatenate([trk, np.full(len(nrow), -1)])
    det = geo.detid[rows]
    order = np.lexsort((np.arange(len(det)), det))  # hits grouped by module, stable
    det, pos, trk = det[order], pos[order], trk[order]
    # at most 127 hits per sensor (hit counts are int8 in SDL/Hit.h:31-32)
    _, first, counts = np.unique(det, return_index=True, return_counts=True)
    rank = np.arange(len(det)) - np.repeat(first, counts)
    keep = rank < max_module_hits
    det, pos, trk = det[keep], pos[keep], trk[keep]

    # pixel seeds
    A = _pixel_crossings(h)
    nx = np.isfinite(A).sum(1)
    seed_p = 0.8 if kind == "pu" else 1.0
    make = (nx >= 3) & (rng.uniform(size=nt) < seed_p) & ((pt > 0.6) if kind == "pu" else True)
    seeds = []
    pix_id = 0
    for i in np.nonzero(make)[0]:
        nh = 4 if nx[i] >= 4 else 3
        copies = [nh]
        if kind == "pu" and nh == 4 and rng.uniform() < 0.25:
            copies.append(3)  # a triplet seed sharing the first three pixel hits
        ids = list(range(pix_id, pix_id + nh))
        pix_id += nh
        for nsh in copies:
            a_last = A[i, nsh - 1]
            r3 = h.pos(np.array([a_last]), [i])[0]
            p3 = h.mom(np.array([a_last]), [i])[0]
            sm = 1.0 + (0.0 if exact else 1.0) * rng.normal(0.0, 0.01)
            seeds.append(dict(
                px=h.pt[i] * np.cos(h.phi0[i]) * sm, py=h.pt[i] * np.sin(h.phi0[i]) * sm, pz=h.pt[i] * h.sh[i] * sm,
                dxy=rng.normal(0.0, 0.002) * (1.0 - exact), dz=h.vz[i] + rng.normal(0.0, 0.005) * (1.0 - exact),
                ptErr=0.015 * h.pt[i] + 0.005, etaErr=0.002,
                gx=r3[0], gy=r3[1], gz=r3[2], gpx=p3[0] * sm, gpy=p3[1] * sm, gpz=p3[2] * sm,
                q=int(h.q[i]), hits=ids[:nsh], trk=int(i)))
    ns = len(seeds)
    f32 = lambda k: np.array([s[k] for s in seeds], np.float32).reshape(ns)
    hit_off = np.zeros(ns + 1, np.uint32)
    hit_off[1:] = np.cumsum([len(s["hits"]) for s in seeds])
    return dict(
        ph2_detId=det.astype(np.uint32), ph2_x=pos[:, 0].astype(np.float32), ph2_y=pos[:, 1].astype(np.float32),
        ph2_z=pos[:, 2].astype(np.float32),
        see_px=f32("px"), see_py=f32("py"), see_pz=f32("pz"), see_dxy=f32("dxy"), see_dz=f32("dz"),
        see_ptErr=f32("ptErr"), see_etaErr=f32("etaErr"), see_stateTrajGlbX=f32("gx"),
        see_stateTrajGlbY=f32("gy"), see_stateTrajGlbZ=f32("gz"), see_stateTrajGlbPx=f32("gpx"),
        see_stateTrajGlbPy=f32("gpy"), see_stateTrajGlbPz=f32("gpz"),
        see_q=np.array([s["q"] for s in seeds], np.int32).reshape(ns),
        see_hitIdx=np.array([x for s in seeds for x in s["hits"]], np.int32),
        see_hitIdx_off=hit_off,
        truth_hit_trk=trk.astype(np.int64), truth_seed_trk=np.array([s["trk"] for s in seeds], np.int64),
        truth_pix_trk=np.array([trk for trk in _pixel_hit_tracks(seeds, pix_id)], np.int64),
        truth_pt=h.pt, truth_eta=h.eta, truth_phi=h.phi0, truth_vz=h.vz, truth_q=h.q, truth_is_hard=hard,
    )


def _pixel_hit_tracks(seeds, n_pix):
    """Track of every pixel hit (pixel hits are created per seeded helix)."""
    trk = np.full(n_pix, -1, np.int64)
    for s in seeds:
        trk[s["hits"]] = s["trk"]
    return trk


_GEO = {}


def _geometry_cache(name):
    if name not in _GEO:
        _GEO[name] = SynthGeometry(geometry_dir(name))
    return _GEO[name]


# ---------------------------------------------------------------- uneven batches
C3_CONFIG = 3  # config id of the uneven-pileup batches


def c3_pileups(batch: int, n: int = 64) -> np.ndarray:
    """Pileup of each event of uneven batch `batch`: U{50..250}, keyed by (config 3, batch)."""
    return np.random.default_rng([C3_CONFIG, batch, 0xC3]).integers(50, 251, n)


def _gen_one(spec):
    kind, index, n_pu, cid, geom_name = spec
    return generate_event(kind, index, n_pu=n_pu, config_id=cid, geom_name=geom_name)


def generate_many(specs, workers: int = 1) -> list:
    """generate_event over (kind, index, n_pu, config_id, geom_name) tuples, in order.  workers > 1 runs a
    process pool: forked while the process has not used the GPU, spawned (fresh interpreters that never
    touch it) once it has."""
    if workers <= 1 or len(specs) < 2:
        return [_gen_one(s) for s in specs]
    import multiprocessing as mp
    import sys
    torch, pkg = sys.modules.get("torch"), sys.modules.get("tracklooper_amd")
    gpu_used = (torch is not None and torch.cuda.is_initialized()) or getattr(pkg, "_lib", None) is not None
    pool = mp.get_context("spawn" if gpu_used else "fork").Pool(min(workers, len(specs)))
    try:
        return pool.map(_gen_one, specs, chunksize=1)
    finally:  # the workers exit on their own: no SIGTERM (which a profiler's signal handler would take)
        pool.close()
        pool.join()


def c3_batch(batch: int, n: int = 64, geom_name: str = "OT800_IT615_pt0.8", workers: int = 1) -> list:
    """The n events of uneven batch `batch`: pileup U{50..250} each (35 charged helices per
    interaction, the PU spectra of generate_event), seeded by (config 3, batch, position)."""
    pus = c3_pileups(batch, n)
    return generate_many([("pu", batch * n + i, int(pu), C3_CONFIG, geom_name) for i, pu in enumerate(pus)], workers)
