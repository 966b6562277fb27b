"""MI355X-native Line Segment Tracking engine.

Python mirror of the reference host interface for the LST path
(``SDL::loadAndFillES*``, ``SDL::Event``, ``SDL::LST::run``), calling the HIP
engine in ``_lib/liblst_amd.so`` through its C ABI (``include/lst_amd.h``).
There is no CPU fallback: importing works anywhere, but creating an event
setup requires the built library and a GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import synth  # noqa: F401  (seeded synthetic events)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "liblst_amd.so")
DEFAULT_GEOMETRY = os.path.join(HERE, "data", "OT800_IT615_pt0.8")

_DT = {0: np.float32, 1: np.uint32, 2: np.int32, 3: np.uint16, 4: np.int16, 5: np.uint8, 6: np.int8}
STAGES = {"hits": 0, "pls": 1, "md": 2, "ls": 3, "t3": 4, "t5": 5, "plsclean": 6, "pt5": 7, "pt3": 8, "tc": 9}
LST_T5_PLAIN_CHI2, LST_T5_DNN = 0, 1
_lib = None


class LSTError(RuntimeError):
    pass


class _PixelSegments(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "hit_indices0", "hit_indices1", "hit_indices2", "hit_indices3", "delta_phi_change", "pt_in", "pt_err",
        "px", "py", "pz", "eta", "eta_err", "phi", "charge", "seed_idx", "superbin", "pixel_type", "is_quad")] + [
        ("n_pls", C.c_uint32)]


class _Input(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "see_px", "see_py", "see_pz", "see_dxy", "see_dz", "see_ptErr", "see_etaErr", "see_stateTrajGlbX",
        "see_stateTrajGlbY", "see_stateTrajGlbZ", "see_stateTrajGlbPx", "see_stateTrajGlbPy", "see_stateTrajGlbPz",
        "see_q", "see_hitIdx", "see_hitIdx_offsets")] + [("n_seeds", C.c_uint32)] + [
        (n, C.c_void_p) for n in ("ph2_detId", "ph2_x", "ph2_y", "ph2_z")] + [("n_hits", C.c_uint32)]


def library_path() -> str:
    """The in-tree build, or the file named by $LST_LIBRARY (an alternative build of the same sources)."""
    return os.environ.get("LST_LIBRARY", LIB_PATH)


def lib():
    """Load liblst_amd.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        path = library_path()
        if not os.path.exists(path):
            raise LSTError(f"{path} is missing: run `python -m tracklooper_amd.build` (HIP, gfx950)")
        # PyTorch ships its own HIP runtime.  Loading it first lets the library
        # bind to that same runtime, so torch tensors, streams and events and the
        # library's device memory belong to one runtime in the process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(path)
        P = C.c_void_p
        L.lst_last_error.restype = C.c_char_p
        L.lst_es_create.argtypes = [C.c_char_p, C.c_float, C.c_uint32, C.POINTER(P)]
        L.lst_es_destroy.argtypes = [P]
        L.lst_es_scalar.argtypes = [P, C.c_char_p]
        L.lst_es_scalar.restype = C.c_uint32
        L.lst_event_create.argtypes = [P, P, C.c_int, C.POINTER(P)]
        L.lst_event_destroy.argtypes = [P]
        L.lst_event_add_hits.argtypes = [P, P, P, P, P, P, C.c_uint32, C.c_int]
        L.lst_event_add_pixel_segments.argtypes = [P, C.POINTER(_PixelSegments), C.c_int]
        for n in ("create_mini_doublets", "create_segments", "create_triplets", "create_quintuplets",
                  "pixel_line_segment_cleaning", "create_pixel_quintuplets", "create_pixel_triplets",
                  "create_track_candidates", "reset", "synchronize"):
            getattr(L, "lst_event_" + n).argtypes = [P]
        L.lst_event_count.argtypes = [P, C.c_char_p]
        L.lst_event_count.restype = C.c_uint32
        L.lst_event_get_field.argtypes = [P, C.c_char_p, P, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]
        L.lst_run.argtypes = [P, C.POINTER(_Input), C.c_int]
        L.lst_event_get_output.argtypes = [P, C.POINTER(C.c_uint32), P, P, P, P, C.c_uint32]
        _lib = L
    return _lib


def _check(status):
    if status & 0xFF:
        raise LSTError(f"lst status {status}: {lib().lst_last_error().decode()}")
    return status


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# C types of the lst_input arrays (include/lst_amd.h)
_INPUT_TYPES = {"see_q": np.int32, "see_hitIdx": np.int32, "see_hitIdx_off": np.uint32, "ph2_detId": np.uint32}


def _make_input(ev: dict):
    """lst_input over an event dict's see_* / ph2_* arrays (kept alive by the caller); every array is
    converted to the C type the struct names (float32 unless listed in _INPUT_TYPES)."""
    a = {k: np.ascontiguousarray(v, dtype=_INPUT_TYPES.get(k, np.float32))
         for k, v in ev.items() if k.startswith(("see_", "ph2_"))}
    n_seeds = len(a["see_px"])
    if len(a["see_hitIdx_off"]) != n_seeds + 1:
        raise LSTError(f"see_hitIdx_off has {len(a['see_hitIdx_off'])} entries, {n_seeds + 1} expected")
    if n_seeds and (a["see_hitIdx_off"][-1] > len(a["see_hitIdx"]) or np.any(np.diff(a["see_hitIdx_off"].astype(np.int64)) < 0)):
        raise LSTError("see_hitIdx_off is not a non-decreasing list of offsets into see_hitIdx")
    inp = _Input(
        *[_p(a[k]) for k in ("see_px", "see_py", "see_pz", "see_dxy", "see_dz", "see_ptErr", "see_etaErr",
                             "see_stateTrajGlbX", "see_stateTrajGlbY", "see_stateTrajGlbZ",
                             "see_stateTrajGlbPx", "see_stateTrajGlbPy", "see_stateTrajGlbPz", "see_q",
                             "see_hitIdx", "see_hitIdx_off")],
        len(a["see_px"]),
        *[_p(a[k]) for k in ("ph2_detId", "ph2_x", "ph2_y", "ph2_z")],
        len(a["ph2_x"]))
    return inp, list(a.values())


class LSTESData:
    """Geometry resident in HBM (``SDL::loadAndFillESDevice``, SDL/LSTESData.cc:87-113)."""

    def __init__(self, geom_dir: str = DEFAULT_GEOMETRY, pt_cut: float = 0.8, module_map_stride: int = 0):
        self.h = C.c_void_p()
        _check(lib().lst_es_create(geom_dir.encode(), pt_cut, module_map_stride, C.byref(self.h)))
        self.geom_dir = geom_dir
        self.pt_cut = pt_cut

    def scalar(self, name: str) -> int:
        return lib().lst_es_scalar(self.h, name.encode())

    def close(self):
        if self.h:
            lib().lst_es_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """``SDL::Event`` (SDL/Event.h:26-205) backed by the HIP engine."""

    def __init__(self, es: LSTESData, stream: int | None = None, t5_mode: int = LST_T5_PLAIN_CHI2):
        self.es = es
        self.h = C.c_void_p()
        _check(lib().lst_event_create(es.h, stream, t5_mode, C.byref(self.h)))
        self._keep = []

    # --- SDL::Event stage API -------------------------------------------------
    def addHitToEvent(self, x, y, z, detId, idxInNtuple):
        arrs = [np.ascontiguousarray(a, t) for a, t in ((x, np.float32), (y, np.float32), (z, np.float32),
                                                        (detId, np.uint32), (idxInNtuple, np.uint32))]
        self._keep = arrs
        _check(lib().lst_event_add_hits(self.h, *[_p(a) for a in arrs], len(arrs[0]), 0))

    def addPixelSegmentToEvent(self, hitIndices0, hitIndices1, hitIndices2, hitIndices3, dPhiChange, ptIn, ptErr,
                               px, py, pz, eta, etaErr, phi, charge, seedIdx, superbin, pixelType, isQuad):
        types = [np.uint32] * 4 + [np.float32] * 9 + [np.int32, np.uint32, np.int32, np.int8, np.int8]
        vals = [hitIndices0, hitIndices1, hitIndices2, hitIndices3, dPhiChange, ptIn, ptErr, px, py, pz, eta, etaErr,
                phi, charge, seedIdx, superbin, pixelType, isQuad]
        arrs = [np.ascontiguousarray(a, t) for a, t in zip(vals, types)]
        self._keep = arrs
        ps = _PixelSegments(*[_p(a) for a in arrs], len(arrs[5]))
        return _check(lib().lst_event_add_pixel_segments(self.h, C.byref(ps), 0))

    def createMiniDoublets(self):
        _check(lib().lst_event_create_mini_doublets(self.h))

    def createSegmentsWithModuleMap(self):
        _check(lib().lst_event_create_segments(self.h))

    def createTriplets(self):
        _check(lib().lst_event_create_triplets(self.h))

    def createQuintuplets(self):
        _check(lib().lst_event_create_quintuplets(self.h))

    def pixelLineSegmentCleaning(self):
        _check(lib().lst_event_pixel_line_segment_cleaning(self.h))

    def createPixelQuintuplets(self):
        _check(lib().lst_event_create_pixel_quintuplets(self.h))

    def createPixelTriplets(self):
        _check(lib().lst_event_create_pixel_triplets(self.h))

    def createTrackCandidates(self):
        _check(lib().lst_event_create_track_candidates(self.h))

    def getOutput(self):
        """LST::getOutput (SDL/LST.cc:366-391): per track candidate its ntuple hit
        indices, their number, the pixel seed index and the candidate type."""
        n = C.c_uint32(0)
        _check(lib().lst_event_get_output(self.h, C.byref(n), None, None, None, None, 0))
        k = n.value
        hits = np.zeros((k, 14), np.uint32)
        length = np.zeros(k, np.uint32)
        seed = np.zeros(k, np.int32)
        typ = np.zeros(k, np.int16)
        if k:
            _check(lib().lst_event_get_output(self.h, C.byref(n), _p(hits), _p(length), _p(seed), _p(typ), k))
        return {"tc_hitIdxs": [list(hits[i, :length[i]]) for i in range(k)], "tc_len": length,
                "tc_seedIdx": seed, "tc_trackCandidateType": typ}

    def record_t5dnn(self, capacity: int = 1 << 19):
        """Record the T5 DNN's inputs for every candidate at its call site (DNN mode)."""
        lib().lst_event_record_t5dnn.argtypes = [C.c_void_p, C.c_uint32]
        _check(lib().lst_event_record_t5dnn(self.h, capacity))
        return self

    def t5dnn_record(self) -> dict:
        """The recorded rows: features [n, 38], score [n], t3_pair [n, 2], hit_idxs [n, 10]."""
        L = lib()
        L.lst_event_get_t5dnn_record.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)] + [C.c_void_p] * 4 + [C.c_uint32]
        n = C.c_uint32(0)
        _check(L.lst_event_get_t5dnn_record(self.h, C.byref(n), None, None, None, None, 0))
        k = n.value
        f = np.zeros((k, 38), np.float32)
        sc = np.zeros(k, np.float32)
        pr = np.zeros((k, 2), np.uint32)
        hi = np.zeros((k, 10), np.uint32)
        if k:
            _check(L.lst_event_get_t5dnn_record(self.h, C.byref(n), _p(f), _p(sc), _p(pr), _p(hi), k))
        return {"features": f, "score": sc, "t3_pair": pr, "hit_idxs": hi}

    def resetEvent(self):
        _check(lib().lst_event_reset(self.h))

    def synchronize(self):
        _check(lib().lst_event_synchronize(self.h))

    def getNumberOfMiniDoublets(self):
        return lib().lst_event_count(self.h, b"miniDoublets")

    def getNumberOfSegments(self):
        return lib().lst_event_count(self.h, b"segments")

    def getNumberOfTriplets(self):
        return lib().lst_event_count(self.h, b"triplets")

    def getNumberOfQuintuplets(self):
        return lib().lst_event_count(self.h, b"quintuplets")

    def getNumberOfTrackCandidates(self):
        return lib().lst_event_count(self.h, b"trackCandidates")

    def getNumberOfPixelSegments(self):
        return lib().lst_event_count(self.h, b"pixelSegments")

    # --- host copies of SoA fields (SDL::Event::get*, SDL/Event.h:185-204) -----
    def field(self, name: str) -> np.ndarray:
        dt, n = C.c_int(), C.c_size_t()
        _check(lib().lst_event_get_field(self.h, name.encode(), None, 0, C.byref(dt), C.byref(n)))
        out = np.empty(n.value, _DT[dt.value])
        _check(lib().lst_event_get_field(self.h, name.encode(), _p(out), out.nbytes, C.byref(dt), C.byref(n)))
        return out

    # --- SDL::LST::run ----------------------------------------------------------
    def run_lst(self, ev: dict, upto: str = "t3"):
        """``SDL::LST::run`` (SDL/LST.cc:12-161) on one synthetic/real event dict."""
        inp, self._keep = _make_input(ev)
        _check(lib().lst_run(self.h, C.byref(inp), STAGES[upto]))
        return self

    def close(self):
        if self.h:
            lib().lst_event_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Output(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("n_tc", C.c_uint32), ("hit_idxs", C.c_void_p), ("len", C.c_void_p),
                ("seed_idx", C.c_void_p), ("type", C.c_void_p), ("status", C.c_int32)]


class Batch:
    """Batched engine (lst_batch_* in include/lst_amd.h): up to ``slots`` events
    run in one launch set on one GPU, each kernel launched once for all of them,
    inputs copied to the device inside the run.  Per event the output lists
    equal ``Event.run_lst(ev, "tc").getOutput()``."""

    def __init__(self, es: LSTESData, slots: int = 64, t5_mode: int = LST_T5_PLAIN_CHI2):
        L = lib()
        P = C.c_void_p
        L.lst_batch_create.argtypes = [P, C.c_uint32, C.c_int, C.POINTER(P)]
        L.lst_batch_destroy.argtypes = [P]
        L.lst_run_batch.argtypes = [P, P, C.c_uint32, P]
        L.lst_batch_stage.argtypes = [P, P, C.c_uint32]
        L.lst_batch_run_staged.argtypes = [P, C.c_uint32, C.c_uint32, P]
        L.lst_batch_autotune.argtypes = [P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.lst_batch_get_tuning.argtypes = [P, P, C.c_uint32]
        L.lst_batch_set_tuning.argtypes = [P, P, C.c_uint32]
        L.lst_tune_count.restype = C.c_uint32
        L.lst_tune_name.argtypes = [C.c_uint32]
        L.lst_tune_name.restype = C.c_char_p
        self.es = es
        self.slots = slots
        self.h = C.c_void_p()
        _check(L.lst_batch_create(es.h, slots, t5_mode, C.byref(self.h)))

    @staticmethod
    def _inputs(events):
        arr = (_Input * len(events))()
        keep = []
        for i, ev in enumerate(events):
            arr[i], k = _make_input(ev)
            keep.append(k)
        return arr, keep

    @staticmethod
    def _outputs(n, capacity):
        bufs = [{"hits": np.zeros((capacity, 14), np.uint32), "len": np.zeros(capacity, np.uint32),
                 "seed": np.zeros(capacity, np.int32), "type": np.zeros(capacity, np.int16)} for _ in range(n)]
        outs = (_Output * n)()
        for o, b in zip(outs, bufs):
            o.capacity = capacity
            o.hit_idxs, o.len, o.seed_idx, o.type = (b["hits"].ctypes.data, b["len"].ctypes.data,
                                                    b["seed"].ctypes.data, b["type"].ctypes.data)
        return outs, bufs

    @staticmethod
    def _collect(outs, bufs):
        res = []
        for o, b in zip(outs, bufs):
            k = min(o.n_tc, o.capacity)
            res.append({"n_tc": o.n_tc, "status": o.status,
                        "tc_hitIdxs": [list(b["hits"][i, :b["len"][i]]) for i in range(k)],
                        "tc_len": b["len"][:k].copy(), "tc_seedIdx": b["seed"][:k].copy(),
                        "tc_trackCandidateType": b["type"][:k].copy()})
        return res

    def run(self, events, capacity: int = 8192):
        """lst_run_batch on host event dicts; returns one output dict per event."""
        arr, keep = self._inputs(events)
        outs, bufs = self._outputs(len(events), capacity)
        _check(lib().lst_run_batch(self.h, arr, len(events), outs))
        return self._collect(outs, bufs)

    def stage(self, events):
        """Prepare events once and keep them in pinned host memory (staged events 0..len-1)."""
        arr, keep = self._inputs(events)
        _check(lib().lst_batch_stage(self.h, arr, len(events)))

    def run_staged(self, first: int, n: int, capacity: int | None = 8192):
        """Run n staged events from `first` on (cycling modulo the staged count); capacity None
        skips the output lists (the output headers still come back)."""
        if capacity is None:
            _check(lib().lst_batch_run_staged(self.h, first, n, None))
            return None
        outs, bufs = self._outputs(n, capacity)
        _check(lib().lst_batch_run_staged(self.h, first, n, outs))
        return self._collect(outs, bufs)

    def set_sets_in_flight(self, sets: int):
        """lst_batch_set_sets_in_flight: 1 or 2 launch sets in flight (2: twice the arenas)."""
        L = lib()
        L.lst_batch_set_sets_in_flight.argtypes = [C.c_void_p, C.c_uint32]
        _check(L.lst_batch_set_sets_in_flight(self.h, sets))

    def autotune(self, first: int = 0, repeats: int = 2):
        """Choose each grid-stride kernel's work-group count on one launch set of staged events;
        returns (summed kernel ms at the defaults, at the chosen counts, {kernel: count})."""
        L = lib()
        d, t = C.c_float(), C.c_float()
        _check(L.lst_batch_autotune(self.h, first, repeats, C.byref(d), C.byref(t)))
        return d.value, t.value, self.tuning()

    def tuning(self) -> dict:
        L = lib()
        n = L.lst_tune_count()
        g = np.zeros(n, np.uint32)
        _check(L.lst_batch_get_tuning(self.h, _p(g), n))
        return {L.lst_tune_name(k).decode(): int(g[k]) for k in range(n)}

    def set_tuning(self, table: dict):
        L = lib()
        n = L.lst_tune_count()
        names = [L.lst_tune_name(k).decode() for k in range(n)]
        g = np.array([int(table.get(nm, 0)) for nm in names], np.uint32)
        _check(L.lst_batch_set_tuning(self.h, _p(g), n))

    def close(self):
        if self.h:
            lib().lst_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
