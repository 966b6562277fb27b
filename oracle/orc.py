"""ORACLE — test infrastructure only.

ctypes binding of oracle/_build/liblst_oracle.so, the serial CPU restatement
of the reference LST chain.  Imported only by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, as the checker; never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liblst_oracle.so")
_DT = {0: np.float32, 1: np.uint32, 2: np.int32, 3: np.uint16, 4: np.int16, 5: np.uint8, 6: np.int8, 7: np.float64}
STAGES = {"hits": 1, "pls": 2, "md": 3, "ls": 4, "t3": 5, "t5": 6, "plsclean": 7, "pt5": 8, "pt3": 9, "tc": 10}
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with make (gcc, IEEE, no FP contraction)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE, "-j8"] + (["-B"] if force else []))
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_geometry_load.restype = C.c_void_p
        L.orc_geometry_load.argtypes = [C.c_char_p, C.c_uint, C.c_char_p, C.c_int]
        L.orc_geometry_free.argtypes = [C.c_void_p]
        L.orc_geometry_field.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                         C.POINTER(C.c_long)]
        L.orc_geometry_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_geometry_scalar.restype = C.c_uint
        L.orc_event_new.restype = C.c_void_p
        L.orc_event_new.argtypes = [C.c_void_p, C.c_float, C.c_int]
        L.orc_event_free.argtypes = [C.c_void_p]
        L.orc_prepare.argtypes = [C.c_void_p, C.c_uint] + [C.c_void_p] * 16 + [C.c_uint] + [C.c_void_p] * 4
        L.orc_run.argtypes = [C.c_void_p, C.c_int]
        L.orc_event_field.argtypes = L.orc_geometry_field.argtypes
        L.orc_event_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_event_scalar.restype = C.c_uint
        L.orc_t5dnn_forward.argtypes = [C.c_void_p]
        L.orc_t5dnn_forward.restype = C.c_float
        L.orc_t5dnn_features.argtypes = [C.c_void_p, C.c_uint, C.c_void_p]
        L.orc_t5dnn_features.restype = C.c_int
        L.orc_t5dnn_set_weights.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _field(fn, h, name):
    p, t, n = C.c_void_p(), C.c_int(), C.c_long()
    if not fn(h, name.encode(), C.byref(p), C.byref(t), C.byref(n)):
        raise KeyError(name)
    dt = np.dtype(_DT[t.value])
    if n.value == 0:
        return np.zeros(0, dt)
    buf = (C.c_char * (n.value * dt.itemsize)).from_address(p.value)
    return np.frombuffer(buf, dt).copy()


class Geometry:
    def __init__(self, geom_dir: str, stride: int = 40):
        err = C.create_string_buffer(512)
        self.h = lib().orc_geometry_load(geom_dir.encode(), stride, err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def field(self, name):
        return _field(lib().orc_geometry_field, self.h, name)

    def scalar(self, name):
        return lib().orc_geometry_scalar(self.h, name.encode())

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_geometry_free(self.h)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Event:
    """One serial run of the chain; fields are read back by reference name."""

    def __init__(self, geom: Geometry, pt_cut: float = 0.8, t5_mode: int = 0):
        self.geom = geom
        self.h = lib().orc_event_new(geom.h, pt_cut, t5_mode)

    def run_lst(self, ev: dict, upto: str = "tc"):
        a = {k: np.ascontiguousarray(v) for k, v in ev.items()}
        ns = len(a["see_px"])
        keys = ["see_px", "see_py", "see_pz", "see_dxy", "see_dz", "see_ptErr", "see_etaErr", "see_stateTrajGlbX",
                "see_stateTrajGlbY", "see_stateTrajGlbZ", "see_stateTrajGlbPx", "see_stateTrajGlbPy",
                "see_stateTrajGlbPz", "see_q", "see_hitIdx", "see_hitIdx_off"]
        args = [_ptr(a[k]) for k in keys]
        hits = [_ptr(a[k]) for k in ("ph2_detId", "ph2_x", "ph2_y", "ph2_z")]
        lib().orc_prepare(self.h, ns, *args, len(a["ph2_x"]), *hits)
        rc = lib().orc_run(self.h, STAGES[upto])
        if rc != 0:
            raise RuntimeError(f"oracle run failed ({rc})")
        return self

    def field(self, name):
        return _field(lib().orc_event_field, self.h, name)

    def scalar(self, name):
        return lib().orc_event_scalar(self.h, name.encode())

    def t5dnn_features(self, t5: int) -> np.ndarray:
        """The 38 DNN inputs of stored quintuplet slot t5 (SDL/NeuralNetwork.h:36-114)."""
        out = np.zeros(38, np.float32)
        if lib().orc_t5dnn_features(self.h, t5, out.ctypes.data_as(C.c_void_p)) != 0:
            raise IndexError(t5)
        return out

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_event_free(self.h)
