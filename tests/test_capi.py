"""The C-ABI library loads and exports every entry point include/lst_amd.h declares (CPU only)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "lst_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(lst_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("lst_es_create", "lst_event_create", "lst_event_add_hits", "lst_event_add_pixel_segments",
              "lst_event_create_mini_doublets", "lst_event_create_segments", "lst_event_create_triplets",
              "lst_run", "lst_event_get_field"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import tracklooper_amd as lst
    if not os.path.exists(lst.LIB_PATH):
        pytest.skip("liblst_amd.so not built (run python -m tracklooper_amd.build)")
    lib = C.CDLL(lst.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_paths_without_device():
    """Argument errors are reported as status codes, without touching a GPU."""
    import tracklooper_amd as lst
    if not os.path.exists(lst.LIB_PATH):
        pytest.skip("liblst_amd.so not built")
    L = lst.lib()
    assert L.lst_event_create_mini_doublets(None) == 4  # LST_E_ORDER
    assert L.lst_event_get_field(None, b"hits.xs", None, 0, None, None) == 3  # LST_E_INPUT
    L.lst_batch_set_sets_in_flight.argtypes = [C.c_void_p, C.c_uint32]
    assert L.lst_batch_set_sets_in_flight(None, 2) == 3  # LST_E_INPUT
    es = C.c_void_p()
    assert L.lst_es_create(b"/nonexistent/geometry", 0.8, 0, C.byref(es)) == 1  # LST_E_IO
    assert b"cannot read" in L.lst_last_error()
