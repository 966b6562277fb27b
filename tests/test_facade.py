"""The C++ facade with the reference's host interface (include/lst_sdl.hpp:
SDL::loadAndFillESHost/Device, SDL::Event, SDL::LST::run) compiles against
the C ABI (CPU) and, driven by tests/cpp/lst_facade_main.cpp, gives the
oracle's output lists and object counts (GPU)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(ROOT, "tracklooper_amd", "_lib")


def build_driver(out_dir) -> str:
    exe = os.path.join(str(out_dir), "lst_facade_main")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "lst_facade_main.cpp"), "-L", LIB_DIR, "-llst_amd",
           f"-Wl,-rpath,{LIB_DIR}", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_facade_compiles(tmp_path):
    if not os.path.exists(os.path.join(LIB_DIR, "liblst_amd.so")):
        pytest.skip("liblst_amd.so is not built")
    assert os.path.exists(build_driver(tmp_path))


def _write_run_args(path, ev):
    keys = ["see_px", "see_py", "see_pz", "see_dxy", "see_dz", "see_ptErr", "see_etaErr", "see_stateTrajGlbX",
            "see_stateTrajGlbY", "see_stateTrajGlbZ", "see_stateTrajGlbPx", "see_stateTrajGlbPy",
            "see_stateTrajGlbPz"]
    with open(path, "wb") as f:
        f.write(np.array([len(ev["see_px"]), len(ev["ph2_x"])], np.uint32).tobytes())
        for k in keys:
            f.write(np.asarray(ev[k], np.float32).tobytes())
        f.write(np.asarray(ev["see_q"], np.int32).tobytes())
        f.write(np.asarray(ev["see_hitIdx_off"], np.uint32).tobytes())
        f.write(np.asarray(ev["see_hitIdx"], np.int32).tobytes())
        f.write(np.asarray(ev["ph2_detId"], np.uint32).tobytes())
        for k in ("ph2_x", "ph2_y", "ph2_z"):
            f.write(np.asarray(ev[k], np.float32).tobytes())


def _write_prepared(path, ref):
    f_ = ref.field
    with open(path, "wb") as f:
        f.write(np.array([len(f_("in_x")), len(f_("in_ptIn"))], np.uint32).tobytes())
        for k, t in (("in_x", np.float32), ("in_y", np.float32), ("in_z", np.float32), ("in_detId", np.uint32),
                     ("in_idx", np.uint32), ("in_hitIdx0", np.uint32), ("in_hitIdx1", np.uint32),
                     ("in_hitIdx2", np.uint32), ("in_hitIdx3", np.uint32), ("in_dPhiChange", np.float32),
                     ("in_ptIn", np.float32), ("in_ptErr", np.float32), ("in_px", np.float32), ("in_py", np.float32),
                     ("in_pz", np.float32), ("in_eta", np.float32), ("in_etaErr", np.float32), ("in_phi", np.float32),
                     ("in_charge", np.int32), ("in_seedIdx", np.uint32), ("in_superbin", np.int32),
                     ("in_pixelType", np.int8), ("in_isQuad", np.int8)):
            f.write(np.asarray(f_(k), t).tobytes())


@pytest.mark.gpu
def test_facade_matches_oracle(tmp_path, oracle_geometry, geom_dir):
    from oracle import orc
    from tracklooper_amd import synth
    exe = build_driver(tmp_path)
    ev = synth.generate_event("pu", 11, n_pu=200)
    ref = orc.Event(oracle_geometry).run_lst(ev, "tc")
    _write_run_args(tmp_path / "args.bin", ev)
    _write_prepared(tmp_path / "prep.bin", ref)
    out = tmp_path / "out.txt"
    p = subprocess.run([exe, geom_dir, str(tmp_path / "args.bin"), str(tmp_path / "prep.bin"), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = out.read_text().splitlines()
    assert lines[0] == "order_error 1"
    counts = [int(x) for x in lines[1].split()[1:]]
    n = ref.scalar("nTrackCandidates")
    # getNumberOfMiniDoublets / Segments count outer-tracker objects (SDL/Event.cc:1290-1330): the pixel
    # module's entry (the last) is left out
    want = [len(ref.field("in_x")), int(ref.field("mds.nMDs")[:-1].sum()), int(ref.field("segs.nSegments")[:-1].sum()),
            int(ref.field("t3.nTriplets").sum()), int(ref.field("t5.nQuintuplets").sum()),
            ref.scalar("nPixelQuintuplets"), ref.scalar("nPixelTriplets"), n, n - ref.scalar("nTrackCandidatesT5")]
    assert counts == want
    assert int(lines[3].split()[1]) == n
    flat, off = ref.field("out.hitIdxs"), ref.field("out.hitOffsets")
    typ, seed = ref.field("out.trackCandidateType"), ref.field("out.seedIdx")
    for i, line in enumerate(lines[4:4 + n]):
        v = [int(x) for x in line.split()[1:]]
        assert v[0] == typ[i] and v[1] == seed[i] and v[2] == off[i + 1] - off[i]
        assert v[3:] == list(flat[off[i]:off[i + 1]])
