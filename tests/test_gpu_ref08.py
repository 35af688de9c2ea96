"""08_city_scale_uav.cpp's own run, compiled from the reference sources: the FP64 engine
reproduces its VTK frames.

The fixture (tests/golden/ref08_city, written by tests/golden/make_ref08_golden.py) holds
sampled values of 08's ASCII frames 0 and 1 (after 1 and 393 steps) for the synthetic city
of city_stl.reference08_city_stl(): DX = 1.5 on 1100x767x167 cells, the 20 Hz sine source
moving at 30 m/s (08:1062-1092), tau 0.55, the reference's sponge and NSCBC faces.  This
pins the composed moving-sine path (08:540-628) over hundreds of steps, buildings crossing
four faces and the top plane included.  The engine's field is printed as 08 prints it
(`std::scientific`, precision 6, of the float value) and compared as text.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import nlos_engine as nlos

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_f64_reproduces_reference08_frames(gpu):
    import sys
    sys.path.insert(0, GOLDEN)
    import city_stl
    frames = json.load(open(os.path.join(GOLDEN, "ref08_city", "vtk_frames.json")))
    nx, ny, nz = frames["grid"]
    cfg = nlos.reference08_config()
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    e = nlos.Engine(cfg)
    assert tuple(e.dims) == (nx, ny, nz)
    e.voxelize_stl(city_stl.reference08_city_stl(), 1.0, -22248.0, -4123.0, -14.0)
    assert e.info().n_solid_boundary > 0
    e.set_source(nlos.reference08_source())
    done = 0
    for fr in sorted(frames["frames"], key=int):
        steps = frames["steps"][fr]
        print(f"[08 frames] frame {fr}: step to {steps}", flush=True)
        e.step(steps - done, capture_field=True)
        done = steps
        field = e.read_field()
        want = frames["frames"][fr]
        st = want["stride"]
        for k, exp in want["slices"].items():
            got = ["%.6e" % float(v) for v in field[int(k), ::st, ::st].ravel()]
            assert len(got) == len(exp)
            bad = [u for u, (a, b) in enumerate(zip(got, exp)) if a != b]
            assert not bad, f"frame {fr} layer {k}: {len(bad)} of {len(exp)} differ, first {got[bad[0]]} vs {exp[bad[0]]}"
        assert np.abs(field).max() > 0
    e.close()


def test_cpp_city_scale_uav_program_reproduces_reference08_frames(gpu, tmp_path):
    """host/city_scale_uav.cpp: 08's main on the C++ LBMSolver surface (lbm_solver.hpp), run
    like the reference in a directory holding csx_model_binary.stl, frames written as legacy
    VTK BINARY and the run stopped after frame 1.  The header (but the format line) and the
    sampled values of frames 0 and 1 equal the compiled 08's."""
    import subprocess
    import sys
    sys.path.insert(0, GOLDEN)
    import city_stl
    exe = os.path.join(os.path.dirname(nlos.LIB_PATH), "city_scale_uav")
    assert os.path.exists(exe), "build the package (make) first"
    frames = json.load(open(os.path.join(GOLDEN, "ref08_city", "vtk_frames.json")))
    nx, ny, nz = frames["grid"]
    (tmp_path / "csx_model_binary.stl").write_bytes(city_stl.reference08_city_stl())
    print("[08 C++ program] run to frame 1", flush=True)
    out = subprocess.run([exe, "--binary", "--max-frames", "2"], cwd=tmp_path, capture_output=True, text=True,
                         timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    for fr, want in frames["frames"].items():
        raw = (tmp_path / "output_mpi" / f"flight_{int(fr):05d}.vtk").read_bytes()
        end = raw.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
        header = raw[:end].decode().rstrip("\n").split("\n")
        assert header[2] == "BINARY"
        assert header[:2] + header[3:] == want["header"][:2] + want["header"][3:]
        field = np.frombuffer(raw, dtype=">f4", count=nx * ny * nz, offset=end).reshape(nz, ny, nx)
        st = want["stride"]
        for k, exp in want["slices"].items():
            got = ["%.6e" % float(v) for v in field[int(k), ::st, ::st].ravel()]
            bad = [u for u, (a, b) in enumerate(zip(got, exp)) if a != b]
            assert len(got) == len(exp) and not bad, f"frame {fr} layer {k}: {len(bad)} differ"
