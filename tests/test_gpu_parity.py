"""GPU parity of the HIP engine against the oracle and the compiled reference's goldens.

FP64 (full populations, reference operation order) must be bit-identical to the oracle
and reproduce the reference's probe files byte for byte.  FP32 and FP16S (deviation
storage, FP32 arithmetic) are held to the relative-L2 tolerances written in each test.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np
import pytest

import nlos_engine as nlos
import scenes
from oracle_ffi import Oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu

# Reduced-precision tolerances on probe traces and fields, relative L2 against FP64.
TOL_F32 = 1e-4
TOL_F16 = 3e-2


def _engine_run(cfg, solid, src, probes, steps, rho=None):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if src is not None:
        e.set_source(src)
    e.set_probes(probes)
    e.step(steps, capture_field=True)
    e.synchronize()
    return e


def _oracle_run(cfg, solid, src, probes, steps, rho=None):
    o = Oracle(cfg)
    o.set_mask(solid)
    if rho is not None:
        o.set_density(rho)
    if src is not None:
        o.set_source(src)
    o.set_probes(probes)
    o.run(steps)
    return o


def _oracle_interior_f(o):
    return o.read_f()[1:-1, 1:-1, 1:-1, :]


PATTERNS = ["inplace", "two_array"]


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("scene", ["box", "box64", "narrow", "wide", "pulse", "reflect"])
def test_f64_bit_identical_to_oracle(gpu, scene, pattern):
    print(f"[f64 vs oracle] scene={scene} pattern={pattern}", flush=True)
    if scene == "box":
        cfg, solid, src, probes = scenes.box_scene("f64")
        rho, steps = None, 60
    elif scene == "box64":      # nx a multiple of 64: the odd sweep's lane-shift path
        cfg, solid, src, probes = scenes.box_scene("f64", nx=64, ny=32, nz=20)
        rho, steps = None, 60
    elif scene == "narrow":     # rows shorter than 19 cells: x-face records from one lane
        cfg, solid, src, probes = scenes.box_scene("f64", nx=12, ny=20, nz=14)
        rho, steps = None, 40
    elif scene == "wide":       # a partial last row segment, columns clipped at nz
        cfg, solid, src, probes = scenes.box_scene("f64", nx=100, ny=10, nz=37)
        rho, steps = None, 40
    elif scene == "pulse":
        cfg, solid, rho, probes = scenes.pulse_scene("f64")
        src, steps = None, 40
    else:
        cfg, solid, src, probes = scenes.reflect_scene("f64")
        rho, steps = None, 50
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = _engine_run(cfg, solid, src, probes, steps, rho)
    assert e.info().stream_pattern == nlos.PATTERNS[pattern]
    o = _oracle_run(cfg, solid, src, probes, steps, rho)
    got, want = e.read_probes(), o.read_probes()
    assert got.shape == want.shape
    assert np.abs(want).max() > 0.0
    assert np.array_equal(got, want), np.abs(got - want).max()
    assert np.array_equal(e.read_field(), o.read_field())
    fe, fo = e.debug_read_f(), _oracle_interior_f(o)
    fluid = solid == 0
    assert np.array_equal(fe[fluid], fo[fluid])
    e.close()


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("precision,tol", [("f32", TOL_F32), ("f16s", TOL_F16)])
@pytest.mark.parametrize("scene", ["box", "reflect"])
def test_reduced_precision_vs_f64(gpu, precision, tol, scene, pattern):
    print(f"[{precision} vs f64] scene={scene} pattern={pattern}", flush=True)
    make = scenes.box_scene if scene == "box" else scenes.reflect_scene
    cfg, solid, src, probes = make(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = _engine_run(cfg, solid, src, probes, 80)
    cfg64, _, _, _ = make("f64")
    o = _oracle_run(cfg64, solid, src, probes, 80)
    err = scenes.rel_l2(e.read_probes(), o.read_probes())
    ferr = scenes.rel_l2(e.read_field(), o.read_field())
    print(f"  probe rel-L2 {err:.3e}  field rel-L2 {ferr:.3e}")
    assert err <= tol and ferr <= tol
    e.close()


@pytest.mark.parametrize("precision,pattern", [("f64", "inplace"), ("f64", "two_array"),
                                               ("f16s", "two_array"), ("b16", "two_array")])
def test_three_phase_api_equals_fused_step(gpu, precision, pattern):
    """apply_source, exchange_halos and step_collision driven one by one equal nlos_step
    (the x-halo records of the 16-bit two-array runs included)."""
    cfg, solid, src, probes = scenes.box_scene(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    a = _engine_run(cfg, solid, src, probes, 25)
    b = nlos.Engine(cfg)
    b.set_mask(solid)
    b.set_source(src)
    b.set_probes(probes)
    dt = b.dt
    for s in range(25):
        b.apply_source(s * dt)
        b.exchange_halos()
        b.step_collision_bgk()
    b.synchronize()
    assert np.array_equal(a.read_probes(), b.read_probes())
    assert np.array_equal(a.debug_read_f()[solid == 0], b.debug_read_f()[solid == 0])


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("precision", ["f32", "f16s"])
def test_rest_state_is_exact_fixed_point(gpu, precision, pattern):
    cfg, solid, _, probes = scenes.box_scene(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_probes(probes)
    e.step(31, capture_field=True)
    e.synchronize()
    assert not np.any(e.read_probes())
    assert not np.any(e.read_field())
    f = e.debug_read_f()[solid == 0]
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    assert np.array_equal(f, np.broadcast_to(w, f.shape))


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("precision", ["f64", "f16s"])
def test_checkpoint_restore_is_bitwise(gpu, precision, pattern):
    cfg, solid, src, probes = scenes.box_scene(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = _engine_run(cfg, solid, src, probes, 17)
    blob = e.checkpoint()
    e.step(20)
    e.synchronize()
    p1, f1 = e.read_probes(), e.debug_read_f()
    e.close()
    r = nlos.Engine(cfg)
    r.set_mask(solid)
    r.set_source(src)
    r.set_probes(probes)
    r.restore(blob)
    r.step(20)
    r.synchronize()
    assert np.array_equal(r.read_probes(), p1)
    assert np.array_equal(r.debug_read_f(), f1, equal_nan=True)


@pytest.mark.parametrize("precision,scene", [("f32", "box"), ("f32", "wide"), ("f32", "reflect"),
                                             ("f16s", "box"), ("f16s", "wide")])
def test_inplace_equals_two_array_bitwise(gpu, precision, scene):
    """The in-place A-A sweep and the two-array pull sweep are two kernels and two layouts
    of the same step: probes, field and populations agree bit for bit (rows of 40, 100 and
    30 cells, so no row is a whole number of 64-cell blocks; partial reflection in FP32).
    Scaled FP16 with partial reflection is not bitwise: the in-place even sweep stores the
    reflected wall copy (rounding after reflecting), the pull sweep reflects the stored,
    rounded value; test_inplace_and_two_array_fp16_reflection_agree_to_rounding covers it."""
    if scene == "reflect":
        cfg, solid, src, probes = scenes.reflect_scene(precision)
    else:
        dims = {} if scene == "box" else {"nx": 100, "ny": 10, "nz": 37}
        cfg, solid, src, probes = scenes.box_scene(precision, **dims)
    runs = []
    for pattern in PATTERNS:
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        e = _engine_run(cfg, solid, src, probes, 33)
        runs.append((e.read_probes(), e.read_field(), e.debug_read_f()[solid == 0]))
        e.close()
    assert np.abs(runs[0][1]).max() > 0
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


def _rows_of_three_waves(precision):
    """Rows of 192 cells, three 64-cell waves: solids ending on both sides of the wave
    boundaries at x = 64 and 128 and at the row ends, a density field that differs from
    cell to cell (so a value moved to the wrong lane shows at once) and a moving source."""
    nx, ny, nz = 192, 20, 16
    cfg, solid, _, _ = scenes.box_scene(precision, nx=nx, ny=ny, nz=nz)
    solid[:6, 3:7, 62:66] = 1
    solid[:9, 9:12, 127:129] = 1
    solid[:4, 14:18, 188:192] = 1
    solid[:3, 0:3, 0:2] = 1
    solid[5:8, 12:15, 63:64] = 1
    rng = np.random.default_rng(11)
    rho = 1.0 + 1e-3 * rng.random((nz, ny, nx))
    src = nlos.default_source(sx=60.3, sy=9.7, sz=6.2, ex=131.1, ey=12.4, ez=8.0, speed=900.0,
                              amp=2e-2, freq=120.0, waveform=nlos.WAVE_SINE)
    probes = [(63.5, 10.5, 4.5), (64.5, 5.5, 9.5), (127.5, 12.5, 2.5), (190.5, 3.5, 12.5),
              (0.5, 18.5, 6.5)]
    return cfg, solid, src, probes, rho


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s"])
def test_inplace_rows_of_whole_waves(gpu, precision):
    """Rows a whole number of 64-cell waves, solids at the wave boundaries and a density
    that differs from cell to cell: the in-place A-A pattern equals the two-array pattern
    bit for bit (partial reflection in FP64 and FP32; scaled FP16 with partial reflection
    rounds the wall copies in another order in the two patterns, so it runs at 1), and
    FP64 equals the oracle."""
    cfg, solid, src, probes, rho = _rows_of_three_waves(precision)
    cfg.wall_reflect = 1.0 if precision == "f16s" else 0.8
    runs = []
    for pattern in ("inplace", "two_array"):
        print(f"[rows of whole waves] {precision} {pattern}", flush=True)
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        e = _engine_run(cfg, solid, src, probes, 41, rho)
        runs.append((e.read_probes(), e.read_field(), e.debug_read_f()[solid == 0]))
        e.close()
    assert np.abs(runs[0][1]).max() > 0
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    if precision == "f64":
        o = _oracle_run(cfg, solid, src, probes, 41, rho)
        assert np.array_equal(runs[0][0], o.read_probes())
        assert np.array_equal(runs[0][1], o.read_field())
        assert np.array_equal(runs[0][2], _oracle_interior_f(o)[solid == 0])


def test_inplace_and_two_array_fp16_reflection_agree_to_rounding(gpu):
    """Scaled FP16 with wall_reflect = 0.8: the two patterns round the reflected wall values
    in a different order, so they agree to FP16 rounding noise, far inside the FP16S
    tolerance against FP64 (TOL_F16)."""
    cfg, solid, src, probes = scenes.reflect_scene("f16s")
    runs = []
    for pattern in PATTERNS:
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        e = _engine_run(cfg, solid, src, probes, 33)
        runs.append((e.read_probes(), e.read_field()))
        e.close()
    perr = scenes.rel_l2(runs[0][0], runs[1][0])
    ferr = scenes.rel_l2(runs[0][1], runs[1][1])
    print(f"  in place vs two arrays: probe {perr:.2e}, field {ferr:.2e}")
    assert 0.0 < ferr and perr <= 1e-2 * TOL_F16 and ferr <= 1e-2 * TOL_F16


@pytest.mark.parametrize("precision", ["f64", "f16s", "b16"])
def test_field_capture_leaves_the_state_unchanged(gpu, precision):
    """The capture variants of the sweep write rho-1 besides the step: the same populations
    and probes with the field captured on the last step or never."""
    cfg, solid, src, probes = scenes.box_scene(precision)
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    a = _engine_run(cfg, solid, src, probes, 19)            # captures the last step
    b = nlos.Engine(cfg)
    b.set_mask(solid)
    b.set_source(src)
    b.set_probes(probes)
    b.step(19)
    b.synchronize()
    assert np.array_equal(a.read_probes(), b.read_probes())
    assert np.array_equal(a.debug_read_f(), b.debug_read_f(), equal_nan=True)
    assert np.abs(a.read_field()).max() > 0
    a.close()
    b.close()


def test_checkpoint_of_another_format_version_is_refused(gpu):
    """Format version 3 holds the 16-bit modes' populations blocked by 64 cells; a blob that
    claims another version (bytes 8..11 after the magic) is refused, whatever its size."""
    cfg, solid, src, probes = scenes.box_scene("f16s")
    e = _engine_run(cfg, solid, src, probes, 2)
    blob = bytearray(e.checkpoint())
    assert struct.unpack_from("<i", blob, 8)[0] == 3
    struct.pack_into("<i", blob, 8, 2)
    with pytest.raises(nlos.EngineError) as ex:
        e.restore(bytes(blob))
    assert ex.value.code == 1 and "version 2" in str(ex.value)
    e.close()


def test_nonfinite_population_is_reported(gpu):
    cfg, solid, src, probes = scenes.box_scene("f32")
    e = _engine_run(cfg, solid, src, probes, 3)
    e.debug_poke(20, 3, 20, 7, float("nan"))
    e.step(2)
    with pytest.raises(nlos.EngineError) as ex:
        e.synchronize()
    assert ex.value.code == 4


def test_voxelizer_matches_oracle_bitwise(gpu):
    import city_stl
    cfg = nlos.reference05_config()
    stl = city_stl.reference_city_stl()
    e = nlos.Engine(cfg)
    e.voxelize_stl(stl, 1.0, -22248.0, -4123.0, -14.0)
    o = Oracle(cfg)
    assert o.load_stl(stl, 1.0, -22248.0, -4123.0, -14.0) == 0
    assert np.array_equal(e.ground_heights(), o.ground_heights())
    m = e.get_mask()
    assert np.array_equal(m, o.get_mask())
    assert m.sum() > 0
    assert e.info().n_solid_boundary > 0


def _reference05_engine(scene, pattern="auto"):
    import city_stl
    cfg = nlos.reference05_config()
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = nlos.Engine(cfg)
    if scene == "city":
        e.voxelize_stl(city_stl.reference_city_stl(), 1.0, -22248.0, -4123.0, -14.0)
    else:
        assert e.load_stl("/nonexistent/csx_model_binary.stl") == 5   # empty geometry, 05:407-410
    names, pts = nlos.reference05_probes()
    e.set_probes(pts, names)
    e.set_source(nlos.reference05_source())
    return e


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("scene", ["flat", "city"])
def test_f64_reproduces_reference_probe_files(gpu, scene, pattern, tmp_path):
    """05_sensor_illusion's own run (393+1 steps) compiled from the reference sources."""
    print(f"[reference goldens] scene={scene} pattern={pattern}", flush=True)
    e = _reference05_engine(scene, pattern)
    frames = json.load(open(os.path.join(GOLDEN, f"ref05_{scene}", "vtk_frames.json")))
    total = int(round(1.0 / e.dt))
    assert total == 393
    done = 0
    for fr in ("50", "98"):
        s = int(fr) * 4                      # out_interval = 393 // 80 = 4
        e.step(s + 1 - done, capture_field=True)
        done = s + 1
        e.write_vtk(str(tmp_path / "f.vtk"))
        lines = open(tmp_path / "f.vtk").read().split("\n")
        want = frames["frames"][fr]
        assert lines[:10] == want["header"]
        vals = lines[10:]
        nx, ny, nz = frames["grid"]
        st = want["stride"]
        for k, exp in want["slices"].items():
            k = int(k)
            got = [vals[(k * ny + j) * nx + i] for j in range(0, ny, st) for i in range(0, nx, st)]
            assert got == exp, f"frame {fr} slice {k}"
    e.step(total + 1 - done)
    e.write_probe_files(str(tmp_path))
    for n in ("B1", "B2", "C1", "C2"):
        got = open(tmp_path / f"probe_{n}.txt").read()
        want = open(os.path.join(GOLDEN, f"ref05_{scene}", f"probe_{n}.txt")).read()
        assert got == want, n
    e.close()


def test_run_loop_frames(gpu, tmp_path):
    """Engine.run mirrors LBMSolver::run's frame cadence (05:1064-1097)."""
    cfg, solid, src, probes = scenes.box_scene("f32")
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_source(src)
    e.set_probes(probes, [f"P{i}" for i in range(len(probes))])
    total = e.run(40 * e.dt, 5, vtk_dir=str(tmp_path / "vtk"), probe_dir=str(tmp_path))
    assert total == 40
    assert len(os.listdir(tmp_path / "vtk")) == 40 // (40 // 5) + 1
    assert len(open(tmp_path / "probe_P0.txt").read().split()) == 41


@pytest.mark.parametrize("scene", ["city", "flat"])
def test_cpp_lbmsolver_program_reproduces_reference(gpu, scene, tmp_path):
    """host/sensor_illusion.cpp: 05's main on the C++ LBMSolver surface (lbm_solver.hpp),
    run like the reference in a directory holding csx_model_binary.stl."""
    import subprocess
    import city_stl
    exe = os.path.join(os.path.dirname(nlos.LIB_PATH), "sensor_illusion")
    assert os.path.exists(exe), "build the package (make) first"
    if scene == "city":
        (tmp_path / "csx_model_binary.stl").write_bytes(city_stl.reference_city_stl())
    out = subprocess.run([exe, "--no-vtk"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    for n in ("B1", "B2", "C1", "C2"):
        got = (tmp_path / f"probe_{n}.txt").read_text()
        want = open(os.path.join(GOLDEN, f"ref05_{scene}", f"probe_{n}.txt")).read()
        assert got == want, n


@pytest.mark.parametrize("precision", ["f64", "f16s"])
def test_lane_shift_path_equals_plain_loads(gpu, precision, monkeypatch):
    """The odd sweep's aligned-load + lane-shift path (64x4 tiles of full 64-cell rows)
    gives the same bits as plain shifted loads, on 64x4 tiles and on the default 256x1
    row segments."""
    cfg, solid, src, probes = scenes.box_scene(precision, nx=128, ny=24, nz=16)
    cfg.stream_pattern = nlos.PATTERN_INPLACE
    runs = []
    for block, flag in (("64x4", "1"), ("64x4", "0"), ("256x1", "1")):
        monkeypatch.setenv("NLOS_BLOCK", block)
        monkeypatch.setenv("NLOS_DPP", flag)
        e = _engine_run(cfg, solid, src, probes, 21)
        runs.append((e.read_probes(), e.debug_read_f()))
        e.close()
    for r in runs[1:]:
        assert np.array_equal(runs[0][0], r[0])
        assert np.array_equal(runs[0][1], r[1], equal_nan=True)


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s"])
def test_two_array_launch_shapes_agree(gpu, precision, monkeypatch):
    """The two-array sweep gives the same bits for every workgroup shape (partial last row
    segments and partial tile rows included)."""
    cfg, solid, src, probes = scenes.box_scene(precision, nx=100, ny=13, nz=37)
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    runs = []
    for block in ("256x1", "64x4", "128x2"):
        monkeypatch.setenv("NLOS_BLOCK", block)
        e = _engine_run(cfg, solid, src, probes, 23)
        runs.append((e.read_probes(), e.debug_read_f()))
        e.close()
    for r in runs[1:]:
        assert np.array_equal(runs[0][0], r[0])
        assert np.array_equal(runs[0][1], r[1], equal_nan=True)
