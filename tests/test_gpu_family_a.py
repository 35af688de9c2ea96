"""Legacy family-A mode (NLOS_FAMILY_A: the step_pull of 01_acoustic_scattering.cpp and
04_hpc_scalability.cpp) on the GPU.

* 01's own run reproduces the compiled reference's probe_<A|B|C>.dat files byte for byte
  (all 402 records) and its VTK frames at steps 104 and 208 (sampled values, header);
* on smaller scenes the engine's populations equal the oracle's (oracle/ref_cpu_a.cpp)
  bit for bit, with the default and with a radial (04-style) sponge;
* checkpoint/restore is bitwise; family-B-only features are refused.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import family_a as fa
import nlos_engine as nlos
from oracle_ffi import OracleA

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref01")


def _engine(cfg, solid, rho, names=None, pts=None):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if pts is not None:
        e.set_probes(pts, names)
    return e


def test_ref01_probe_files_are_byte_identical(gpu, tmp_path):
    cfg = fa.config()
    solid, rho = fa.cylinder_and_pulse(cfg)
    names, pts = fa.ref01_probes()
    e = _engine(cfg, solid, rho, names, pts)
    n = fa.ref01_steps(cfg)                    # records s = 0 .. n
    print(f"[family A] 01 run: {n} steps", flush=True)
    e.step(n)
    e.synchronize()
    e.write_probe_dat(str(tmp_path), scale=100.0)
    for name in names:
        got = (tmp_path / f"probe_{name}.dat").read_text()
        want = open(os.path.join(GOLDEN, f"probe_{name}.dat")).read()
        assert got == want, name


def test_ref01_vtk_frames(gpu, tmp_path):
    frames = json.load(open(os.path.join(GOLDEN, "vtk_frames.json")))
    nx, ny, nz = frames["grid"]
    cfg = fa.config()
    solid, rho = fa.cylinder_and_pulse(cfg)
    e = _engine(cfg, solid, rho)
    done = 0
    for f in ("104", "208"):
        want = frames["frames"][f]
        e.step(int(f) - done, capture_field=True)
        done = int(f)
        path = tmp_path / f"cyl_{f}.vtk"
        e.write_vtk(str(path))
        lines = path.read_text().split("\n")
        assert lines[:10] == want["header"], f
        k, st = want["layer"], want["stride"]
        got = [lines[10 + (k * ny + j) * nx + i] for j in range(0, ny, st) for i in range(0, nx, st)]
        assert got == want["values"], f


def test_cpp_acoustic_scattering_program_reproduces_reference(gpu, tmp_path):
    """host/acoustic_scattering.cpp: 01's run through the C ABI.  Its probe files equal the
    compiled 01's byte for byte, its frames at steps 104 and 208 hold 01's sampled values, and
    its frame of step 0 is the initial pulse in 01's format."""
    import subprocess
    exe = os.path.join(os.path.dirname(nlos.LIB_PATH), "acoustic_scattering")
    assert os.path.exists(exe), "build the package (make) first"
    out = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    for name in ("A", "B", "C"):
        got = (tmp_path / "output_mpi" / f"probe_{name}.dat").read_text()
        assert got == open(os.path.join(GOLDEN, f"probe_{name}.dat")).read(), name
    frames = json.load(open(os.path.join(GOLDEN, "vtk_frames.json")))
    nx, ny, nz = frames["grid"]
    for f in ("104", "208"):
        want = frames["frames"][f]
        lines = (tmp_path / "output_mpi" / f"cyl_{int(f):05d}.vtk").read_text().split("\n")
        assert lines[:10] == want["header"], f
        k, st = want["layer"], want["stride"]
        got = [lines[10 + (k * ny + j) * nx + i] for j in range(0, ny, st) for i in range(0, nx, st)]
        assert got == want["values"], f
    lines = (tmp_path / "output_mpi" / "cyl_00000.vtk").read_text().split("\n")
    assert lines[:10] == frames["frames"]["104"]["header"]
    cfg = fa.config()
    solid, rho = fa.cylinder_and_pulse(cfg)
    rho[solid == 1] = 1.0
    assert lines[10:-1] == ["%.6e" % ((v - 1.0) * 100.0) for v in rho.ravel()]


def test_cpp_hpc_scalability_program_prints_04s_report(gpu, tmp_path):
    """host/hpc_scalability.cpp: 04's benchmark (208^3, line source, 100 + 2000 steps) through
    the C ABI, reporting as 04 does (04:366-401).  04 prints only its timing, so this checks
    the report and a rate no CPU run of 04 reaches."""
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(nlos.LIB_PATH), "hpc_scalability")
    assert os.path.exists(exe), "build the package (make) first"
    out = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == "=== START BENCHMARK ===" and lines[-1] == "=== END BENCHMARK ==="
    assert "Grid: 208 x 208 x 208 (dx=0.125)" in lines
    mlups = float(re.search(r"^MLUPS: (\S+)$", out.stdout, re.M).group(1))
    secs = float(re.search(r"^BENCHMARK_TIME: (\S+) s$", out.stdout, re.M).group(1))
    print(f"[04 C++ program] {mlups:.0f} MLUPS, {secs:.3f} s", flush=True)
    assert abs(208 ** 3 * 2000 / secs / 1e6 - mlups) <= 1e-3 * mlups
    assert mlups > 2000.0


def _small_scene(seed=4):
    cfg = fa.config(x=(-1.2, 1.2), y=(-1.0, 1.0), z=(-0.15, 0.15), dx=0.05, tau=0.52, sponge=0.3, strength=0.8)
    nx, ny, nz = fa.dims(cfg)
    solid, rho = fa.cylinder_and_pulse(cfg, radius=0.25, centre=(0.5, 0.1), width=0.3, amp=0.05)
    rng = np.random.default_rng(seed)
    solid[:, 3:7, 30:36] = 1                     # a block touching nothing, z-invariant
    solid[1, 15:18, 8:12] = 1                    # a block in one z plane (periodic neighbours)
    rho = rho + 1e-4 * rng.standard_normal(rho.shape) * (solid == 0)
    pts = [(0.0, 0.5, 0.0), (-0.7, -0.4, 0.1), (0.9, 0.8, -0.1)]
    return cfg, solid, rho, pts


@pytest.mark.parametrize("radial", [False, True])
def test_engine_equals_oracle_bitwise(gpu, radial):
    cfg, solid, rho, pts = _small_scene()
    e = _engine(cfg, solid, rho, None, pts)
    o = OracleA(cfg)
    o.set_mask(solid)
    o.set_density(rho)
    o.set_probes(pts)
    if radial:
        # 04:182-206: sigma = s * min(1, (r - R0)/T)^2 beyond R0
        nx, ny, nz = fa.dims(cfg)
        sg = np.zeros((nz, ny, nx))
        for j in range(ny):
            for i in range(nx):
                x = cfg.x_min + (i + 0.5) * cfg.dx
                y = cfg.y_min + (j + 0.5) * cfg.dx
                r = math.sqrt(x * x + y * y)
                if r > 0.6:
                    ratio = min(1.0, (r - 0.6) / 0.3)
                    sg[:, j, i] = 0.5 * ratio * ratio
        sg[solid == 1] = 0.0
        e.set_sponge(sg)
        o.set_sponge(sg)
    n = 57
    e.step(n)
    o.run(n)
    e.synchronize()
    fe, fo = e.debug_read_f(), o.read_f()
    fluid = solid == 0
    assert np.array_equal(fe[fluid], fo[fluid])
    po = o.read_probes()
    assert np.array_equal(e.read_probes(), po[:, 1:])
    assert np.abs(po).max() > 0


def test_checkpoint_restore_is_bitwise(gpu):
    cfg, solid, rho, pts = _small_scene()
    e = _engine(cfg, solid, rho, None, pts)
    e.step(11)
    blob = e.checkpoint()
    e.step(13)
    e.synchronize()
    p1, f1 = e.read_probes(), e.debug_read_f()
    r = _engine(cfg, solid, None, None, pts)
    r.restore(blob)
    r.step(13)
    r.synchronize()
    assert np.array_equal(r.read_probes(), p1)
    assert np.array_equal(r.debug_read_f(), f1, equal_nan=True)


def test_family_b_features_are_refused(gpu):
    cfg = fa.config(x=(-0.5, 0.5), y=(-0.5, 0.5), z=(-0.1, 0.1), dx=0.05)
    cfg32 = fa.config(x=(-0.5, 0.5), y=(-0.5, 0.5), z=(-0.1, 0.1), dx=0.05)
    cfg32.precision = nlos.F32
    with pytest.raises(nlos.EngineError) as ex:
        nlos.Engine(cfg32)
    assert ex.value.code == 7
    e = nlos.Engine(cfg)
    with pytest.raises(nlos.EngineError) as ex:
        e.set_source(nlos.default_source())
    assert ex.value.code == 7
    with pytest.raises(nlos.EngineError) as ex:
        nlos.Batch([cfg])
    assert ex.value.code == 7
    b = nlos.Engine(nlos.default_config(x_max=10.0, y_max=10.0, z_max=10.0, dx=1.0))
    with pytest.raises(nlos.EngineError) as ex:
        b.set_sponge(np.zeros((10, 10, 10)))
    assert ex.value.code == 7


def test_checkpoint_of_the_other_family_is_refused(gpu):
    """Both families store FP64 populations, but in different velocity orderings."""
    cfg, solid, rho, pts = _small_scene()
    a = _engine(cfg, solid, rho, None, pts)
    a.step(2)
    blob = a.checkpoint()
    cfg_b = fa.config(x=(-1.2, 1.2), y=(-1.0, 1.0), z=(-0.15, 0.15), dx=0.05)
    cfg_b.family = nlos.FAMILY_B
    cfg_b.stream_pattern = nlos.PATTERN_TWO_ARRAY
    cfg_b.ground = 0
    cfg_b.nscbc_faces = 0
    b = _engine(cfg_b, solid, None, None, pts)
    with pytest.raises(nlos.EngineError) as ex:
        b.restore(blob)
    assert ex.value.code == 1                    # NLOS_ERR_ARG


def test_line_source_of_04_bitwise_against_oracle(gpu):
    """04's scene at dx = 0.5 (cylinder, radial sponge, z periodic) driven by 04's line
    source, with 04's restart of the source time after a warm-up (04:369-388: t counts
    from 0 again for the timed loop): probes and populations equal the oracle's bit for bit,
    and every z layer stays identical."""
    cfg = fa.coarse04_config()
    solid, sigma = fa.cylinder_04(cfg), fa.radial_sponge_04(cfg)
    src = nlos.reference04_line_source(0)
    src.dt = 0.577350269 * cfg.dx / 1.0
    pts = [(0.0, 9.0, 0.0), (7.1, -2.3, 0.4), (-3.3, -11.1, -1.0)]
    print("[family A] 04 line source at dx=0.5: 30 + 30 steps", flush=True)
    e = _engine(cfg, solid, None, None, pts)
    e.set_sponge(sigma)
    e.set_line_source(src)
    o = OracleA(cfg)
    o.set_mask(solid)
    o.set_sponge(sigma)
    o.set_probes(pts)
    o.set_line_source(src)
    e.step(30)
    o.run(30)
    src.step0 = 30
    e.set_line_source(src)
    o.set_line_source(src)
    e.step(30, capture_field=True)
    o.run(30)
    e.synchronize()
    got, want = e.read_probes(), o.read_probes()[:, 1:]      # the oracle also keeps s = 0
    assert got.shape == (3, 60)
    assert np.abs(want).max() > 1e-6
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    f = e.debug_read_f()
    fo = o.read_f()
    fluid = solid == 0
    assert np.array_equal(f[fluid], fo[fluid])
    field = e.read_field()
    for k in range(1, field.shape[0]):
        assert np.array_equal(field[k], field[0]), k
    e.close()


def test_line_source_is_refused_in_family_b(gpu):
    cfg = nlos.default_config(x_max=16.0, y_max=16.0, z_max=8.0, precision="f64")
    e = nlos.Engine(cfg)
    with pytest.raises(nlos.EngineError):
        e.set_line_source(nlos.reference04_line_source())
    e.close()


@pytest.mark.parametrize("chunks", [[1], [2, 1, 17, 4], [24]])
def test_fused_step_equals_two_pass_step(gpu, chunks, monkeypatch):
    """nlos_step runs a family-A step as one pass (stream, then the next step's line-source
    add and collision of the same cell); NLOS_FUSED_A=0 runs 01's two passes.  Both give the
    same bits (probes, populations, captured field) on 04's scene with its line source,
    however the steps are split into calls, and across a restart of the source time."""
    cfg = fa.coarse04_config()
    solid, sigma = fa.cylinder_04(cfg), fa.radial_sponge_04(cfg)
    pts = [(0.0, 9.0, 0.0), (7.1, -2.3, 0.4), (-3.3, -11.1, -1.0)]
    print(f"[family A] fused step, chunks {chunks}", flush=True)
    runs = []
    for fused in ("0", "1"):
        monkeypatch.setenv("NLOS_FUSED_A", fused)
        src = nlos.reference04_line_source(0)
        src.dt = 0.577350269 * cfg.dx / 1.0
        e = _engine(cfg, solid, None, None, pts)
        e.set_sponge(sigma)
        e.set_line_source(src)
        for u, n in enumerate(chunks if fused == "1" else [sum(chunks)]):
            e.step(n)
        src.step0 = e.info().step
        e.set_line_source(src)
        e.step(5, capture_field=True)
        e.synchronize()
        runs.append((e.read_probes(), e.debug_read_f(), e.read_field()))
        e.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b, equal_nan=True)      # solid cells read as NaN
    assert np.abs(runs[0][0]).max() > 1e-9               # the line source has reached the probes
