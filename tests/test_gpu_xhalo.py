"""x-halo records (two-array pattern, every storage mode).

The halo cells i = 0 and i = nx + 1 sit alone in their cache lines of the lattice, so the
engine keeps the five slots that interior cells pull from each of them in compact per-row
records: the ghost kernel writes them; the sweep and the solid kernel read them.  With
NLOS_XHALO=0 the same values go through the lattice instead.  Both ways must give the same
bits: probes, populations and the captured field.  The scenes cover:
* NSCBC x faces;
* buildings crossing both x faces (tracked solid cells at i = 1 and i = nx);
* partially reflecting walls;
* rows shorter than one wave and rows of whole 64-cell waves;
* grids one and two cells across (one lane reads both faces' records).
"""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu


def _thin(precision, nx):
    """nx = 1 or 2 interior cells across: one lane can read both faces' records."""
    ny, nz = 10, 8
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(nx), y_min=0.0, y_max=float(ny),
                              z_min=0.0, z_max=float(nz), sponge_width=2.0, sponge_strength=0.1,
                              tau=0.55, sky_z=1e9, ground=0, sponge_faces=nlos.FACES_ALL,
                              nscbc_faces=nlos.FACES_ALL, precision=precision)
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[2:4, 6:8, :] = 1
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    rho = 1.0 + 1e-3 * np.exp(-((y - 3.0) ** 2 + (z - 4.0) ** 2 + (x - 0.3) ** 2) / 4.0)
    return cfg, solid, None, rho, [(nx / 2, 2.5, 4.5), (nx / 2, 8.5, 1.5)]


def _scene(name, precision):
    if name.startswith("thin"):
        return _thin(precision, int(name[4:]))
    if name == "box":
        cfg, solid, src, probes = scenes.box_scene(precision)
        return cfg, solid, src, None, probes
    if name == "box128":
        cfg, solid, src, probes = scenes.box_scene(precision, nx=128, ny=24, nz=16)
        return cfg, solid, src, None, probes
    if name == "pulse":
        cfg, solid, rho, probes = scenes.pulse_scene(precision)
        return cfg, solid, None, rho, probes
    cfg, solid, src, probes = scenes.reflect_scene(precision)
    return cfg, solid, src, None, probes


def _run(cfg, solid, src, rho, probes, steps):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if src is not None:
        e.set_source(src)
    e.set_probes(probes)
    e.step(steps, capture_field=True)
    e.synchronize()
    out = (e.read_probes(), e.debug_read_f(), e.read_field(), e.info().device_bytes)
    e.close()
    return out


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "b16"])
@pytest.mark.parametrize("scene", ["box", "box128", "pulse", "reflect", "thin1", "thin2"])
def test_xhalo_records_equal_lattice_halo(gpu, precision, scene, monkeypatch):
    print(f"[x-halo records] scene={scene} precision={precision}", flush=True)
    cfg, solid, src, rho, probes = _scene(scene, precision)
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NLOS_XHALO", flag)
        runs.append(_run(cfg, solid, src, rho, probes, 41))
    rec, lat = runs
    assert rec[3] > lat[3]                  # the records were allocated and used
    for a, b in zip(rec[:3], lat[:3]):
        assert np.array_equal(a, b, equal_nan=True)
    assert np.abs(rec[0]).max() > 0
