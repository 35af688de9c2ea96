"""x-halo records (two-array pattern, every storage mode).

The halo cells i = 0 and i = nx + 1 sit alone in their cache lines of the lattice, so the
engine keeps the five slots that interior cells pull from each of them in compact per-row
records: the ghost kernel writes them; the sweep and the solid kernel read them.  With
NLOS_XHALO=0 the same values go through the lattice instead.  Both ways must give the same
bits: probes, populations and the captured field.  The scenes cover:
* NSCBC x faces;
* buildings crossing both x faces (tracked solid cells at i = 1 and i = nx);
* partially reflecting walls;
* rows shorter than one wave and rows of whole 64-cell waves.
"""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu


def _scene(name, precision):
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
@pytest.mark.parametrize("scene", ["box", "box128", "pulse", "reflect"])
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
