"""The legacy family-A step (01/04 step_pull semantics) on seeded scenes of random shape and
options: box and grid size, dx, tau, the default or a random per-cell sponge, random solid
columns and blocks, an initial density bump, 04's line source at a random place, strength
and frequency (sometimes none), random probes, runs in chunks with a field capture and a
restart of the source time.  Populations, probes and field equal the oracle's bit for bit,
with the fused pass and with the two passes."""
from __future__ import annotations

import math

import numpy as np
import pytest

import family_a as fa
import nlos_engine as nlos
from oracle_ffi import OracleA

pytestmark = pytest.mark.gpu


def _scene(seed):
    rng = np.random.default_rng(3000 + seed)
    dx = float(rng.choice([0.25, 0.5, 1.0]))
    hx, hy = float(rng.uniform(3.0, 12.0)), float(rng.uniform(3.0, 12.0))
    hz = float(rng.choice([0.5, 1.0, 3.0])) * dx * float(rng.integers(2, 8))
    cfg = fa.config(x=(-hx, hx), y=(-hy, hy), z=(-hz, hz), dx=dx, tau=float(rng.uniform(0.502, 0.8)),
                    sponge=float(rng.uniform(0.2, 3.0)), strength=float(rng.uniform(0.2, 1.0)))
    nx, ny, nz = fa.dims(cfg)
    solid = np.zeros((nz, ny, nx), np.uint8)
    for _ in range(int(rng.integers(0, 5))):
        i0, j0 = int(rng.integers(0, nx)), int(rng.integers(0, ny))
        w, d = int(rng.integers(1, max(2, nx // 4))), int(rng.integers(1, max(2, ny // 4)))
        if rng.random() < 0.6:
            solid[:, j0:j0 + d, i0:i0 + w] = 1               # a column through the periodic z
        else:
            k0 = int(rng.integers(0, nz))
            solid[k0:k0 + int(rng.integers(1, 3)), j0:j0 + d, i0:i0 + w] = 1
    rho = None
    if rng.random() < 0.5:
        cx, cy = float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy))
        z, y, x = np.meshgrid(cfg.z_min + (np.arange(nz) + 0.5) * dx, cfg.y_min + (np.arange(ny) + 0.5) * dx,
                              cfg.x_min + (np.arange(nx) + 0.5) * dx, indexing="ij")
        rho = 1.0 + 1e-3 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 2.0)
    sigma = None
    if rng.random() < 0.5:
        r0, ramp, s = float(rng.uniform(1.0, 0.8 * min(hx, hy))), float(rng.uniform(0.5, 3.0)), float(rng.uniform(0.1, 0.6))
        sigma = np.zeros((nz, ny, nx))
        for j in range(ny):
            for i in range(nx):
                xx = cfg.x_min + (i + 0.5) * dx
                yy = cfg.y_min + (j + 0.5) * dx
                r = math.sqrt(xx * xx + yy * yy)
                if r > r0:
                    sigma[:, j, i] = s * min(1.0, (r - r0) / ramp) ** 2
    line = None
    if rng.random() < 0.7:
        line = nlos.reference04_line_source(0)
        line.x0, line.y0 = float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy))
        line.amp = float(rng.uniform(1e-3, 2e-2))
        line.b = float(rng.uniform(0.5, 3.0))
        line.omega = float(rng.uniform(0.05, 0.5))
        line.r2_max = float(rng.uniform(1.0, 16.0))
        line.dt = 0.577350269 * dx
    pts = [(float(rng.uniform(-hx, hx)), float(rng.uniform(-hy, hy)), float(rng.uniform(-hz, hz)))
           for _ in range(int(rng.integers(1, 5)))]
    return cfg, solid, rho, sigma, line, pts


@pytest.mark.parametrize("fused", ["1", "0"])
@pytest.mark.parametrize("seed", list(range(12)))
def test_random_family_a_scene_equals_oracle(gpu, seed, fused, monkeypatch):
    monkeypatch.setenv("NLOS_FUSED_A", fused)
    cfg, solid, rho, sigma, line, pts = _scene(seed)
    print(f"[random family A {seed}] grid {fa.dims(cfg)}, dx {cfg.dx}, tau {cfg.tau:.3f}, "
          f"sponge {'radial' if sigma is not None else 'default'}, line source {line is not None}, "
          f"fused {fused}", flush=True)
    e = nlos.Engine(cfg)
    o = OracleA(cfg)
    for x in (e, o):
        x.set_mask(solid)
        if rho is not None:
            x.set_density(rho)
        if sigma is not None:
            x.set_sponge(sigma)
        if line is not None:
            x.set_line_source(line)
        x.set_probes(pts)
    e.step(17)
    e.step(14, capture_field=True)
    o.run(31)
    e.synchronize()
    fluid = solid == 0
    assert np.array_equal(e.debug_read_f()[fluid], o.read_f()[fluid])
    want = o.read_probes()[:, 1:]        # the oracle also keeps step 0
    assert np.array_equal(e.read_probes(), want)
    # the captured field is rho - 1, formed in FP64 and stored as FP32
    assert np.array_equal(e.read_field()[fluid], (o.read_rho()[fluid] - 1.0).astype(np.float32))
    if line is not None:                 # 04's timed loop counts t from 0 again
        line.step0 = 31
        e.set_line_source(line)
        o.set_line_source(line)
    e.step(9)
    o.run(9)
    e.synchronize()
    assert np.array_equal(e.debug_read_f()[fluid], o.read_f()[fluid])
    assert np.array_equal(e.read_probes(), o.read_probes()[:, 1:])
    e.close()
