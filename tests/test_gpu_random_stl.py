"""Seeded random meshes through the reference's column top-hit voxeliser (05:113-192):
boxes, pyramid roofs and free-standing tilted triangles, some with vertices on the cell-
centre lattice (ties decided by EPS and the candidate order), at a random scale, shift,
domain origin and dx.  Ground heights and the mask equal the oracle's bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def _mesh(rng, x0, y0, w, d, dx):
    import city_stl
    tris = []
    for _ in range(int(rng.integers(3, 40))):
        kind = rng.random()
        ax, ay = x0 + rng.uniform(-0.1, 1.0) * w, y0 + rng.uniform(-0.1, 1.0) * d
        bw, bd = rng.uniform(0.5, 0.3 * w), rng.uniform(0.5, 0.3 * d)
        if rng.random() < 0.3:   # corners on the cell-centre lattice
            ax = x0 + (np.floor((ax - x0) / dx) + 0.5) * dx
            ay = y0 + (np.floor((ay - y0) / dx) + 0.5) * dx
            bw = max(dx, np.round(bw / dx) * dx)
            bd = max(dx, np.round(bd / dx) * dx)
        h = rng.uniform(0.5, 12.0)
        if kind < 0.4:
            tris += city_stl.box_triangles(ax, ax + bw, ay, ay + bd, 0.0, h)
        elif kind < 0.7:
            tris += city_stl.pyramid_block_triangles(ax, ax + bw, ay, ay + bd, 0.0, h, h + rng.uniform(0.5, 6.0))
        else:
            p = [(ax + rng.uniform(-2, 2) * dx, ay + rng.uniform(-2, 2) * dx, rng.uniform(0.0, 15.0)) for _ in range(3)]
            tris.append(tuple(p))
    return tris


@pytest.mark.parametrize("seed", list(range(12)))
def test_random_mesh_heights_and_mask_equal_oracle(gpu, seed):
    import city_stl
    rng = np.random.default_rng(500 + seed)
    dx = float(rng.choice([0.5, 1.0, 1.5, 2.0]))
    nx, ny, nz = (int(v) for v in rng.integers(8, 70, 3))
    x0, y0 = (float(v) for v in rng.uniform(-30.0, 30.0, 2))
    z0 = float(rng.choice([0.0, -dx, rng.uniform(-3.0, 1.0)]))
    cfg = nlos.default_config(dx=dx, x_min=x0, x_max=x0 + nx * dx, y_min=y0, y_max=y0 + ny * dx,
                              z_min=z0, z_max=z0 + nz * dx, sky_z=float(rng.choice([1e9, z0 + 0.6 * nz * dx])))
    scale = float(rng.choice([1.0, 0.5, 2.0]))
    sx, sy, sz = float(rng.uniform(-5, 5)), float(rng.uniform(-5, 5)), float(rng.uniform(-1, 1))
    # vertices map to v * scale + shift (05:70-111): the mesh spans the domain after the map
    tris = _mesh(rng, (x0 - sx) / scale, (y0 - sy) / scale, nx * dx / scale, ny * dx / scale, dx / scale)
    stl = city_stl.stl_bytes(tris)
    print(f"[random mesh {seed}] {len(tris)} triangles, grid {nx}x{ny}x{nz}, dx {dx}, scale {scale}", flush=True)
    e = nlos.Engine(cfg)
    e.voxelize_stl(stl, scale, sx, sy, sz)
    o = Oracle(cfg)
    assert o.load_stl(stl, scale, sx, sy, sz) == 0
    assert np.array_equal(e.ground_heights(), o.ground_heights())
    assert np.array_equal(e.get_mask(), o.get_mask())
    e.close()
