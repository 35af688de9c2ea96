"""Crossing voxelisers (NLOS_VOXEL_PARITY / WINDING, optionally on three axes) and the
chunked STL file reader.

The reference's height-field voxeliser is pinned bit for bit in test_gpu_parity.py.  The
crossing modes have no reference counterpart; they are pinned here against analytic
containment of cell centres in axis-aligned boxes whose faces sit a quarter cell off the
centre lattice (so no centre is on a face), with overlapping shells, a floating box
(an overhang), an inward-oriented cavity and a box with a missing face.
"""
from __future__ import annotations

import struct

import numpy as np
import pytest

import nlos_engine as nlos
import workloads

pytestmark = pytest.mark.gpu

NX, NY, NZ = 48, 40, 32

QUADS = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
BOTTOM = 0


def _box_tris(b, inward=False, skip=()):
    x0, x1, y0, y1, z0, z1 = b
    p = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    out = []
    for n, (a, b_, c, d) in enumerate(QUADS):
        if n in skip:
            continue
        for tri in ((p[a], p[b_], p[c]), (p[a], p[c], p[d])):
            out.append(tri[::-1] if inward else tri)
    return out


def _stl(tris):
    buf = bytearray(b"voxel modes".ljust(80, b"\0"))
    buf += struct.pack("<I", len(tris))
    rec = struct.Struct("<12fH")
    for a, b, c in tris:
        buf += rec.pack(0.0, 0.0, 0.0, *a, *b, *c, 0)
    return bytes(buf)


def _inside(b):
    x0, x1, y0, y1, z0, z1 = b
    z, y, x = np.meshgrid(np.arange(NZ) + 0.5, np.arange(NY) + 0.5, np.arange(NX) + 0.5, indexing="ij")
    return (x > x0) & (x < x1) & (y > y0) & (y < y1) & (z > z0) & (z < z1)


def _mask(stl, voxelizer):
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(NX), y_min=0.0, y_max=float(NY), z_min=0.0,
                              z_max=float(NZ), sky_z=1e9, precision="f32")
    cfg.voxelizer = voxelizer
    e = nlos.Engine(cfg)
    e.voxelize_stl(stl)
    m = e.get_mask().reshape(NZ, NY, NX).astype(bool)
    h = e.ground_heights()
    e.close()
    return m, h


# faces on quarter-cell offsets (k + 0.25 or k + 0.75)
OVERLAP = [(4.25, 20.75, 5.25, 18.75, 0.0, 12.25), (14.25, 30.75, 10.25, 25.75, 0.0, 20.75)]
FLOATING = (32.25, 44.75, 4.25, 15.75, 10.25, 16.75)
ROOM_OUT = (6.25, 24.75, 26.25, 38.75, 0.0, 18.25)
ROOM_IN = (9.25, 21.75, 29.25, 35.75, 2.25, 14.75)


def _scene():
    tris = [t for b in OVERLAP + [FLOATING, ROOM_OUT] for t in _box_tris(b)]
    tris += _box_tris(ROOM_IN, inward=True)
    return _stl(tris)


def test_winding_mode_is_union_minus_cavities(gpu):
    m, _ = _mask(_scene(), nlos.VOXEL_WINDING)
    want = _inside(OVERLAP[0]) | _inside(OVERLAP[1]) | _inside(FLOATING) | (_inside(ROOM_OUT) & ~_inside(ROOM_IN))
    assert m.sum() > 0
    assert np.array_equal(m, want)


def test_parity_mode_is_odd_containment(gpu):
    m, _ = _mask(_scene(), nlos.VOXEL_PARITY)
    count = sum(_inside(b).astype(int) for b in OVERLAP + [FLOATING, ROOM_OUT, ROOM_IN])
    assert np.array_equal(m, count % 2 == 1)


def test_height_mode_fills_under_overhangs_crossing_modes_do_not(gpu):
    stl = _stl(_box_tris(FLOATING))
    mh, hh = _mask(stl, nlos.VOXEL_HEIGHT)
    mw, hw = _mask(stl, nlos.VOXEL_WINDING)
    box = _inside(FLOATING)
    under = _inside(FLOATING[:4] + (0.0, FLOATING[5])) & ~box
    assert np.array_equal(mw, box)
    assert mh[under].all() and mh[box].all()
    assert np.array_equal(hh, hw)          # column heights are the reference's in every mode


def test_three_axis_vote_closes_a_missing_face(gpu):
    b = (10.25, 30.75, 8.25, 28.75, 4.25, 20.75)
    stl = _stl(_box_tris(b, skip=(BOTTOM,)))
    single, _ = _mask(stl, nlos.VOXEL_PARITY)
    voted, _ = _mask(stl, nlos.VOXEL_PARITY | nlos.VOXEL_3AXIS)
    assert not np.array_equal(single, _inside(b))     # z rays alone miss the open bottom
    assert np.array_equal(voted, _inside(b))


def test_crossing_modes_match_height_mode_on_a_city_of_boxes_on_the_ground(gpu):
    """Boxes standing on the ground have no overhangs: every mode gives the same mask."""
    boxes = workloads.seeded_city(NX, NY, seed=11, fp_lo=4.0, fp_hi=12.0, h_mu=np.log(8.0))
    stl = workloads.stl_bytes(boxes)
    mh, _ = _mask(stl, nlos.VOXEL_HEIGHT)
    for mode in (nlos.VOXEL_WINDING, nlos.VOXEL_WINDING | nlos.VOXEL_3AXIS):
        m, _ = _mask(stl, mode)
        assert np.array_equal(m, mh), mode


def test_bad_voxelizer_is_refused(gpu):
    with pytest.raises(nlos.EngineError):
        _mask(_stl(_box_tris(FLOATING)), nlos.VOXEL_3AXIS)     # height field has no 3-axis form


def test_stl_file_reader_matches_in_memory_parse(gpu, tmp_path):
    """nlos_load_stl reads the file in chunks of 65536 records; 6000 boxes (72000 triangles)
    cross a chunk boundary.  A truncated file gives empty geometry and NLOS_ERR_IO."""
    rng = np.random.default_rng(2)
    boxes = []
    for _ in range(6000):
        x0, y0 = rng.uniform(0, NX - 2), rng.uniform(0, NY - 2)
        boxes.append((x0, x0 + rng.uniform(0.3, 2), y0, y0 + rng.uniform(0.3, 2), 0.0, rng.uniform(1, 20)))
    stl = workloads.stl_bytes(boxes)
    path = tmp_path / "city.stl"
    path.write_bytes(stl)
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(NX), y_min=0.0, y_max=float(NY), z_min=0.0,
                              z_max=float(NZ), sky_z=1e9, precision="f32")
    a = nlos.Engine(cfg)
    a.voxelize_stl(stl)
    b = nlos.Engine(cfg)
    assert b.load_stl(str(path)) == 0
    assert np.array_equal(a.ground_heights(), b.ground_heights())
    assert np.array_equal(a.get_mask(), b.get_mask())
    assert a.get_mask().sum() > 0
    (tmp_path / "cut.stl").write_bytes(stl[:-25])
    c = nlos.Engine(cfg)
    assert c.load_stl(str(tmp_path / "cut.stl")) == 5    # NLOS_ERR_IO, empty geometry kept
    assert c.get_mask().sum() == 0
