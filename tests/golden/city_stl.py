"""Binary-STL writer for synthetic cities (boxes and pyramid-roofed blocks).

The format is the one `GeometryEngine::load` reads (05_sensor_illusion.cpp:49-58, 70-111):
an 80-byte header, a little-endian uint32 triangle count, then 50-byte records of
float32 normal[3], v1[3], v2[3], v3[3] and a uint16 attribute.

Coordinates are written in the STL frame.  The reference adds its `stl_sx/sy/sz` shift on
load (05:90-92), so a caller that wants a box at domain position `p` writes `p - shift`.
Integer-valued corners are exact in float32 (below 2**24), so the loaded geometry equals
the intended one bit for bit.
"""
from __future__ import annotations

import struct
from typing import Iterable, List, Sequence, Tuple

Vec = Tuple[float, float, float]
Tri = Tuple[Vec, Vec, Vec]


def box_triangles(x0: float, x1: float, y0: float, y1: float, z0: float, z1: float) -> List[Tri]:
    """Twelve triangles of an axis-aligned box, outward winding."""
    p = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    tris: List[Tri] = []
    for a, b, c, d in quads:
        tris.append((p[a], p[b], p[c]))
        tris.append((p[a], p[c], p[d]))
    return tris


def pyramid_block_triangles(x0: float, x1: float, y0: float, y1: float, z0: float,
                            wall: float, apex: float) -> List[Tri]:
    """Block with vertical walls up to `wall` and a four-sided pyramid roof up to `apex`."""
    cx, cy = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
    b = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)]
    t = [(x0, y0, wall), (x1, y0, wall), (x1, y1, wall), (x0, y1, wall)]
    top = (cx, cy, apex)
    tris: List[Tri] = [(b[0], b[2], b[1]), (b[0], b[3], b[2])]
    for i in range(4):
        j = (i + 1) % 4
        tris.append((b[i], b[j], t[j]))
        tris.append((b[i], t[j], t[i]))
        tris.append((t[i], t[j], top))
    return tris


def _normal(tri: Tri) -> Vec:
    (ax, ay, az), (bx, by, bz), (cx, cy, cz) = tri
    ux, uy, uz = bx - ax, by - ay, bz - az
    vx, vy, vz = cx - ax, cy - ay, cz - az
    nx, ny, nz = uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx
    n = (nx * nx + ny * ny + nz * nz) ** 0.5 or 1.0
    return (nx / n, ny / n, nz / n)


def stl_bytes(tris: Sequence[Tri], header: bytes = b"synthetic city") -> bytes:
    out = bytearray(header.ljust(80, b"\0")[:80])
    out += struct.pack("<I", len(tris))
    for tri in tris:
        out += struct.pack("<3f", *_normal(tri))
        for v in tri:
            out += struct.pack("<3f", *v)
        out += struct.pack("<H", 0)
    return bytes(out)


def shifted(tris: Iterable[Tri], sx: float, sy: float, sz: float) -> List[Tri]:
    return [tuple((v[0] - sx, v[1] - sy, v[2] - sz) for v in t) for t in tris]  # type: ignore


# ---------------------------------------------------------------------------------------
# The scene used for the compiled-reference golden traces.  05_sensor_illusion.cpp's main
# (05:1128-1170) fixes the domain x in [668,1064], y in [17,314], z in [0,106] at DX=1.5,
# the STL shift (-22248, -4123, -14), the source at (912, 66.6, 86) and probes B/C at
# (858, 220, 86) / (778, 190, 86).  The buildings below are given in domain coordinates:
# two between source and probes, one pyramid roof (sloped faces exercise the ray test's
# t computation), and three that cross the x-, x+ and y+ faces, one of them reaching the
# top plane, so solid cells sit on open boundaries.
REF_SHIFT = (-22248.0, -4123.0, -14.0)

REF_CITY_BOXES = [
    (870.0, 900.0, 120.0, 160.0, 0.0, 60.0),
    (790.0, 820.0, 150.0, 175.0, 0.0, 45.0),
    (650.0, 700.0, 230.0, 270.0, 0.0, 40.0),
    (950.0, 990.0, 290.0, 330.0, 0.0, 110.0),
    (1040.0, 1080.0, 100.0, 130.0, 0.0, 30.0),
]
REF_CITY_PYRAMIDS = [
    (700.0, 740.0, 40.0, 80.0, 0.0, 25.0, 40.0),
]


def reference_city_triangles() -> List[Tri]:
    tris: List[Tri] = []
    for b in REF_CITY_BOXES:
        tris += box_triangles(*b)
    for p in REF_CITY_PYRAMIDS:
        tris += pyramid_block_triangles(*p)
    # STL frame = domain - shift  (the reference adds the shift back on load)
    return shifted(tris, *REF_SHIFT)


def reference_city_stl() -> bytes:
    return stl_bytes(reference_city_triangles(), b"nlos reference-city golden scene")


if __name__ == "__main__":
    import sys
    path = sys.argv[1] if len(sys.argv) > 1 else "csx_model_binary.stl"
    with open(path, "wb") as fh:
        fh.write(reference_city_stl())
    print(path)


# ---------------------------------------------------------------------------------------
# The scene for 08_city_scale_uav.cpp's own run (08:1062-1092): domain x in [0,1650],
# y in [0,1150], z in [0,250] at DX=1.5, the same STL shift as 05, a 20 Hz source moving
# from (912, 66.6, 86) along +y at 30 m/s.  Buildings in domain coordinates: three within
# the first second's wave travel of the source (one next to its path, one pyramid roof),
# one crossing the y- face close to the source (tracked solid cells that the wave reaches),
# and blocks crossing the x-, x+ and y+ faces and the top plane.
REF08_CITY_BOXES = [
    (850.0, 890.0, 120.0, 170.0, 0.0, 70.0),
    (940.0, 990.0, 20.0, 60.0, 0.0, 120.0),
    (1000.0, 1060.0, -20.0, 30.0, 0.0, 40.0),
    (960.0, 1000.0, 300.0, 340.0, 0.0, 240.0),
    (-30.0, 40.0, 500.0, 560.0, 0.0, 60.0),
    (1600.0, 1680.0, 300.0, 360.0, 0.0, 90.0),
    (400.0, 460.0, 1100.0, 1200.0, 0.0, 300.0),
]
REF08_CITY_PYRAMIDS = [
    (700.0, 760.0, 200.0, 260.0, 0.0, 35.0, 55.0),
]


def reference08_city_stl() -> bytes:
    tris: List[Tri] = []
    for b in REF08_CITY_BOXES:
        tris += box_triangles(*b)
    for p in REF08_CITY_PYRAMIDS:
        tris += pyramid_block_triangles(*p)
    return stl_bytes(shifted(tris, *REF_SHIFT), b"nlos 08 city golden scene")
