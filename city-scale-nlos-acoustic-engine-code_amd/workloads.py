"""Seeded synthetic inputs: box cities as binary STL, moving sources, ground probes.

The city follows the moving-source city workload: axis-aligned boxes with footprints
uniform in [10, 40] m, heights lognormal(mu = ln 20, sigma = 0.5) capped at 200 m, added
until 35 % of the ground is covered, a rigid ground plane, sky height above the tallest
building.  Randomness comes from numpy's PCG64 generator with an explicit seed, so every
machine builds the same city.  The STL layout is the one GeometryEngine::load reads
(05_sensor_illusion.cpp:49-58, 70-111).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

import nlos_engine as nlos

Box = Tuple[float, float, float, float, float, float]   # x0, x1, y0, y1, z0, z1


def box_triangles(b: Box):
    x0, x1, y0, y1, z0, z1 = b
    p = [(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
         (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)]
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]
    out = []
    for a, b_, c, d in quads:
        out.append((p[a], p[b_], p[c]))
        out.append((p[a], p[c], p[d]))
    return out


def stl_bytes(boxes: Sequence[Box], header: bytes = b"seeded city") -> bytes:
    tris = [t for b in boxes for t in box_triangles(b)]
    buf = bytearray(header.ljust(80, b"\0")[:80])
    buf += struct.pack("<I", len(tris))
    rec = struct.Struct("<12fH")
    for (a, b, c) in tris:
        buf += rec.pack(0.0, 0.0, 0.0, *a, *b, *c, 0)
    return bytes(buf)


def seeded_city(width_m: float, depth_m: float, seed: int = 0, coverage: float = 0.35,
                fp_lo: float = 10.0, fp_hi: float = 40.0, h_mu: float = math.log(20.0),
                h_sigma: float = 0.5, h_cap: float = 200.0, margin_m: float = 0.0) -> List[Box]:
    """Boxes placed uniformly over [margin, width-margin] x [margin, depth-margin] until
    the covered ground fraction of the whole area reaches `coverage` (1 m raster)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    W, D = int(math.ceil(width_m)), int(math.ceil(depth_m))
    cover = np.zeros((D, W), bool)
    boxes: List[Box] = []
    target = coverage * W * D
    covered = 0
    while covered < target:
        fx, fy = rng.uniform(fp_lo, fp_hi, 2)
        x0 = rng.uniform(margin_m, max(margin_m, width_m - margin_m - fx))
        y0 = rng.uniform(margin_m, max(margin_m, depth_m - margin_m - fy))
        h = min(h_cap, float(np.exp(rng.normal(h_mu, h_sigma))))
        boxes.append((x0, x0 + fx, y0, y0 + fy, 0.0, h))
        i0, i1 = int(x0), min(W, int(math.ceil(x0 + fx)))
        j0, j1 = int(y0), min(D, int(math.ceil(y0 + fy)))
        before = cover[j0:j1, i0:i1].sum()
        cover[j0:j1, i0:i1] = True
        covered += int(cover[j0:j1, i0:i1].sum() - before)
    return boxes


@dataclass
class CityWorkload:
    """The moving-source city scene on an nx x ny x nz grid of spacing dx."""
    nx: int = 1024
    ny: int = 1024
    nz: int = 256
    dx: float = 1.0
    seed: int = 0
    sponge_cells: int = 32
    c_phys: float = 343.0
    tau: float = 0.5 + 1e-2
    freq: float = 200.0
    amp_rho: float = 1e-4          # injected mass per step at the waveform peak, in rho0
    height: float = 30.0
    speed: float = 20.0
    n_probes: int = 64
    boxes: List[Box] = field(default_factory=list)

    def config(self, precision="f16s") -> nlos.Config:
        return nlos.default_config(
            x_min=0.0, x_max=self.nx * self.dx, y_min=0.0, y_max=self.ny * self.dx,
            z_min=0.0, z_max=self.nz * self.dx, dx=self.dx, c_phys=self.c_phys, tau=self.tau,
            sponge_width=self.sponge_cells * self.dx, sponge_strength=0.1, sky_z=1e9,
            precision=precision, world_offset=(0.0, 0.0, 0.0))

    def build_boxes(self) -> List[Box]:
        if not self.boxes:
            self.boxes = seeded_city(self.nx * self.dx, self.ny * self.dx, seed=self.seed)
        return self.boxes

    def stl(self) -> bytes:
        return stl_bytes(self.build_boxes())

    def source(self) -> nlos.Source:
        dt = nlos.dt_of(self.config())
        x0, y0 = 0.30 * self.nx * self.dx, 0.45 * self.ny * self.dx
        x1 = 0.70 * self.nx * self.dx
        return nlos.default_source(sx=x0, sy=y0, sz=self.height, ex=x1, ey=y0, ez=self.height,
                                   speed=self.speed, amp=self.amp_rho / dt, freq=self.freq,
                                   waveform=nlos.WAVE_SINE, injection=nlos.INJECT_TRILINEAR)

    def probes(self, heights: np.ndarray) -> np.ndarray:
        """Ground-level probes: seeded columns within 120 m of the source start whose
        ground cell is fluid, at the centre of the first cell above ground."""
        rng = np.random.Generator(np.random.PCG64(self.seed + 1))
        x0, y0 = 0.30 * self.nx * self.dx, 0.45 * self.ny * self.dx
        pts = []
        while len(pts) < self.n_probes:
            r = 120.0 * math.sqrt(rng.uniform(0.0, 1.0))
            a = rng.uniform(0.0, 2.0 * math.pi)
            x, y = x0 + r * math.cos(a), y0 + r * math.sin(a)
            i, j = int(x / self.dx), int(y / self.dx)
            if not (0 <= i < self.nx and 0 <= j < self.ny):
                continue
            if heights[j + 1, i + 1] > 0.5 * self.dx:   # column (with halo offset) is built
                continue
            pts.append((x, y, 0.5 * self.dx))
        return np.array(pts)


@dataclass
class ScatterWorkload:
    """Rigid-cube scattering: n^3 cells at dx = 1, tau = 0.5 + 1e-4, a 24^3 solid cube at
    the centre ([52,76)^3 for n = 128), an initial Gaussian density pulse of amplitude
    1e-3 rho0 and sigma 4 cells centred 30 cells off the cube centre along +x, a 16-cell
    sponge ramping quadratically to 0.1 on all six faces, open faces with NSCBC on all six,
    no ground, 8 probes at seeded positions in the fluid outside the sponge."""
    n: int = 128
    seed: int = 0
    cube: int = 24
    sponge_cells: int = 16
    tau: float = 0.5 + 1e-4
    n_probes: int = 8

    def config(self, precision="f64") -> nlos.Config:
        n = float(self.n)
        return nlos.default_config(x_min=0.0, x_max=n, y_min=0.0, y_max=n, z_min=0.0, z_max=n,
                                   dx=1.0, c_phys=343.0, tau=self.tau,
                                   sponge_width=float(self.sponge_cells), sponge_strength=0.1,
                                   sky_z=1e9, ground=0, sponge_faces=nlos.FACES_ALL,
                                   nscbc_faces=nlos.FACES_ALL, precision=precision,
                                   world_offset=(0.0, 0.0, 0.0))

    def mask(self) -> np.ndarray:
        n, h = self.n, self.cube // 2
        m = np.zeros((n, n, n), np.uint8)
        m[n // 2 - h:n // 2 + h, n // 2 - h:n // 2 + h, n // 2 - h:n // 2 + h] = 1
        return m

    def density(self) -> np.ndarray:
        n = self.n
        c = np.arange(n) + 0.5
        z, y, x = np.meshgrid(c, c, c, indexing="ij")
        cx, cy, cz = n / 2 + 30.0, n / 2, n / 2
        r2 = (x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2
        return 1.0 + 1e-3 * np.exp(-r2 / (2.0 * 4.0 ** 2))

    def probes(self) -> np.ndarray:
        rng = np.random.Generator(np.random.PCG64(self.seed))
        n, s, h = self.n, self.sponge_cells, self.cube // 2
        pts = []
        while len(pts) < self.n_probes:
            p = rng.uniform(s, n - s, 3)
            if np.all(np.abs(p - n / 2) < h + 1):
                continue
            pts.append(tuple(p))
        return np.array(pts)


@dataclass
class BatchScene:
    cfg: nlos.Config
    source: nlos.Source
    probes: np.ndarray
    dims: Tuple[int, int, int]


@dataclass
class BatchWorkload:
    """Independent small scenes: dims uniform in [64,192] x [64,192] x [64,96] cells at
    dx = 1 m, a point source (Ricker, seeded centre frequency 40-120 Hz) at a seeded height
    of 2-20 m over a partially reflecting ground (bounce-back weight uniform 0.6-1.0), half
    of the sources moving at uniform 1-10 m/s in a seeded horizontal direction, 4 probes at
    1.5 m height, 16-cell sponge on the five open faces, no buildings."""
    n_scenes: int = 256
    seed: int = 0
    precision: str = "f32"
    tau: float = 0.5 + 1e-2
    lo: Tuple[int, int, int] = (64, 64, 64)
    hi: Tuple[int, int, int] = (192, 192, 96)
    sponge_cells: int = 16

    def scenes(self) -> List[BatchScene]:
        rng = np.random.Generator(np.random.PCG64(self.seed))
        out = []
        for s in range(self.n_scenes):
            nx, ny, nz = (int(rng.integers(self.lo[a], self.hi[a] + 1)) for a in range(3))
            cfg = nlos.default_config(x_min=0.0, x_max=float(nx), y_min=0.0, y_max=float(ny),
                                      z_min=0.0, z_max=float(nz), dx=1.0, c_phys=343.0, tau=self.tau,
                                      sponge_width=float(self.sponge_cells), sponge_strength=0.1,
                                      sky_z=1e9, wall_reflect=float(rng.uniform(0.6, 1.0)),
                                      precision=self.precision, world_offset=(0.0, 0.0, 0.0))
            h = float(rng.uniform(2.0, 20.0))
            x0 = float(rng.uniform(0.35, 0.65)) * nx
            y0 = float(rng.uniform(0.35, 0.65)) * ny
            moving = s % 2 == 1
            speed = float(rng.uniform(1.0, 10.0)) if moving else 0.0
            ang = float(rng.uniform(0.0, 2.0 * math.pi))
            x1, y1 = (x0 + 1000.0 * math.cos(ang), y0 + 1000.0 * math.sin(ang)) if moving else (x0, y0)
            f0 = float(rng.uniform(40.0, 120.0))
            src = nlos.default_source(sx=x0, sy=y0, sz=h, ex=x1, ey=y1, ez=h, speed=speed, amp=1e-2,
                                      freq=f0, t0=1.5 / f0, waveform=nlos.WAVE_RICKER)
            pr = []
            for _ in range(4):
                pr.append((float(rng.uniform(self.sponge_cells, nx - self.sponge_cells)),
                           float(rng.uniform(self.sponge_cells, ny - self.sponge_cells)), 1.5))
            out.append(BatchScene(cfg, src, np.array(pr), (nx, ny, nz)))
        return out


@dataclass
class CanyonWorkload:
    """Adjoint source localisation in an urban canyon: nx x ny x nz cells at dx = 1 m, two
    parallel rows of seeded buildings along x (footprint lengths uniform 10-40 m, depth
    uniform 15-30 m, heights lognormal(ln 20, 0.5) capped at nz - 20) on either side of a
    street of `street` cells, rigid ground, 16-cell sponge on five faces, open faces at the
    rest state, or with nscbc=True the reference's open boundary behind the sponge (NSCBC on
    x-, x+, y-, y+ and the top, 05:684-835, which the adjoint transposes; no building touches
    those faces).  The source is a static point source in the
    street, spread by a Gaussian stencil of sigma 2 cells (radius 3), driven by a sampled
    Ricker waveform (peak 100 Hz) of n_samples samples; 16 probes at 1.5 m in the street."""
    nx: int = 256
    ny: int = 256
    nz: int = 128
    street: int = 20
    seed: int = 0
    n_samples: int = 3000
    n_probes: int = 16
    tau: float = 0.5 + 1e-2
    nscbc: bool = False

    def config(self, precision="f32") -> nlos.Config:
        cfg = nlos.default_config(x_min=0.0, x_max=float(self.nx), y_min=0.0, y_max=float(self.ny),
                                  z_min=0.0, z_max=float(self.nz), dx=1.0, c_phys=343.0, tau=self.tau,
                                  sponge_width=16.0, sponge_strength=0.1, sky_z=1e9,
                                  stream_pattern="two_array", precision=precision,
                                  world_offset=(0.0, 0.0, 0.0))
        if not self.nscbc:
            cfg.nscbc_faces = 0
        return cfg

    def boxes(self) -> List[Box]:
        rng = np.random.Generator(np.random.PCG64(self.seed))
        yc = self.ny / 2.0
        out = []
        for side in (-1, 1):
            x = 16.0
            while x < self.nx - 16.0:
                length = float(rng.uniform(10.0, 40.0))
                depth = float(rng.uniform(15.0, 30.0))
                h = min(self.nz - 20.0, float(np.exp(rng.normal(math.log(20.0), 0.5))))
                x1 = min(x + length, self.nx - 16.0)
                if side < 0:
                    y0, y1 = yc - self.street / 2.0 - depth, yc - self.street / 2.0
                else:
                    y0, y1 = yc + self.street / 2.0, yc + self.street / 2.0 + depth
                out.append((x, x1, y0, y1, 0.0, h))
                x = x1 + float(rng.uniform(2.0, 8.0))
        return out

    def samples(self, f0=100.0, t0=0.015) -> np.ndarray:
        t = np.arange(self.n_samples) * nlos.dt_of(self.config())
        a = (np.pi * f0 * (t - t0)) ** 2
        return (1.0 - 2.0 * a) * np.exp(-a)

    def source(self, x=None, y=None, z=3.37, amp=1.0) -> nlos.Source:
        x = 0.41 * self.nx + 0.27 if x is None else x
        y = self.ny / 2.0 + 2.21 if y is None else y
        return nlos.default_source(sx=x, sy=y, sz=z, ex=x, ey=y, ez=z, speed=0.0, amp=amp,
                                   stencil_sigma_cells=2.0, stencil_radius=3,
                                   injection=nlos.INJECT_GAUSS5, waveform=nlos.WAVE_SAMPLED)

    def probes(self) -> np.ndarray:
        rng = np.random.Generator(np.random.PCG64(self.seed + 7))
        yc = self.ny / 2.0
        xs = rng.uniform(24.0, self.nx - 24.0, self.n_probes)
        ys = rng.uniform(yc - self.street / 2.0 + 1.0, yc + self.street / 2.0 - 1.0, self.n_probes)
        return np.stack([xs, ys, np.full(self.n_probes, 1.5)], 1)
