"""Small seeded scenes shared by the oracle and GPU parity tests."""
from __future__ import annotations

import numpy as np

import nlos_engine as nlos


def box_scene(precision="f64", nx=40, ny=32, nz=24, seed=3):
    """Buildings that cross the x-, y+ and top faces, a moving sine source, six probes
    (one inside a solid), reference sponge/NSCBC face sets and ground."""
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(nx), y_min=0.0, y_max=float(ny),
                              z_min=0.0, z_max=float(nz), sponge_width=6.0, sponge_strength=0.35,
                              tau=0.55, sky_z=1e9, precision=precision)
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:8, 5:11, 0:6] = 1            # crosses x-
    solid[:, 26:32, 30:36] = 1          # crosses y+ and reaches the top plane
    solid[:12, 12:18, 16:22] = 1
    solid[:5, 2:6, 34:40] = 1           # crosses x+
    rng = np.random.default_rng(seed)
    for _ in range(6):
        i, j = rng.integers(0, nx - 3), rng.integers(0, ny - 3)
        solid[: rng.integers(1, 6), j:j + 3, i:i + 3] = 1
    src = nlos.default_source(sx=14.3, sy=7.7, sz=9.2, ex=25.1, ey=20.4, ez=11.0, speed=60.0,
                              amp=2e-2, freq=120.0, waveform=nlos.WAVE_SINE)
    probes = [(3.5, 20.5, 1.5), (37.2, 15.3, 7.7), (20.0, 16.0, 20.9), (8.5, 8.5, 3.5),
              (33.5, 29.5, 10.5), (1.2, 30.8, 22.5)]
    probes = [p for p in probes if p[0] < nx and p[1] < ny and p[2] < nz]
    return cfg, solid, src, probes


def pulse_scene(precision="f64", n=28, tau=0.5 + 1e-4):
    """Config-1 shape at small size: open bottom, sponge on all six faces, NSCBC on all six
    faces, a cube scatterer, an initial Gaussian density pulse and no source."""
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(n), y_min=0.0, y_max=float(n),
                              z_min=0.0, z_max=float(n), sponge_width=4.0, sponge_strength=0.1,
                              tau=tau, sky_z=1e9, ground=0, sponge_faces=nlos.FACES_ALL,
                              nscbc_faces=nlos.FACES_ALL, precision=precision)
    solid = np.zeros((n, n, n), np.uint8)
    a, b = n // 2 - 3, n // 2 + 3
    solid[a:b, a:b, a:b] = 1
    z, y, x = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 0.5, np.arange(n) + 0.5, indexing="ij")
    c = (n / 2 + 7.0, n / 2, n / 2)
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    rho = 1.0 + 1e-3 * np.exp(-r2 / (2 * 2.0 ** 2))
    probes = [(4.5, 4.5, 4.5), (n - 5.5, n / 2, n / 2), (n / 2, 5.5, n / 2), (n / 2, n / 2, n - 5.5)]
    return cfg, solid, rho, probes


def reflect_scene(precision="f64", nx=30, ny=26, nz=20):
    """Partially reflecting walls, trilinear injection of a Ricker wavelet, NSCBC off on
    one face, a sponge on four faces."""
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(nx), y_min=0.0, y_max=float(ny),
                              z_min=0.0, z_max=float(nz), sponge_width=5.0, sponge_strength=0.2,
                              tau=0.6, sky_z=1e9, wall_reflect=0.8,
                              sponge_faces=nlos.FACE_XM | nlos.FACE_XP | nlos.FACE_YM | nlos.FACE_ZP,
                              nscbc_faces=nlos.FACE_XM | nlos.FACE_XP | nlos.FACE_YP | nlos.FACE_ZP,
                              precision=precision)
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:7, 8:14, 10:16] = 1
    solid[:3, 0:4, 0:30] = 1
    src = nlos.default_source(sx=6.6, sy=16.2, sz=4.4, ex=6.6, ey=16.2, ez=4.4, amp=5e-2,
                              freq=90.0, t0=0.012, waveform=nlos.WAVE_RICKER,
                              injection=nlos.INJECT_TRILINEAR)
    probes = [(20.5, 20.5, 1.5), (25.0, 5.0, 4.0), (12.5, 11.5, 8.5)]
    return cfg, solid, src, probes


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def canyon_scene(precision="f64", nx=28, ny=24, nz=14, waveform="ricker", n_samples=0, open_faces=False):
    """Street canyon for the adjoint: two building rows along x with a 6-cell street, ground,
    sponge on five faces, open faces at the rest state (no NSCBC), partially reflecting walls,
    a static Gaussian-stencil source (sigma 2 cells, radius 3) in the street, four probes.
    open_faces: NSCBC on the reference's five faces and no sponge, so that the waves leave
    through the characteristic ghosts, and two more probes next to the x- and y+ faces."""
    cfg = nlos.default_config(dx=1.0, x_min=0.0, x_max=float(nx), y_min=0.0, y_max=float(ny),
                              z_min=0.0, z_max=float(nz), sponge_width=4.0, sponge_strength=0.2,
                              tau=0.6, sky_z=1e9, wall_reflect=0.9, nscbc_faces=0,
                              stream_pattern="two_array", precision=precision)
    if open_faces:
        cfg.nscbc_faces = nlos.FACES_REFERENCE
        cfg.sponge_faces = 0
    solid = np.zeros((nz, ny, nx), np.uint8)
    c = ny // 2
    solid[:7, c - 3 - 5:c - 3, 4:nx - 4] = 1
    solid[:9, c + 3:c + 3 + 5, 6:nx - 6] = 1
    kw = dict(sx=11.37, sy=c + 0.21, sz=3.63, ex=11.37, ey=c + 0.21, ez=3.63, speed=0.0, amp=1.0,
              stencil_sigma_cells=2.0, stencil_radius=3, injection=nlos.INJECT_GAUSS5)
    if waveform == "sampled":
        t = np.arange(n_samples) * nlos.dt_of(cfg)
        f0, t0 = 90.0, 0.012
        a = (np.pi * f0 * (t - t0)) ** 2
        samples = (1.0 - 2.0 * a) * np.exp(-a)
        src = nlos.default_source(waveform=nlos.WAVE_SAMPLED, **kw)
        src_samples = samples
    else:
        src = nlos.default_source(waveform=nlos.WAVE_RICKER, freq=90.0, t0=0.012, **kw)
        src_samples = None
    probes = [(3.5, c + 0.5, 1.5), (nx - 3.5, c - 1.5, 2.5), (15.5, c + 1.5, 10.5), (20.2, c - 0.4, 4.4)]
    if open_faces:
        probes += [(0.5, c + 0.5, 3.5), (13.5, ny - 0.5, 11.5)]
    return cfg, solid, src, src_samples, probes


def random_scene(seed, precision="f64"):
    """A seeded scene of random shape and options: grid 1-200 x 2-40 x 3-28 cells at a random
    origin and dx, random tau, sponge width/strength/faces, NSCBC faces, ground, sky height,
    wall reflection, solid boxes (some crossing faces), a moving source with a random
    waveform and injection, probes (some inside solids), and sometimes an initial density
    field.  Returns cfg, solid, src, samples, probes, rho, steps."""
    rng = np.random.default_rng(1000 + seed)
    nx = int(rng.choice([int(rng.integers(1, 5)), int(rng.integers(5, 73)), int(rng.integers(100, 201))]))
    ny = int(rng.integers(2, 41))
    nz = int(rng.integers(3, 29))
    dx = float(rng.choice([0.5, 1.0, 1.5]))
    x0, y0, z0 = (float(v) for v in rng.uniform(-20.0, 20.0, 3))
    faces = lambda: int(rng.integers(0, 64))
    cfg = nlos.default_config(dx=dx, x_min=x0, x_max=x0 + nx * dx, y_min=y0, y_max=y0 + ny * dx,
                              z_min=z0, z_max=z0 + nz * dx,
                              sponge_width=float(rng.uniform(0.0, 6.0)) * dx,
                              sponge_strength=float(rng.uniform(0.05, 0.4)),
                              tau=float(rng.uniform(0.5005, 0.9)),
                              sky_z=float(rng.choice([1e9, z0 + rng.uniform(0.3, 0.8) * nz * dx])),
                              wall_reflect=float(rng.choice([1.0, rng.uniform(0.5, 1.0)])),
                              sponge_faces=faces(), nscbc_faces=faces(), ground=int(rng.integers(0, 2)),
                              precision=precision)
    solid = np.zeros((nz, ny, nx), np.uint8)
    for _ in range(int(rng.integers(0, 7))):
        i0, j0 = int(rng.integers(-2, nx)), int(rng.integers(-2, ny))
        w, d, h = (int(v) for v in rng.integers(1, 9, 3))
        k0 = int(rng.choice([0, rng.integers(0, nz)]))
        solid[k0:k0 + h, max(j0, 0):j0 + d, max(i0, 0):i0 + w] = 1
    inside = lambda: (x0 + rng.uniform(0.05, 0.95) * nx * dx, y0 + rng.uniform(0.05, 0.95) * ny * dx,
                      z0 + rng.uniform(0.05, 0.95) * nz * dx)
    s, e = inside(), inside()
    waveform = int(rng.choice([nlos.WAVE_GAUSS_PULSE, nlos.WAVE_SINE, nlos.WAVE_SAMPLED, nlos.WAVE_RICKER]))
    dt = dx / (np.sqrt(3.0) * 343.0)
    steps = int(rng.integers(25, 51))
    src = nlos.default_source(sx=s[0], sy=s[1], sz=s[2], ex=e[0], ey=e[1], ez=e[2],
                              speed=float(rng.uniform(0.0, 400.0)), amp=float(rng.uniform(5e-3, 5e-2)),
                              freq=float(rng.uniform(50.0, 400.0)), t0=float(rng.uniform(0.0, 20.0)) * dt,
                              sigma_t=float(rng.uniform(3.0, 10.0)) * dt, waveform=waveform,
                              injection=int(rng.integers(0, 2)))
    samples = rng.standard_normal(steps + 8) if waveform == nlos.WAVE_SAMPLED else None
    probes = [inside() for _ in range(int(rng.integers(1, 7)))]
    rho = None
    if rng.random() < 0.3:
        rho = 1.0 + 1e-3 * rng.random((nz, ny, nx))
    return cfg, solid, src, samples, probes, rho, steps
