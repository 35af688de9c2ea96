"""Pins of the oracle (oracle/ref_cpu.cpp) before it is trusted as the engine's checker.

1. Against the compiled, unmodified reference program 05_sensor_illusion (fixtures in
   tests/golden/ref05_*, written by tests/golden/make_ref_goldens.py): probe files byte for
   byte and VTK field samples, on flat ground and on a city whose buildings cross three
   open faces and the top plane.
2. Closed forms and invariants for what those runs leave open: the sine waveform and the
   moving source (08_city_scale_uav.cpp:543, 05:221-238), the source normalisation
   (05:631), the trilinear and Ricker/sampled extensions, mass conservation in a closed
   box, mirror symmetry (pins the direction/inverse tables), the sponge ramp at known
   distances, and the voxeliser on analytic boxes.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

import city_stl
import nlos_engine as nlos
from oracle_ffi import Oracle

# Sums of injected populations are differences of values near w_q, so they carry
# cancellation of order 1e-16/|injection|; the source tolerances below allow for it.
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1])


@pytest.mark.parametrize("scene", ["city", "flat"])
def test_oracle_reproduces_compiled_reference(scene):
    o = Oracle(nlos.reference05_config())
    assert o.dims == (264, 198, 71)
    if scene == "city":
        assert o.load_stl(city_stl.reference_city_stl(), 1.0, *city_stl.REF_SHIFT) == 0
    names, pts = nlos.reference05_probes()
    o.set_probes(pts)
    o.set_source(nlos.reference05_source())
    frames = json.load(open(os.path.join(GOLDEN, f"ref05_{scene}", "vtk_frames.json")))
    nx, ny, nz = frames["grid"]
    done = 0
    for fr in ("50", "98"):
        s = int(fr) * 4
        o.run(s + 1 - done)
        done = s + 1
        field = o.read_field()
        want = frames["frames"][fr]
        st = want["stride"]
        for k, exp in want["slices"].items():
            got = ["%.6e" % float(v) for v in field[int(k), ::st, ::st].ravel()]
            assert got == exp, f"frame {fr} slice {k}"
    o.run(394 - done)
    probes = o.read_probes()
    assert probes.shape == (4, 394)
    for n, row in zip(names, probes):
        text = "".join("%.8e\n" % v for v in row)
        assert text == open(os.path.join(GOLDEN, f"ref05_{scene}", f"probe_{n}.txt")).read(), n


def _small_cfg(**kw):
    base = dict(dx=1.0, x_min=0.0, x_max=16.0, y_min=0.0, y_max=14.0, z_min=0.0, z_max=12.0,
                sponge_width=3.0, tau=0.6, sky_z=1e9)
    base.update(kw)
    return nlos.default_config(**base)


def _injected(o, t):
    before = o.read_f()
    o.apply_source(t)
    return o.read_f() - before


@pytest.mark.parametrize("t", [0.0131, 0.0522, 0.3])
def test_sine_source_normalisation_and_trajectory(t):
    """Sum of injected populations = amp*sin(2 pi f t)*dt (05:631 with 08:543); the stencil
    is the 5^3 block around the cell of the linearly moving source (05:221-238, 591-603)."""
    o = Oracle(_small_cfg())
    src = nlos.default_source(sx=3.3, sy=4.1, sz=5.2, ex=12.7, ey=9.9, ez=6.1, speed=40.0,
                              amp=0.7, freq=37.0, waveform=nlos.WAVE_SINE)
    o.set_source(src)
    d = _injected(o, t)
    val = 0.7 * math.sin(2 * math.pi * 37.0 * t)
    assert math.isclose(d.sum(), val * o.dt, rel_tol=1e-8)
    # per-cell distribution proportional to w_q
    cell = d.sum(axis=3)
    nz_cells = np.argwhere(np.abs(cell) > 0)
    for (k, j, i) in nz_cells[:5]:
        np.testing.assert_allclose(d[k, j, i], W * cell[k, j, i], rtol=1e-7)
    dist = math.dist((3.3, 4.1, 5.2), (12.7, 9.9, 6.1))
    r = min(1.0, t / (dist / 40.0))
    pos = (3.3 + r * 9.4, 4.1 + r * 5.8, 5.2 + r * 0.9)
    centre = [math.floor(p) + 1 for p in pos]          # +1: halo offset of read_f
    ks, js, is_ = nz_cells.T
    assert (is_.min(), is_.max()) == (centre[0] - 2, centre[0] + 2)
    assert (js.min(), js.max()) == (centre[1] - 2, centre[1] + 2)
    assert (ks.min(), ks.max()) == (centre[2] - 2, centre[2] + 2)


def test_gaussian_pulse_weights_closed_form():
    o = Oracle(_small_cfg())
    src = nlos.default_source(sx=8.2, sy=7.1, sz=6.6, ex=8.2, ey=7.1, ez=6.6, amp=3.0)
    o.set_source(src)
    t = 0.047
    d = _injected(o, t).sum(axis=3)
    val = 3.0 * math.exp(-0.5 * ((t - 0.05) / 0.005) ** 2)
    assert math.isclose(d.sum(), val * o.dt, rel_tol=1e-8)
    # weights exp(-r^2/(2 (1.25 dx)^2)) relative to the centre cell
    i, j, k = 8 + 1, 7 + 1, 6 + 1
    for di, dj, dk in [(1, 0, 0), (-2, 1, 0), (1, 1, -2)]:
        c0 = np.array([8.5, 7.5, 6.5]) - (8.2, 7.1, 6.6)
        c1 = c0 + (di, dj, dk)
        ratio = math.exp(-0.5 * (c1 @ c1 - c0 @ c0) / 1.25 ** 2)
        assert math.isclose(d[k + dk, j + dj, i + di] / d[k, j, i], ratio, rel_tol=1e-8)


def test_solid_cells_receive_no_source_and_normalisation_is_over_fluid():
    o = Oracle(_small_cfg())
    m = np.zeros((12, 14, 16), np.uint8)
    m[:7, 5:9, 6:10] = 1
    o.set_mask(m)
    o.set_source(nlos.default_source(sx=8.0, sy=8.0, sz=7.5, ex=8.0, ey=8.0, ez=7.5, amp=1.0))
    d = _injected(o, 0.05).sum(axis=3)[1:-1, 1:-1, 1:-1]
    assert not np.any(d[m == 1])
    assert math.isclose(d.sum(), 1.0 * o.dt, rel_tol=1e-8)


def test_trilinear_injection_weights():
    o = Oracle(_small_cfg())
    src = nlos.default_source(sx=5.8, sy=6.25, sz=4.5, ex=5.8, ey=6.25, ez=4.5, amp=2.0, freq=50.0,
                              waveform=nlos.WAVE_SINE, injection=nlos.INJECT_TRILINEAR)
    o.set_source(src)
    t = 0.004
    d = _injected(o, t).sum(axis=3)
    assert math.isclose(d.sum(), 2.0 * math.sin(2 * math.pi * 50 * t) * o.dt, rel_tol=1e-8)
    nzc = np.argwhere(np.abs(d) > 0)
    assert len(nzc) == 4          # z sits on a cell centre: tz = 0
    tx, ty = 5.8 - 5.5, 6.25 - 5.5
    want = {(5, 5): (1 - tx) * (1 - ty), (6, 5): tx * (1 - ty), (5, 6): (1 - tx) * ty, (6, 6): tx * ty}
    for (k, j, i) in nzc:
        assert k == 4 + 1
        assert math.isclose(d[k, j, i] / d.sum(), want[(i - 1, j - 1)], rel_tol=1e-8)


def test_ricker_and_sampled_waveforms():
    o = Oracle(_small_cfg())
    src = nlos.default_source(sx=8, sy=7, sz=6, ex=8, ey=7, ez=6, amp=1.5, freq=80.0, t0=0.01,
                              waveform=nlos.WAVE_RICKER)
    o.set_source(src)
    t = 0.0123
    a = (math.pi * 80.0 * (t - 0.01)) ** 2
    assert math.isclose(_injected(o, t).sum(), 1.5 * (1 - 2 * a) * math.exp(-a) * o.dt, rel_tol=1e-8)
    o2 = Oracle(_small_cfg())
    samples = np.linspace(-1, 1, 9)
    o2.set_source(nlos.default_source(sx=8, sy=7, sz=6, ex=8, ey=7, ez=6, amp=2.0,
                                      waveform=nlos.WAVE_SAMPLED), samples)
    assert math.isclose(_injected(o2, 3 * o2.dt).sum(), 2.0 * samples[3] * o2.dt, rel_tol=1e-8)


def _closed_box(n=12, tau=0.6):
    cfg = _small_cfg(x_max=float(n), y_max=float(n), z_max=float(n), sponge_faces=0,
                     nscbc_faces=0, tau=tau)
    m = np.ones((n, n, n), np.uint8)
    m[:-1, 1:-1, 1:-1] = 0          # walls on four sides and the top layer; ground below
    m[-1] = 1
    return cfg, m


def test_mass_conservation_in_closed_box():
    cfg, m = _closed_box()
    o = Oracle(cfg)
    o.set_mask(m)
    n = m.shape[0]
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    rho = 1.0 + 1e-3 * np.exp(-((x - 5.1) ** 2 + (y - 6.3) ** 2 + (z - 4.2) ** 2) / 4.0)
    o.set_density(rho)
    fluid = np.pad(m == 0, 1)
    m0 = o.read_f()[fluid].sum()
    o.run(60)
    m1 = o.read_f()[fluid].sum()
    assert abs(m1 - m0) < 1e-12 * m0
    assert np.abs(o.read_rho()[m == 0] - 1).max() > 1e-7      # something moved


def test_mirror_symmetry_pins_direction_tables():
    """A pulse symmetric under x -> -x in a symmetric closed box stays symmetric; a
    mismatched direction/inverse table breaks it."""
    cfg, m = _closed_box(n=14)
    o = Oracle(cfg)
    o.set_mask(m)
    n = 14
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    rho = 1.0 + 1e-3 * np.exp(-((x - 7.0) ** 2 + (y - 5.2) ** 2 + (z - 4.6) ** 2) / 3.0)
    o.set_density(rho)
    o.run(40)
    r = o.read_rho() - 1.0
    assert np.abs(r).max() > 1e-6
    np.testing.assert_allclose(r, r[:, :, ::-1], atol=1e-13, rtol=0)
    f = o.read_f()[1:-1, 1:-1, 1:-1]
    mirror = [int(np.flatnonzero((EX == -EX[q]) & (EY == EY[q]) & (EZ == EZ[q]))[0]) for q in range(19)]
    np.testing.assert_allclose(f, f[:, :, ::-1][..., mirror], atol=1e-13, rtol=0)


def test_rest_state_without_source_stays_at_rest():
    o = Oracle(_small_cfg(sponge_faces=nlos.FACES_ALL, ground=1))
    m = np.zeros((12, 14, 16), np.uint8)
    m[:5, 3:7, 4:9] = 1
    o.set_mask(m)
    o.set_probes([(2.5, 2.5, 2.5), (10.0, 10.0, 9.0)])
    o.run(30)
    assert np.abs(o.read_probes()).max() < 1e-14
    f = o.read_f()[1:-1, 1:-1, 1:-1][m == 0]
    np.testing.assert_allclose(f, np.broadcast_to(W, f.shape), rtol=1e-14)


def test_sponge_ramp_at_known_distances():
    cfg = _small_cfg(x_max=40.0, y_max=30.0, z_max=20.0, sponge_width=8.0, sponge_strength=0.25,
                     sponge_faces=nlos.FACES_REFERENCE)
    s = Oracle(cfg).sponge()
    # cell centre (i+0.5) from x-: d = 0.5 -> 0.25*(7.5/8)^2; centre at 15.5/14.5 has d>=8 -> 0
    assert s[10, 15, 0] == 0.25 * (7.5 / 8.0) * (7.5 / 8.0)
    assert s[10, 15, 20] == 0.0
    assert s[0, 15, 20] == 0.0                     # no sponge on the ground face
    assert s[19, 15, 20] == 0.25 * (7.5 / 8.0) * (7.5 / 8.0)
    assert s[10, 15, 39] == s[10, 15, 0]
    assert np.all(np.diff(s[10, 15, :20]) <= 0)


def test_voxelizer_analytic_boxes():
    boxes = [(2.3, 7.6, 1.2, 5.9, 0.0, 4.4), (9.1, 13.7, 6.3, 12.8, 0.0, 7.7)]
    tris = [t for b in boxes for t in city_stl.box_triangles(*b)]
    o = Oracle(_small_cfg())
    assert o.load_stl(city_stl.stl_bytes(tris)) == 0
    m = o.get_mask()
    nz, ny, nx = m.shape
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    want = np.zeros_like(m, dtype=bool)
    for (x0, x1, y0, y1, z0, z1) in boxes:
        want |= (x > x0) & (x < x1) & (y > y0) & (y < y1) & (z < z1)
    assert np.array_equal(m.astype(bool), want)
    h = o.ground_heights()[1:-1, 1:-1]
    # heights are the float32 STL tops (05:90-92), hit from z = 1e4, so they carry the
    # rounding of 1e4 - t (ulp 1.8e-12)
    assert abs(h[3, 4] - float(np.float32(4.4))) < 1e-11
    assert abs(h[10, 11] - float(np.float32(7.7))) < 1e-11
    assert h[0, 15] == -1e30


def test_missing_stl_keeps_empty_geometry():
    o = Oracle(_small_cfg())
    assert o.load_stl(b"short") == 1
    assert o.get_mask().sum() == 0


# ---- branches the compiled reference does not run: wall reflection < 1, the z- face as an
# open boundary (NSCBC and sponge) and an open (fluid) ground halo.  C1-, C4- and C5-shaped
# workloads use them, so each is pinned by a closed form or an exact symmetry that ties it
# to golden-pinned code.

def _feq_np(rho, u):
    """Textbook D3Q19 equilibrium w (rho + 3 e.j + 4.5 (e.j)^2/rho - 1.5 j^2/rho), j = rho u."""
    e = np.stack([EX, EY, EZ], axis=1).astype(np.float64)
    cu = e @ np.asarray(u, np.float64)
    return W * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * float(np.dot(u, u)))


def _reflect_scene(alpha, ground=1):
    n = (14, 12, 10)
    cfg = _small_cfg(x_max=float(n[0]), y_max=float(n[1]), z_max=float(n[2]), sponge_faces=0,
                     nscbc_faces=0, wall_reflect=alpha, ground=ground)
    m = np.zeros(n[::-1], np.uint8)
    m[:4, 3:7, 4:8] = 1            # a block on the ground
    m[2:6, 8:10, 10:12] = 1        # a floating block (walls above and below fluid)
    m[7, 1:3, 1:3] = 1
    return cfg, m


def _pull_sources(m, ground):
    """Per interior cell and direction: 0 fluid interior source, 1 solid source (bounce),
    2 source on the open halo (holds the rest state w)."""
    nz, ny, nx = m.shape
    pad = np.full((nz + 2, ny + 2, nx + 2), 2, np.int8)     # halo: rest state
    pad[1:-1, 1:-1, 1:-1] = m
    if ground:
        pad[0] = 1                                          # ground halo is solid
    kinds = np.empty((nz, ny, nx, 19), np.int8)
    for q in range(19):
        kinds[..., q] = pad[1 - EZ[q]:1 - EZ[q] + nz, 1 - EY[q]:1 - EY[q] + ny, 1 - EX[q]:1 - EX[q] + nx]
    return kinds


@pytest.mark.parametrize("alpha", [1.0, 0.8, 0.35])
def test_wall_reflection_closed_form_uniform_state(alpha):
    """From a uniform state f = w (1 + delta), one step pulls w (1 + delta) across fluid
    links, w from the open halo, and w (1 + alpha delta) off a wall (the bounced deviation
    scaled by alpha, the rest part kept).  So every fluid cell's density after one step is
    1 + delta (1 - sum_halo w_q - (1 - alpha) sum_wall w_q).  Reflecting f instead of f - w,
    or scaling by 1 - alpha, fails this."""
    delta = 1e-3
    cfg, m = _reflect_scene(alpha)
    o = Oracle(cfg)
    o.set_mask(m)
    o.set_density(np.full(m.shape, 1.0 + delta))
    o.run(1)
    kinds = _pull_sources(m, ground=1)
    want = 1.0 + delta * (1.0 - (W * (kinds == 2)).sum(-1) - (1.0 - alpha) * (W * (kinds == 1)).sum(-1))
    got = o.read_rho()
    fluid = m == 0
    assert np.any(fluid & (kinds == 1).any(-1))             # walls are hit
    np.testing.assert_allclose(got[fluid], want[fluid], atol=1e-15, rtol=0)


@pytest.mark.parametrize("alpha", [0.8, 0.35])
def test_wall_reflection_pulls_the_inverse_population(alpha):
    """With a uniform moving state f = feq(rho, u) the populations of a cell differ by
    direction.  A wall returns population inv(q) of the cell itself, reflected about rest:
    f_in_q = w_q + alpha (feq_inv(q) - w_q).  The first moments of the pulled set then have
    a closed form per cell; pulling f_q instead of f_inv(q) changes them."""
    cfg, m = _reflect_scene(alpha)
    o = Oracle(cfg)
    o.set_mask(m)
    rho0, u = 1.0 + 2e-3, np.array([3e-3, -2e-3, 1.5e-3])
    fe = _feq_np(rho0, u)
    f = o.read_f()
    f[1:-1, 1:-1, 1:-1][m == 0] = fe
    o.write_f(f)
    o.run(1)
    kinds = _pull_sources(m, ground=1)
    inv = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15])
    pulled = np.where(kinds == 0, fe, np.where(kinds == 2, W, W + alpha * (fe[inv] - W)))
    want = pulled.sum(-1)
    fluid = m == 0
    np.testing.assert_allclose(o.read_rho()[fluid], want[fluid], atol=2e-15, rtol=0)
    # the bounce-back inverse matters: with f_q instead of f_inv(q) some cells differ
    wrong = np.where(kinds == 0, fe, np.where(kinds == 2, W, W + alpha * (fe - W))).sum(-1)
    assert np.abs(wrong - want)[fluid].max() > 1e-6


def test_sponge_values_on_the_bottom_face():
    """The z- sponge uses the distance of the cell centre to z_min, like the other faces."""
    cfg = _small_cfg(x_max=30.0, y_max=30.0, z_max=20.0, sponge_width=6.0, sponge_strength=0.3,
                     sponge_faces=nlos.FACE_ZM | nlos.FACE_ZP, ground=0)
    s = Oracle(cfg).sponge()
    for k in range(6):
        d = k + 0.5
        want = 0.3 * ((6.0 - d) / 6.0) * ((6.0 - d) / 6.0)
        assert s[k, 15, 15] == want, k
        assert s[19 - k, 15, 15] == want, k
    assert np.all(s[6:14] == 0.0)
    cfg2 = _small_cfg(x_max=30.0, y_max=30.0, z_max=20.0, sponge_width=6.0, sponge_strength=0.3,
                      sponge_faces=nlos.FACE_ZM | nlos.FACE_XM, ground=0)
    s2 = Oracle(cfg2).sponge()
    # the nearer face wins: at (i=2, k=4) x- is 2.5 away, z- 4.5
    assert s2[4, 15, 2] == 0.3 * ((6.0 - 2.5) / 6.0) ** 2
    assert s2[2, 15, 20] == 0.3 * ((6.0 - 2.5) / 6.0) ** 2


@pytest.mark.parametrize("faces", ["zonly", "all"])
def test_open_bottom_mirrors_open_top(faces):
    """A scene symmetric under z -> z_max - z, with an open (fluid) ground halo and NSCBC
    plus sponge on both z faces, stays symmetric: the z- ghost (Wp = 0, 05:684-704's rule
    on the bottom face) must mirror the golden-pinned z+ ghost (05:804-835), and the z-
    sponge the z+ one.  Swapping Wp/Wm on z-, dropping the z- sponge or closing the ground
    breaks the symmetry by far more than rounding."""
    n = 16
    zf = nlos.FACE_ZM | nlos.FACE_ZP
    fs = zf if faces == "zonly" else nlos.FACES_ALL
    cfg = _small_cfg(x_max=float(n), y_max=float(n), z_max=float(n), sponge_width=3.0,
                     sponge_strength=0.1, sponge_faces=fs, nscbc_faces=fs, ground=0, tau=0.56)
    m = np.zeros((n, n, n), np.uint8)
    m[6:10, 5:9, 3:7] = 1          # symmetric about the z mid-plane
    o = Oracle(cfg)
    o.set_mask(m)
    z, y, x = np.meshgrid(*(np.arange(n) + 0.5,) * 3, indexing="ij")
    rho = 1.0 + 1e-3 * np.exp(-((x - 10.3) ** 2 + (y - 7.1) ** 2 + (z - 8.0) ** 2) / 4.0)
    o.set_density(rho)
    o.run(30)
    r = o.read_rho() - 1.0
    # the wave has crossed both z faces: the open boundary and the sponge acted
    assert np.abs(r[0]).max() > 1e-6 and np.abs(r[-1]).max() > 1e-6
    # mirrored populations are summed in another order: rounding differences of ~1e-15
    np.testing.assert_allclose(r, r[::-1], atol=1e-13, rtol=0)
    f = o.read_f()[1:-1, 1:-1, 1:-1]
    mirror = [int(np.flatnonzero((EX == EX[q]) & (EY == EY[q]) & (EZ == -EZ[q]))[0]) for q in range(19)]
    np.testing.assert_allclose(f, f[::-1][..., mirror], atol=1e-13, rtol=0)


def test_rest_state_is_fixed_with_open_ground_and_six_open_faces():
    o = Oracle(_small_cfg(sponge_faces=nlos.FACES_ALL, nscbc_faces=nlos.FACES_ALL, ground=0,
                          wall_reflect=0.7))
    m = np.zeros((12, 14, 16), np.uint8)
    m[3:8, 3:7, 4:9] = 1
    o.set_mask(m)
    o.run(20)
    f = o.read_f()[1:-1, 1:-1, 1:-1][m == 0]
    np.testing.assert_allclose(f, np.broadcast_to(W, f.shape), rtol=1e-14)
