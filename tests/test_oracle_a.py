"""The family-A oracle (oracle/ref_cpu_a.cpp) against the compiled reference 01's probe files
(tests/golden/ref01, written by tests/golden/make_ref_goldens.py).  CPU only."""
from __future__ import annotations

import math
import os

import numpy as np

import family_a as fa
from oracle_ffi import OracleA

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref01")


def test_oracle_a_reproduces_01_probe_files():
    """The first 120 steps of 01's run (all 402 records are checked on the GPU), through the
    file format of 01:481-495 (default stream formatting: %g)."""
    cfg = fa.config()
    solid, rho = fa.cylinder_and_pulse(cfg)
    o = OracleA(cfg)
    o.set_mask(solid)
    o.set_density(rho)
    names, pts = fa.ref01_probes()
    o.set_probes(pts)
    n = 120
    o.run(n)
    p = o.read_probes()
    assert p.shape == (3, n + 1)
    dt = math.sqrt(1.0 / 3.0) * cfg.dx / cfg.c_phys
    for i, name in enumerate(names):
        lines = open(os.path.join(GOLDEN, f"probe_{name}.dat")).read().split("\n")
        assert lines[0] == "# Time Pressure"
        assert len(lines) == fa.ref01_steps(cfg) + 3        # header, s = 0..max, final newline
        mine = ["%g %g" % (s * dt, p[i, s] * 100.0) for s in range(n + 1)]
        assert mine == lines[1:n + 2], name


def test_oracle_a_rest_state_moves_only_by_the_weight_sum_rounding():
    """With no pulse every cell stays at w up to rounding: in FP64 the weights sum to
    1 + 2^-52, so each collision moves f by a few ulps (01's frames show the same 2.2e-14
    pressure at rest); a mistake in streaming, bounce-back or the halos would move it far
    more."""
    cfg = fa.config(x=(-1.0, 1.0), y=(-1.0, 1.0), z=(-0.1, 0.1), dx=0.1)
    solid, _ = fa.cylinder_and_pulse(cfg, radius=0.3)
    o = OracleA(cfg)
    o.set_mask(solid)
    o.run(7)
    f = o.read_f()[solid == 0]
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    assert sum(w) == 1.0 + 2.0 ** -52
    assert np.abs(f - w).max() <= 1e-15
    assert np.abs(o.read_rho()[solid == 0] - 1.0).max() <= 1e-14


# ---- 04's line source (04:212-235) ----------------------------------------------------------

def _line(step0=0, **kw):
    import nlos_engine as nlos
    src = nlos.reference04_line_source(step0)
    src.dt = 0.577350269 * 0.5 / 1.0          # 04's dt rule at dx = 0.5
    for k, v in kw.items():
        setattr(src, k, v)
    return src


def test_line_source_injected_mass_closed_form():
    """One step from rest with no sponge and no walls: the populations gain, per fluid cell
    within sqrt(r2_max) of the line in every z layer, amp exp(-ln2 d^2/b) sin(omega t) dt,
    split as w_q, with t counted from step0.  The step itself moves no mass this far from
    the open faces.  A wrong cut-off, time origin, exponent or layer coverage fails this."""
    cfg = fa.coarse04_config()
    nx, ny, nz = fa.dims(cfg)
    solid = np.zeros((nz, ny, nx), np.uint8)
    solid[:, 24:28, 24:28] = 1            # a column the source disk covers: skipped cells
    o = OracleA(cfg)
    o.set_mask(solid)
    o.set_sponge(np.zeros((nz, ny, nx)))
    src = _line(step0=-7)
    o.set_line_source(src)
    f0 = o.read_f()
    o.run(1)
    df = (o.read_f() - f0)[solid == 0].sum(axis=0)
    t = 7 * src.dt
    want = 0.0
    n_cells = 0
    for j in range(ny):
        for i in range(nx):
            x = cfg.x_min + i * cfg.dx + 0.5 * cfg.dx
            y = cfg.y_min + j * cfg.dx + 0.5 * cfg.dx
            d2 = x * x + (y + 6.0) ** 2
            if d2 <= 36.0 and not solid[0, j, i]:
                want += 0.01 * math.exp(-math.log(2.0) * d2 / 2.0) * math.sin(math.pi / 15.0 * t) * src.dt
                n_cells += 1
    want *= nz
    assert 430 < n_cells < 470              # pi * 6^2 / 0.5^2 = 452 cells, less the column
    # the collision redistributes among directions and conserves the sum; the rest state's
    # weight-sum rounding (1 + 2^-52) drifts it by ~1e-11 relative
    np.testing.assert_allclose(df.sum(), want, rtol=1e-9)
    o2 = OracleA(cfg)
    o2.set_mask(solid)
    o2.set_sponge(np.zeros((nz, ny, nx)))
    o2.set_line_source(_line(step0=0))
    f0 = o2.read_f()
    o2.run(1)
    assert np.abs((o2.read_f() - f0)[solid == 0]).max() < 1e-15      # sin(0) = 0 at s = step0


def test_line_source_with_04_geometry_is_z_invariant_and_x_symmetric():
    """04's scene (cylinder r < 4 along z, radial sponge, line source at (0, -6)) is the same
    in every z layer and mirror-symmetric in x: with z periodic every layer repeats the
    same arithmetic (bit for bit), and the x mirror holds up to summation-order rounding."""
    cfg = fa.coarse04_config()
    nx, ny, nz = fa.dims(cfg)
    o = OracleA(cfg)
    o.set_mask(fa.cylinder_04(cfg))
    o.set_sponge(fa.radial_sponge_04(cfg))
    o.set_line_source(_line(step0=0))
    o.run(40)
    r = o.read_rho() - 1.0
    assert np.abs(r).max() > 1e-5
    for k in range(1, nz):
        assert np.array_equal(r[k], r[0]), k
    np.testing.assert_allclose(r, r[:, :, ::-1], atol=1e-13, rtol=0)
