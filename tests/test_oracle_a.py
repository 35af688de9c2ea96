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
