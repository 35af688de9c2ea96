"""Step graph of nlos_step (two-array pattern, every storage mode).

nlos_step launches each step's kernels (source entries, ghost layer, tracked boundary
solids, sweep, counter) with the step read from a device counter, so that a pair of steps
is captured once as a HIP graph and replayed; the graph is captured again whenever an
argument it bakes in moves (a grown probe history or schedule, the other population array
after an odd number of steps).  NLOS_STEP_GRAPH=0 launches every step directly, and timing
mode always does.  All ways must give the same bits (probes, populations, captured field)
however the steps are split into calls.
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
    if name == "pulse":
        cfg, solid, rho, probes = scenes.pulse_scene(precision)
        return cfg, solid, None, rho, probes
    cfg, solid, src, probes = scenes.reflect_scene(precision)
    return cfg, solid, src, None, probes


def _run(cfg, solid, src, rho, probes, chunks, timing=False):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if src is not None:
        e.set_source(src)
    e.set_probes(probes)
    if timing:
        e.timing(True)
    for u, n in enumerate(chunks):
        e.step(n, capture_field=u == len(chunks) - 1)
    e.synchronize()
    out = (e.read_probes(), e.debug_read_f(), e.read_field())
    e.close()
    return out


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "b16"])
@pytest.mark.parametrize("scene", ["box", "pulse", "reflect"])
def test_step_graph_equals_direct_launches(gpu, precision, scene, monkeypatch):
    print(f"[step graph] scene={scene} precision={precision}", flush=True)
    cfg, solid, src, rho, probes = _scene(scene, precision)
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    monkeypatch.setenv("NLOS_STEP_GRAPH", "0")
    want = _run(cfg, solid, src, rho, probes, [43])
    monkeypatch.setenv("NLOS_STEP_GRAPH", "1")
    runs = [([43], False), ([1, 6, 3, 33], False), ([2, 2, 39], False), ([43], True)]
    for chunks, timing in runs:
        got = _run(cfg, solid, src, rho, probes, chunks, timing)
        for a, b in zip(got, want):
            assert np.array_equal(a, b, equal_nan=True), (chunks, timing)
    assert np.abs(want[0]).max() > 0
