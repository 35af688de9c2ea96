"""Seeded scenes of random shape and options (rows from 1 to 200 cells, every face option,
partial reflection, sky height, all waveforms and injections, initial densities): the FP64
engine equals the oracle bit for bit in both stream patterns, FP32 in place equals FP32 with
two arrays bit for bit, and a batch of such scenes equals the scenes stepped alone."""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

SEEDS = list(range(40))


def _engine(cfg, solid, src, samples, probes, rho, steps):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    e.set_source(src, samples)
    e.set_probes(probes)
    e.step(steps, capture_field=True)
    e.synchronize()
    out = (e.read_probes(), e.read_field(), e.debug_read_f()[solid == 0])
    e.close()
    return out


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_fp64_equals_oracle(gpu, seed):
    cfg, solid, src, samples, probes, rho, steps = scenes.random_scene(seed, "f64")
    print(f"[random scene {seed}] {cfg.x_max - cfg.x_min:.1f} m x {cfg.y_max - cfg.y_min:.1f} m x "
          f"{cfg.z_max - cfg.z_min:.1f} m, dx {cfg.dx}, faces sponge {cfg.sponge_faces} nscbc {cfg.nscbc_faces}, "
          f"ground {cfg.ground}, waveform {src.waveform}, injection {src.injection}, steps {steps}", flush=True)
    o = Oracle(cfg)
    o.set_mask(solid)
    if rho is not None:
        o.set_density(rho)
    o.set_source(src, samples)
    o.set_probes(probes)
    o.run(steps)
    want = (o.read_probes(), o.read_field(), o.read_f()[1:-1, 1:-1, 1:-1, :][solid == 0])
    assert np.abs(want[1]).max() > 0.0
    for pattern in ("two_array", "inplace"):
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        got = _engine(cfg, solid, src, samples, probes, rho, steps)
        for g, w in zip(got, want):
            assert g.shape == w.shape
            assert np.array_equal(g, w), (pattern, np.abs(g - w).max())


@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_fp32_inplace_equals_two_array(gpu, seed):
    cfg, solid, src, samples, probes, rho, steps = scenes.random_scene(seed, "f32")
    runs = []
    for pattern in ("two_array", "inplace"):
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        runs.append(_engine(cfg, solid, src, samples, probes, rho, steps))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "f32", "b16"])
def test_random_scenes_batched_equal_alone(gpu, precision):
    defs = [scenes.random_scene(seed, precision) for seed in range(12)]
    for d in defs:
        d[0].stream_pattern = nlos.PATTERN_TWO_ARRAY
    b = nlos.Batch([d[0] for d in defs])
    for i, (cfg, solid, src, samples, probes, rho, steps) in enumerate(defs):
        e = b.scene(i)
        e.set_mask(solid)
        if rho is not None:
            e.set_density(rho)
        e.set_source(src, samples)
        e.set_probes(probes)
    b.step(25)
    b.synchronize()
    for i, (cfg, solid, src, samples, probes, rho, steps) in enumerate(defs):
        want = _engine(cfg, solid, src, samples, probes, rho, 25)
        assert np.array_equal(b.scene(i).read_probes(), want[0]), i
        assert np.array_equal(b.scene(i).debug_read_f()[solid == 0], want[2]), i
    b.close()


@pytest.mark.parametrize("seed", SEEDS[:20])
def test_random_scene_fp16_inplace_equals_two_array(gpu, seed):
    """Scaled FP16 without partial reflection (with it the two patterns round the wall copies
    in another order): in place equals two arrays bit for bit."""
    cfg, solid, src, samples, probes, rho, steps = scenes.random_scene(seed, "f16s")
    cfg.wall_reflect = 1.0
    runs = []
    for pattern in ("two_array", "inplace"):
        cfg.stream_pattern = nlos.PATTERNS[pattern]
        runs.append(_engine(cfg, solid, src, samples, probes, rho, steps))
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["f64", "b16"])
@pytest.mark.parametrize("seed", SEEDS[:12])
def test_random_scene_step_forms_agree(gpu, seed, precision, monkeypatch):
    """The same steps through the replayed step graph, through direct launches
    (NLOS_STEP_GRAPH=0) and through the allocation-free entry point, in uneven chunks with a
    field capture: probes, field and populations bit for bit."""
    cfg, solid, src, samples, probes, rho, steps = scenes.random_scene(seed, precision)
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    chunks = [steps // 3, 1, steps - steps // 3 - 1]
    runs = []
    for form in ("graph", "direct", "noalloc"):
        monkeypatch.setenv("NLOS_STEP_GRAPH", "0" if form == "direct" else "1")
        e = nlos.Engine(cfg)
        e.set_mask(solid)
        if rho is not None:
            e.set_density(rho)
        e.set_source(src, samples)
        e.set_probes(probes)
        e.step(1, capture_field=True)      # a first capture allocates the field buffer
        for n in chunks:
            if form == "noalloc":
                e.step_noalloc(n, capture_field=True)
            else:
                e.step(n, capture_field=True)
        e.synchronize()
        runs.append((e.read_probes(), e.read_field(), e.debug_read_f()[solid == 0]))
        e.close()
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "b16"])
@pytest.mark.parametrize("seed", SEEDS[:8])
def test_random_scene_checkpoint_restore(gpu, seed, precision):
    """A checkpoint taken mid-run and restored into a fresh engine continues bit for bit."""
    cfg, solid, src, samples, probes, rho, steps = scenes.random_scene(seed, precision)
    if precision == "b16":
        cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY

    def fresh():
        e = nlos.Engine(cfg)
        e.set_mask(solid)
        if rho is not None:
            e.set_density(rho)
        e.set_source(src, samples)
        e.set_probes(probes)
        return e

    a = fresh()
    a.step(steps // 2)
    blob = a.checkpoint()
    a.step(steps - steps // 2, capture_field=True)
    a.synchronize()
    want = (a.read_probes(), a.read_field(), a.debug_read_f())
    a.close()
    r = fresh()
    r.restore(blob)
    r.step(steps - steps // 2, capture_field=True)
    r.synchronize()
    got = (r.read_probes(), r.read_field(), r.debug_read_f())
    r.close()
    for g, w in zip(got, want):
        assert np.array_equal(g, w, equal_nan=True)
