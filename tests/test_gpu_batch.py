"""Batched scenes: every scene of a batch gives the same bits as the same scene stepped
alone (the batch changes only how launches are grouped), over runs that mix replayed
two-step graphs with a single trailing step, across consecutive step calls, and after the
workgroup-shape autotuner ran."""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu


def _defs(precision):
    """Uneven scenes: buildings crossing open faces (tracked boundary solids), a partial
    last row segment, rows shorter than 19 cells, reflecting walls with trilinear Ricker
    injection, and a source-free pulse scene."""
    out = []
    cfg, solid, src, probes = scenes.box_scene(precision)
    out.append((cfg, solid, src, probes, None))
    cfg, solid, src, probes = scenes.box_scene(precision, nx=100, ny=12, nz=37, seed=5)
    out.append((cfg, solid, src, probes, None))
    cfg, solid, src, probes = scenes.reflect_scene(precision)
    out.append((cfg, solid, src, probes, None))
    cfg, solid, src, probes = scenes.box_scene(precision, nx=12, ny=20, nz=14, seed=7)
    out.append((cfg, solid, src, probes, None))
    cfg, solid, rho, probes = scenes.pulse_scene(precision)
    out.append((cfg, solid, None, probes, rho))
    for d in out:
        d[0].stream_pattern = nlos.PATTERN_TWO_ARRAY
    return out


def _setup(e, d):
    cfg, solid, src, probes, rho = d
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if src is not None:
        e.set_source(src)
    e.set_probes(probes)


def _single(d, steps):
    e = nlos.Engine(d[0])
    _setup(e, d)
    e.step(steps)
    e.synchronize()
    res = (e.read_probes(), e.debug_read_f())
    e.close()
    return res


@pytest.mark.parametrize("precision", ["f64", "f32", "f16s", "b16"])
def test_batch_equals_scenes_stepped_alone(gpu, precision):
    print(f"[batch vs single] precision={precision}", flush=True)
    defs = _defs(precision)
    b = nlos.Batch([d[0] for d in defs])
    for i, d in enumerate(defs):
        _setup(b.scene(i), d)
    b.step(7)       # three graph replays of two steps and one direct step
    b.step(10)      # the same graph again, from the other phase
    b.synchronize()
    for i, d in enumerate(defs):
        want_p, want_f = _single(d, 17)
        got_p, got_f = b.scene(i).read_probes(), b.scene(i).debug_read_f()
        assert got_p.shape == want_p.shape
        assert np.array_equal(got_p, want_p), (i, np.abs(got_p - want_p).max())
        assert np.array_equal(got_f, want_f, equal_nan=True), i
    b.close()


def test_batch_autotune_keeps_state_and_results(gpu):
    defs = _defs("f32")
    b = nlos.Batch([d[0] for d in defs])
    for i, d in enumerate(defs):
        _setup(b.scene(i), d)
    b.step(3)
    t = b.autotune(reps=1)
    # three widths for all scenes, then three for each of the two row-length classes
    # (rows of one wave: nx = 40, 30, 12, 28; of two: nx = 100), and the classes' choice
    # timed together when it differs from the best single width
    assert 9 <= t.candidates <= 10 and t.best_ms > 0.0
    w = b.tile_widths()
    assert len(w) == len(defs) and set(w) <= {64, 128, 256}
    if t.per_scene:
        assert t.block_x == 0 and t.block_y == 0 and len(set(w)) > 1
    else:
        assert t.block_x * t.block_y == 256 and w == [t.block_x] * len(defs)
    b.step(6)
    b.synchronize()
    for i, d in enumerate(defs):
        want_p, want_f = _single(d, 9)
        assert np.array_equal(b.scene(i).read_probes(), want_p), i
        assert np.array_equal(b.scene(i).debug_read_f(), want_f, equal_nan=True), i
    b.close()


@pytest.mark.parametrize("precision", ["f64", "f16s", "b16"])
def test_batch_with_a_tile_width_per_scene(gpu, precision):
    """Scenes of one launch sweeping with different tile shapes (64x4, 128x2, 256x1 cells)
    still give the bits of each scene stepped alone; widths other than 64, 128, 256 are
    refused."""
    print(f"[batch, per-scene tile widths] precision={precision}", flush=True)
    defs = _defs(precision)
    b = nlos.Batch([d[0] for d in defs])
    for i, d in enumerate(defs):
        _setup(b.scene(i), d)
    widths = [64, 128, 256, 128, 64]
    b.set_tile_widths(widths)
    assert b.tile_widths() == widths
    with pytest.raises(nlos.EngineError):
        b.set_tile_widths([64, 96, 256, 128, 64])
    with pytest.raises(nlos.EngineError):
        b.set_tile_widths([64, 128])
    b.step(7)
    b.set_tile_widths([256, 64, 128, 64, 256])   # a new choice between two step calls
    b.step(10)
    b.synchronize()
    for i, d in enumerate(defs):
        want_p, want_f = _single(d, 17)
        assert np.array_equal(b.scene(i).read_probes(), want_p), i
        assert np.array_equal(b.scene(i).debug_read_f(), want_f, equal_nan=True), i
    b.close()


def test_batch_scene_cannot_step_alone(gpu):
    defs = _defs("f32")[:2]
    b = nlos.Batch([d[0] for d in defs])
    with pytest.raises(nlos.EngineError):
        b.scene(0).step(1)
    b.step(2)
    assert b.scene(1).info().step == 2
    b.close()


def test_batch_of_mixed_precisions_is_refused(gpu):
    """A batch's scenes share one storage precision (one kernel set per launch)."""
    a = _defs("f32")[0][0]
    b = _defs("f16s")[0][0]
    with pytest.raises(nlos.EngineError) as ex:
        nlos.Batch([a, b])
    assert ex.value.code == 1


def test_scene_reads_wait_for_the_batch(gpu):
    """nlos_batch_step returns with its graph replays still running on the batch's stream.
    A scene read issued right after it (no nlos_batch_synchronize) must see every step:
    probes, the field and the state equal the scenes stepped alone, bit for bit."""
    defs = []
    for seed in (11, 12):
        cfg, solid, src, probes = scenes.box_scene("f32", nx=160, ny=128, nz=64, seed=seed)
        cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
        defs.append((cfg, solid, src, probes, None))
    b = nlos.Batch([d[0] for d in defs])
    for i, d in enumerate(defs):
        _setup(b.scene(i), d)
    b.step(300)
    got = [(b.scene(i).read_probes(), b.scene(i).debug_read_f()) for i in range(len(defs))]
    b.close()
    for i, d in enumerate(defs):
        want_p, want_f = _single(d, 300)
        assert got[i][0].shape == want_p.shape == (len(d[3]), 300)
        assert np.array_equal(got[i][0], want_p), i
        assert np.array_equal(got[i][1], want_f, equal_nan=True), i


def test_direct_launches_equal_graph_replays(gpu, monkeypatch):
    """NLOS_BATCH_GRAPH=0 launches every step's kernels directly instead of replaying the
    captured two-step graph (for tools that cannot trace graph launches): the same bits."""
    defs = _defs("f32")
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NLOS_BATCH_GRAPH", flag)
        b = nlos.Batch([d[0] for d in defs])
        for i, d in enumerate(defs):
            _setup(b.scene(i), d)
        b.step(9)
        b.synchronize()
        runs.append([(b.scene(i).read_probes(), b.scene(i).debug_read_f()) for i in range(len(defs))])
        b.close()
    for (pa, fa), (pb, fb) in zip(*runs):
        assert np.array_equal(pa, pb)
        assert np.array_equal(fa, fb, equal_nan=True)


@pytest.mark.parametrize("precision", ["f32", "b16"])
def test_persistent_sweep_equals_one_workgroup_per_tile(gpu, precision, monkeypatch):
    """The persistent multi-scene sweep (a device-sized grid walking the (scene, tile) work
    list) and NLOS_BATCH_PERSISTENT=0 (one workgroup per tile) give the same bits, with tile
    widths that differ from scene to scene."""
    defs = _defs(precision)
    runs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("NLOS_BATCH_PERSISTENT", flag)
        b = nlos.Batch([d[0] for d in defs])
        for i, d in enumerate(defs):
            _setup(b.scene(i), d)
        b.set_tile_widths([128, 64, 256, 64, 128])
        b.step(11)
        b.synchronize()
        runs.append([(b.scene(i).read_probes(), b.scene(i).debug_read_f()) for i in range(len(defs))])
        b.close()
    for (pa, fa), (pb, fb) in zip(*runs):
        assert np.array_equal(pa, pb)
        assert np.array_equal(fa, fb, equal_nan=True)
