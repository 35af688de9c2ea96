"""nlos_step_noalloc (the step for callers that cannot allocate device memory) against
nlos_step, through the C ABI.

The same scene stepped by both entry points must give the same bits: probe histories and
the whole population state.  Two calls in a row on one GPU share its static staging array,
the second with a smaller scene and fewer steps; a canyon scene with the largest source
stencil (radius 3, 343 cells per step) runs 800 steps, more than one staged chunk (764
steps of 343 entries fit the array).  Requests that would need an allocation are refused
and leave the engine as it was.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu


def _engine(cfg, solid, src, probes, samples=None):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if src is not None:
        e.set_source(src, samples)
    e.set_probes(probes)
    return e


def _both(cfg, solid, src, probes, steps, samples=None):
    lib = nlos.load_library()
    a = _engine(cfg, solid, src, probes, samples)
    rc = lib.nlos_step(a.h, steps, 0)
    assert rc == 0, nlos.load_library().nlos_last_error()
    a.synchronize()
    b = _engine(cfg, solid, src, probes, samples)
    rc = lib.nlos_step_noalloc(b.h, steps, 0)
    assert rc == 0, lib.nlos_last_error()
    b.synchronize()
    out = (a.read_probes(), a.debug_read_f(), b.read_probes(), b.debug_read_f(), b.info().step)
    a.close()
    b.close()
    return out


@pytest.mark.parametrize("precision", ["f64", "b16", "f16s"])
def test_noalloc_equals_step_on_two_calls_the_second_smaller(gpu, precision):
    print(f"[noalloc] {precision}: 40x32x24 box scene, 60 steps; then 20x16x12, 25 steps", flush=True)
    for dims, steps in (({}, 60), ({"nx": 20, "ny": 16, "nz": 12}, 25)):
        cfg, solid, src, probes = scenes.box_scene(precision, **dims)
        cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
        pa, fa, pb, fb, step = _both(cfg, solid, src, probes, steps)
        assert step == steps
        assert pa.shape == pb.shape == (len(probes), steps)
        assert np.abs(pa).max() > 0
        assert np.array_equal(pa, pb)
        assert np.array_equal(fa, fb, equal_nan=True)


def test_noalloc_equals_step_over_several_staged_chunks(gpu):
    print("[noalloc] canyon, radius-3 stencil, 800 steps (two staged chunks), f32", flush=True)
    cfg, solid, src, _, probes = scenes.canyon_scene("f32")
    pa, fa, pb, fb, step = _both(cfg, solid, src, probes, 800)
    assert step == 800
    assert np.abs(pa).max() > 0
    assert np.array_equal(pa, pb)
    assert np.array_equal(fa, fb, equal_nan=True)


def test_noalloc_inplace_pattern_and_capture_after_a_first_capture(gpu):
    print("[noalloc] in-place pattern f32, capture after nlos_step captured once", flush=True)
    lib = nlos.load_library()
    cfg, solid, src, probes = scenes.box_scene("f32")
    cfg.stream_pattern = nlos.PATTERN_INPLACE
    a = _engine(cfg, solid, src, probes)
    b = _engine(cfg, solid, src, probes)
    for e in (a, b):
        assert lib.nlos_step(e.h, 3, nlos.STEP_CAPTURE_FIELD) == 0
    assert lib.nlos_step(a.h, 31, nlos.STEP_CAPTURE_FIELD) == 0
    assert lib.nlos_step_noalloc(b.h, 31, nlos.STEP_CAPTURE_FIELD) == 0, lib.nlos_last_error()
    a.synchronize()
    b.synchronize()
    assert np.array_equal(a.read_probes(), b.read_probes())
    assert np.array_equal(a.read_field(), b.read_field())
    a.close()
    b.close()


def test_noalloc_refuses_what_would_allocate(gpu):
    print("[noalloc] refusals", flush=True)
    lib = nlos.load_library()
    cfg, solid, src, probes = scenes.box_scene("f32", nx=20, ny=16, nz=12)
    e = _engine(cfg, solid, src, probes)
    # the probe history holds NLOS_PROBE_HISTORY_RESERVE = 4096 steps after nlos_set_probes
    assert lib.nlos_step_noalloc(e.h, 4097, 0) == 1
    assert b"probe history" in lib.nlos_last_error()
    assert e.info().step == 0
    assert lib.nlos_step_noalloc(e.h, 2, nlos.STEP_CAPTURE_FIELD) == 3
    assert b"field buffer" in lib.nlos_last_error()
    e.timing(True)
    assert lib.nlos_step_noalloc(e.h, 2, 0) == 3
    e.timing(False)
    assert e.info().step == 0
    assert lib.nlos_step_noalloc(e.h, 4096, 0) == 0, lib.nlos_last_error()
    e.synchronize()
    assert e.read_probes().shape == (len(probes), 4096)
    e.close()
