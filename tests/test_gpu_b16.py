"""Block floating point 16-bit storage (NLOS_B16) on the GPU.

B16 keeps the rest population x_0 = g_0 / w_0 in FP32, a 16-bit mantissa of x_q = g_q / w_q
for q = 1..18 and one exponent byte per cell (storage layout in csrc/common.hpp, Mode).  Pinned here:
* the coding itself, bit for bit, on a state set from outside (set_density), against a
  numpy model of the rounding rule;
* the step against FP64 (the oracle) within TOL_B16, and at least twice as accurate as
  FP16S on the same scene (the reason the mode exists);
* the rest state as an exact fixed point, checkpoint/restore and every workgroup shape
  bitwise, non-finite detection, and the refusal of the in-place pattern.
"""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

# Probe and field relative L2 against FP64 after 80 steps of the small scenes: the bound the
# 16-bit storage meets at the benchmark configuration too.
TOL_B16 = 1e-3
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def _run(cfg, solid, src, probes, steps, rho=None):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    if src is not None:
        e.set_source(src)
    e.set_probes(probes)
    e.step(steps, capture_field=True)
    e.synchronize()
    return e


def _two_array(make, precision, **kw):
    out = make(precision, **kw)
    out[0].stream_pattern = nlos.PATTERN_TWO_ARRAY
    return out


@pytest.mark.parametrize("scene", ["box", "reflect"])
def test_b16_vs_f64_and_against_f16s(gpu, scene):
    make = scenes.box_scene if scene == "box" else scenes.reflect_scene
    print(f"[b16 vs f64] scene={scene}", flush=True)
    cfg64, solid, src, probes = make("f64")
    o = Oracle(cfg64)
    o.set_mask(solid)
    o.set_source(src)
    o.set_probes(probes)
    o.run(80)
    err = {}
    for p in ("b16", "f16s"):
        cfg, _, _, _ = _two_array(make, p)
        e = _run(cfg, solid, src, probes, 80)
        err[p] = (scenes.rel_l2(e.read_probes(), o.read_probes()), scenes.rel_l2(e.read_field(), o.read_field()))
        e.close()
    print(f"  b16 probe {err['b16'][0]:.3e} field {err['b16'][1]:.3e};"
          f" f16s probe {err['f16s'][0]:.3e} field {err['f16s'][1]:.3e}")
    assert err["b16"][0] <= TOL_B16 and err["b16"][1] <= TOL_B16
    assert err["b16"][0] <= 0.5 * err["f16s"][0]
    assert err["b16"][1] <= 0.5 * err["f16s"][1]


def _b16_model(x32):
    """The coding rule for a cell whose 19 normalised values all equal x32 (FP32): x_0 kept
    in FP32; for slots 1..18 the exponent e of |x| (|x| < 2^e, at least -113), mantissa
    rint(x * 2^(15-e)) (half to even) saturated to 16 bits, decoded m * 2^(e-15); times w_q
    in FP32, plus w_q in FP64."""
    ax = np.abs(x32.astype(np.float64))
    _, e = np.frexp(ax)
    e = np.maximum(np.where(ax > 0, e, 0), -113)
    m = np.clip(np.rint(x32.astype(np.float64) * np.ldexp(1.0, 15 - e)), -32768, 32767)
    dec = np.ldexp(m, e - 15).astype(np.float32)
    x = np.repeat(dec[..., None], 19, axis=-1)
    x[..., 0] = x32
    g = (x * W.astype(np.float32)).astype(np.float32)
    return g.astype(np.float64) + W


def test_b16_coding_of_a_set_state_is_exact(gpu):
    cfg, solid, rho, probes = scenes.pulse_scene("b16")
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    rng = np.random.default_rng(5)
    # amplitudes over many binades, both signs, exact zeros
    rho = 1.0 + rng.choice([-1.0, 1.0], rho.shape) * 10.0 ** rng.uniform(-12, -1, rho.shape)
    rho[::3, ::2, ::5] = 1.0
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_density(rho)
    f = e.debug_read_f()
    fluid = solid == 0
    x32 = (rho - 1.0).astype(np.float32)
    want = _b16_model(x32)
    assert np.array_equal(f[fluid], want[fluid])
    e.close()


def test_b16_rest_state_is_exact_fixed_point(gpu):
    cfg, solid, _, probes = _two_array(scenes.box_scene, "b16")
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_probes(probes)
    e.step(31, capture_field=True)
    e.synchronize()
    assert not np.any(e.read_probes())
    assert not np.any(e.read_field())
    f = e.debug_read_f()[solid == 0]
    assert np.array_equal(f, np.broadcast_to(W, f.shape))


def test_b16_pulse_with_nscbc_on_all_faces(gpu):
    """Open boundaries everywhere (the x-face records carry FP32 deviations in B16)."""
    cfg64, solid, rho, probes = scenes.pulse_scene("f64")
    o = Oracle(cfg64)
    o.set_mask(solid)
    o.set_density(rho)
    o.set_probes(probes)
    o.run(60)
    cfg, _, _, _ = _two_array(scenes.pulse_scene, "b16")
    e = _run(cfg, solid, None, probes, 60, rho=rho)
    err = scenes.rel_l2(e.read_probes(), o.read_probes())
    ferr = scenes.rel_l2(e.read_field(), o.read_field())
    print(f"  b16 pulse probe {err:.3e} field {ferr:.3e}")
    assert err <= TOL_B16 and ferr <= TOL_B16
    e.close()


def test_b16_checkpoint_restore_is_bitwise(gpu):
    cfg, solid, src, probes = _two_array(scenes.box_scene, "b16")
    e = _run(cfg, solid, src, probes, 17)
    blob = e.checkpoint()
    e.step(20)
    e.synchronize()
    p1, f1 = e.read_probes(), e.debug_read_f()
    e.close()
    r = nlos.Engine(cfg)
    r.set_mask(solid)
    r.set_source(src)
    r.set_probes(probes)
    r.restore(blob)
    r.step(20)
    r.synchronize()
    assert np.array_equal(r.read_probes(), p1)
    assert np.array_equal(r.debug_read_f(), f1, equal_nan=True)


def test_b16_launch_shapes_agree(gpu, monkeypatch):
    cfg, solid, src, probes = _two_array(scenes.box_scene, "b16", nx=100, ny=13, nz=37)
    runs = []
    for block in ("256x1", "64x4", "128x2"):
        monkeypatch.setenv("NLOS_BLOCK", block)
        e = _run(cfg, solid, src, probes, 23)
        runs.append((e.read_probes(), e.debug_read_f()))
        e.close()
    for r in runs[1:]:
        assert np.array_equal(runs[0][0], r[0])
        assert np.array_equal(runs[0][1], r[1], equal_nan=True)


def test_b16_nonfinite_population_is_reported(gpu):
    cfg, solid, src, probes = _two_array(scenes.box_scene, "b16")
    e = _run(cfg, solid, src, probes, 3)
    e.debug_poke(20, 3, 20, 7, float("nan"))
    e.step(2)
    with pytest.raises(nlos.EngineError) as ex:
        e.synchronize()
    assert ex.value.code == 4


def test_b16_refuses_the_inplace_pattern(gpu):
    cfg, _, _, _ = scenes.box_scene("b16")
    cfg.stream_pattern = nlos.PATTERN_INPLACE
    with pytest.raises(nlos.EngineError) as ex:
        nlos.Engine(cfg)
    assert ex.value.code == 7

