"""Gradients by the discrete adjoint against central finite differences, with open faces at
the rest state and through the reference's NSCBC faces.

The loss is L = sum over probes and steps of (rho-1 - obs)^2 with seeded observations.  In
FP64 the adjoint gradient with respect to the source-path shift, the amplitude and the
samples of a sampled waveform must match central differences of the GPU forward and of the
oracle forward (the FP64 CPU restatement of the reference) to 1e-5 relative; FP32 adjoint
gradients must match the FP64 ones to 2e-3 relative."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import nlos_engine as nlos
import scenes
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

STEPS = 70
EVERY = 16           # segments of 16, 16, 16, 16, 6 steps
TOL_FD64 = 1e-5
TOL_F32 = 2e-3


def _obs(n_probes, steps, seed=11):
    rng = np.random.default_rng(seed)
    return 7e-6 * rng.standard_normal((n_probes, steps))


def _shifted(src, d):
    s = nlos.Source()
    C.pointer(s)[0] = src
    s.sx += d[0]; s.ex += d[0]
    s.sy += d[1]; s.ey += d[1]
    s.sz += d[2]; s.ez += d[2]
    return s


def _copy(src, **kw):
    s = nlos.Source()
    C.pointer(s)[0] = src
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _loss_gpu(precision, src, samples=None, record=False, every=EVERY, obs=None, open_faces=False):
    cfg, solid, _, _, probes = scenes.canyon_scene(precision, open_faces=open_faces)
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_source(src, samples)
    e.set_probes(probes)
    if record:
        e.record_forward(STEPS, every)
    else:
        e.step(STEPS)
    p = e.read_probes()
    r = p - obs
    return e, float(np.sum(r * r)), 2.0 * r


def _loss_oracle(src, samples=None, obs=None, open_faces=False):
    cfg, solid, _, _, probes = scenes.canyon_scene("f64", open_faces=open_faces)
    o = Oracle(cfg)
    o.set_mask(solid)
    o.set_source(src, samples)
    o.set_probes(probes)
    o.run(STEPS)
    r = o.read_probes() - obs
    return float(np.sum(r * r))


def _grad(precision, src, samples=None, every=EVERY, obs=None, open_faces=False):
    e, L, dl = _loss_gpu(precision, src, samples, record=True, every=every, obs=obs, open_faces=open_faces)
    g = e.backward(dl, n_samples=0 if samples is None else len(samples))
    e.close()
    return L, g


def _fd4(loss, h):
    """Central difference of fourth order: (8 (L(h) - L(-h)) - (L(2h) - L(-2h))) / 12h."""
    return (8.0 * (loss(h) - loss(-h)) - (loss(2 * h) - loss(-2 * h))) / (12.0 * h)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_adjoint_fp64_matches_central_differences(gpu):
    cfg, solid, src, _, probes = scenes.canyon_scene("f64")
    obs = _obs(len(probes), STEPS)
    L, g = _grad("f64", src, obs=obs)
    assert L > 0.0 and np.abs(g["d_position"]).max() > 0.0 and g["d_amp"] != 0.0
    h = 1e-3
    fd = np.zeros(3)
    for ax in range(3):
        def loss(x):
            d = [0.0, 0.0, 0.0]
            d[ax] = x
            return _loss_gpu("f64", _shifted(src, d), obs=obs)[1]
        fd[ax] = _fd4(loss, h)
        print(f"[adjoint f64] axis {ax}: adjoint {g['d_position'][ax]:.10e} fd {fd[ax]:.10e}", flush=True)
    # relative to the gradient vector (one component can be far smaller than the others)
    assert np.linalg.norm(g["d_position"] - fd) < TOL_FD64 * np.linalg.norm(fd)
    # L is a polynomial of low degree in the amplitude (the response is nearly linear), which
    # the fourth-order difference resolves at any step; a step of 1e-2 keeps roundoff out
    ha = 1e-2 * src.amp
    fd = _fd4(lambda x: _loss_gpu("f64", _copy(src, amp=src.amp + x), obs=obs)[1], ha)
    print(f"[adjoint f64] amp: adjoint {g['d_amp']:.10e} fd {fd:.10e}", flush=True)
    assert _rel(g["d_amp"], fd) < TOL_FD64


def test_adjoint_matches_oracle_central_differences(gpu):
    cfg, solid, src, _, probes = scenes.canyon_scene("f64")
    obs = _obs(len(probes), STEPS, seed=12)
    _, g = _grad("f64", src, obs=obs)
    h = 1e-3
    fd = _fd4(lambda x: _loss_oracle(_shifted(src, [x, 0.0, 0.0]), obs=obs), h)
    assert _rel(g["d_position"][0], fd) < TOL_FD64
    ha = 1e-2 * src.amp
    fd = _fd4(lambda x: _loss_oracle(_copy(src, amp=src.amp + x), obs=obs), ha)
    assert _rel(g["d_amp"], fd) < TOL_FD64


def test_adjoint_sampled_waveform_gradient(gpu):
    cfg, solid, src, samples, probes = scenes.canyon_scene("f64", waveform="sampled", n_samples=STEPS)
    obs = _obs(len(probes), STEPS, seed=13)
    _, g = _grad("f64", src, samples=samples, obs=obs)
    assert g["d_samples"].shape == (STEPS,)
    for k in (3, 17, 40):
        def loss(x):
            sx = samples.copy()
            sx[k] += x
            return _loss_gpu("f64", _copy(src), sx, obs=obs)[1]
        fd = _fd4(loss, 1e-2)
        print(f"[adjoint sampled] k={k}: adjoint {g['d_samples'][k]:.10e} fd {fd:.10e}", flush=True)
        assert _rel(g["d_samples"][k], fd) < TOL_FD64
    # samples after the last probe-affecting step have no influence on the loss
    assert g["d_samples"][-1] == 0.0


def test_adjoint_segments_do_not_change_the_gradient(gpu):
    cfg, solid, src, _, probes = scenes.canyon_scene("f64")
    obs = _obs(len(probes), STEPS)
    _, g1 = _grad("f64", src, every=7, obs=obs)
    _, g2 = _grad("f64", src, every=STEPS, obs=obs)
    # the device work is identical; only the order of the FP64 host sums differs
    assert np.allclose(g1["d_position"], g2["d_position"], rtol=1e-12, atol=0.0)
    assert _rel(g1["d_amp"], g2["d_amp"]) < 1e-12


def test_adjoint_fp32_agrees_with_fp64(gpu):
    cfg, solid, src, _, probes = scenes.canyon_scene("f32")
    obs = _obs(len(probes), STEPS)
    _, g32 = _grad("f32", src, obs=obs)
    _, g64 = _grad("f64", src, obs=obs)
    assert np.linalg.norm(g32["d_position"] - g64["d_position"]) < TOL_F32 * np.linalg.norm(g64["d_position"])
    assert _rel(g32["d_amp"], g64["d_amp"]) < TOL_F32


def test_adjoint_through_nscbc_faces_matches_central_differences(gpu):
    """The reference's open boundary (NSCBC ghosts on x-, x+, y-, y+, z+, 05:684-835) with no
    sponge: the waves leave through the ghosts and two probes sit next to open faces, so the
    transposed ghost fill carries much of the gradient.  FP64 adjoint against fourth-order
    central differences of the GPU forward and of the oracle forward; FP32 against FP64."""
    cfg, solid, src, _, probes = scenes.canyon_scene("f64", open_faces=True)
    obs = _obs(len(probes), STEPS, seed=21)
    L, g = _grad("f64", src, obs=obs, open_faces=True)
    assert L > 0.0
    h = 1e-3
    fd = np.zeros(3)
    for ax in range(3):
        def loss(x):
            d = [0.0, 0.0, 0.0]
            d[ax] = x
            return _loss_gpu("f64", _shifted(src, d), obs=obs, open_faces=True)[1]
        fd[ax] = _fd4(loss, h)
        print(f"[adjoint nscbc] axis {ax}: adjoint {g['d_position'][ax]:.10e} fd {fd[ax]:.10e}", flush=True)
    assert np.linalg.norm(g["d_position"] - fd) < TOL_FD64 * np.linalg.norm(fd)
    ha = 1e-2 * src.amp
    fd = _fd4(lambda x: _loss_gpu("f64", _copy(src, amp=src.amp + x), obs=obs, open_faces=True)[1], ha)
    print(f"[adjoint nscbc] amp: adjoint {g['d_amp']:.10e} fd {fd:.10e}", flush=True)
    assert _rel(g["d_amp"], fd) < TOL_FD64
    fdo = _fd4(lambda x: _loss_oracle(_shifted(src, [0.0, x, 0.0]), obs=obs, open_faces=True), h)
    assert _rel(g["d_position"][1], fdo) < TOL_FD64
    _, g32 = _grad("f32", src, obs=obs, open_faces=True)
    assert np.linalg.norm(g32["d_position"] - g["d_position"]) < TOL_F32 * np.linalg.norm(g["d_position"])


def test_adjoint_rejects_unsupported_setups(gpu):
    # NSCBC faces with solid cells on their planes (tracked boundary solids)
    cfg, solid, src, probes = scenes.box_scene("f64")
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_source(src)
    e.set_probes(probes)
    with pytest.raises(nlos.EngineError):
        e.record_forward(4, 2)
    e.close()
    cfg, solid, src, _, probes = scenes.canyon_scene("f64")
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_source(src)
    e.set_probes(probes)
    e.step(3)
    with pytest.raises(nlos.EngineError):
        e.backward(np.zeros((len(probes), 3)))
    e.close()
