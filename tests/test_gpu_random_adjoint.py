"""The discrete adjoint on seeded scenes of random shape and options (random sponge and NSCBC
faces, ground, wall reflection, sky height, injection and waveform), from rest:
the FP64 gradient of a probe misfit with respect to the source position, the amplitude and
the samples of a sampled waveform against fourth-order central differences of the GPU
forward, and the FP32 gradient against the FP64 one.  The source stands still, away from
the cell centres and stencil boundaries, so the loss is smooth in its position, and the
probes sit a few cells from it."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu

TOL_FD64 = 1e-5
TOL_F32 = 2e-3
FACE_PLANES = ((nlos.FACE_XM, (slice(None), slice(None), 0)), (nlos.FACE_XP, (slice(None), slice(None), -1)),
               (nlos.FACE_YM, (slice(None), 0, slice(None))), (nlos.FACE_YP, (slice(None), -1, slice(None))),
               (nlos.FACE_ZM, (0, slice(None), slice(None))), (nlos.FACE_ZP, (-1, slice(None), slice(None))))


def _scene(seed, precision):
    cfg, solid, src, samples, probes, _, steps = scenes.random_scene(seed, precision)
    # from rest: an initial density field would dominate the probes, and the source's share
    # of them would then sit close to the round-off of the difference quotients
    rho = None
    cfg.stream_pattern = nlos.PATTERN_TWO_ARRAY
    for face, plane in FACE_PLANES:          # no solid cell on an open boundary plane
        if cfg.nscbc_faces & face:
            solid[plane] = 0
    rng = np.random.default_rng(77 + seed)
    dx = cfg.dx
    # a still source a quarter of a cell off the cell centres in every axis
    pos = []
    for lo, n in ((cfg.x_min, solid.shape[2]), (cfg.y_min, solid.shape[1]), (cfg.z_min, solid.shape[0])):
        c = int(rng.integers(0, n))
        pos.append(lo + (c + 0.5 + rng.choice([-0.25, 0.25])) * dx)
    src.sx, src.sy, src.sz = pos
    src.ex, src.ey, src.ez = pos
    src.speed = 0.0
    # probes 2-8 cells from the source (inside the domain), so that they see the wave itself
    # rather than its precursor at round-off level, which no difference quotient resolves
    lo = np.array([cfg.x_min, cfg.y_min, cfg.z_min])
    hi = lo + dx * np.array(solid.shape[::-1])
    n_probes = max(len(probes), 2)
    probes = []
    for _ in range(n_probes):
        d = rng.normal(size=3)
        p = np.array(pos) + rng.uniform(2.0, 8.0) * dx * d / np.linalg.norm(d)
        probes.append(tuple(np.clip(p, lo + 0.05 * dx, hi - 0.05 * dx)))
    return cfg, solid, src, samples, probes, rho, steps


def _copy(src, **kw):
    s = nlos.Source()
    C.pointer(s)[0] = src
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _run(cfg, solid, src, samples, probes, rho, steps, obs, record=False):
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    if rho is not None:
        e.set_density(rho)
    e.set_source(src, samples)
    e.set_probes(probes)
    if record:
        e.record_forward(steps, 9)
    else:
        e.step(steps)
    r = e.read_probes() - obs
    return e, float(np.sum(r * r)), 2.0 * r


def _fd4(loss, h):
    return (8.0 * (loss(h) - loss(-h)) - (loss(2 * h) - loss(-2 * h))) / (12.0 * h)


@pytest.mark.parametrize("seed", list(range(16)))
def test_random_scene_adjoint_matches_central_differences(gpu, seed):
    cfg, solid, src, samples, probes, rho, steps = _scene(seed, "f64")
    e, _, _ = _run(cfg, solid, src, samples, probes, rho, steps, 0.0)
    signal = e.read_probes()
    assert e.info().n_solid_boundary == 0 or cfg.nscbc_faces == 0
    e.close()
    if np.abs(signal).max() < 1e-9:
        pytest.skip("the wave reaches no probe within the steps")
    obs = 0.5 * signal + 0.2 * np.abs(signal).max() * np.random.default_rng(seed).standard_normal(signal.shape)
    print(f"[random adjoint {seed}] grid {solid.shape[::-1]}, faces sponge {cfg.sponge_faces} "
          f"nscbc {cfg.nscbc_faces}, ground {cfg.ground}, reflect {cfg.wall_reflect:.3f}, "
          f"waveform {src.waveform}, injection {src.injection}, steps {steps}", flush=True)
    n_s = 0 if samples is None else len(samples)
    e, L, dl = _run(cfg, solid, src, samples, probes, rho, steps, obs, record=True)
    g = e.backward(dl, n_samples=n_s)
    e.close()
    assert L > 0.0
    loss = lambda s, smp=samples: _run(cfg, solid, s, smp, probes, rho, steps, obs)[1]
    h = 1e-2 * cfg.dx
    fd = np.array([_fd4(lambda x, a=a: loss(_copy(src, **{a: getattr(src, a) + x,
                                                          "e" + a[1]: getattr(src, "e" + a[1]) + x})), h)
                   for a in ("sx", "sy", "sz")])
    print(f"  position adjoint {g['d_position']} fd {fd}", flush=True)
    assert np.linalg.norm(g["d_position"] - fd) <= TOL_FD64 * np.linalg.norm(fd)
    ha = 1e-2 * src.amp
    fda = _fd4(lambda x: loss(_copy(src, amp=src.amp + x)), ha)
    print(f"  amp adjoint {g['d_amp']:.10e} fd {fda:.10e}", flush=True)
    assert abs(g["d_amp"] - fda) <= TOL_FD64 * abs(fda)
    if n_s:
        ref = np.abs(g["d_samples"]).max()
        for k in np.argsort(-np.abs(g["d_samples"]))[:3]:
            def lk(x, k=k):
                s2 = samples.copy()
                s2[k] += x
                return loss(_copy(src), s2)
            fdk = _fd4(lk, 1e-2)
            assert abs(g["d_samples"][k] - fdk) <= TOL_FD64 * ref, (k, g["d_samples"][k], fdk)
    # FP32 against FP64
    cfg32, _, _, _, _, _, _ = _scene(seed, "f32")
    e, _, dl32 = _run(cfg32, solid, src, samples, probes, rho, steps, obs, record=True)
    g32 = e.backward(dl32, n_samples=n_s)
    e.close()
    assert np.linalg.norm(g32["d_position"] - g["d_position"]) <= TOL_F32 * np.linalg.norm(g["d_position"])
    assert abs(g32["d_amp"] - g["d_amp"]) <= TOL_F32 * abs(g["d_amp"])
