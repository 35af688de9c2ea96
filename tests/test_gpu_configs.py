"""The five benchmark configurations at their stated sizes (the other GPU tests use small
scenes of the same construction).

* C1, rigid-cube scattering 128^3 in FP64: all 2000 steps bit for bit against the oracle.
* C2, the moving-source city 1024x1024x256: the precision contract.  Over 2000 steps the
  probe traces of the 16-bit storage the benchmark headlines (B16) stay within relative L2
  1e-3 of FP64 (the FP64 path is bit-identical to the oracle on every smaller scene); FP32
  within 1e-4.  FP16S is run and its error printed: it is not held to 1e-3 and the
  benchmark does not headline it unless it meets the bound.
* C3, the dx = 0.5 m city 2048x2048x512 (2.1e9 cells): size-independent properties (the
  rest state is exact in FP16S and B16; FP16S in place equals FP16S two-array bit for bit
  with the moving source), and the per-probe error of B16 against FP64 on the 512x512x256
  sub-box over the configuration's 4000 steps.
* C4, 256 batched scenes over 3000 steps: a sample of scenes bit for bit against the same
  scenes stepped alone.
* C5, the 256x256x128 canyon: the FP32 adjoint gradient against fourth-order central
  differences of the FP64 forward, relative 1e-4, on a 600-step horizon.

Each test prints what it runs before it runs it.
"""
from __future__ import annotations

import numpy as np
import pytest

import nlos_engine as nlos
import workloads
from oracle_ffi import Oracle
from scenes import rel_l2

pytestmark = pytest.mark.gpu

ERROR_BOUND = 1e-3      # probe rel-L2 of the headline 16-bit storage against FP64
TOL_F32 = 1e-4
TOL_GRAD = 1e-4


def _say(msg):
    print(msg, flush=True)


def test_c1_scatter_128_fp64_bitwise_against_oracle_2000_steps(gpu):
    wl = workloads.ScatterWorkload()
    steps = 2000
    mask, rho, probes = wl.mask(), wl.density(), wl.probes()
    _say(f"[C1] 128^3 FP64, {steps} steps on the GPU")
    e = nlos.Engine(wl.config("f64"))
    e.set_mask(mask)
    e.set_density(rho)
    e.set_probes(probes)
    e.step(steps, capture_field=True)
    e.synchronize()
    got, field = e.read_probes(), e.read_field()
    e.close()
    _say(f"[C1] oracle, {steps} steps on the host cores")
    o = Oracle(wl.config("f64"))
    o.set_mask(mask)
    o.set_density(rho)
    o.set_probes(probes)
    o.run(steps)
    want = o.read_probes()
    assert got.shape == want.shape == (8, steps)
    assert np.abs(want).max() > 1e-7          # the scattered pulse reached the probes
    assert np.array_equal(got, want), float(np.abs(got - want).max())
    assert np.array_equal(field, o.read_field())


# ---- C2 ------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def city():
    wl = workloads.CityWorkload()
    return wl, wl.stl()


def _city_trace(wl, stl, precision, steps, pattern="two_array"):
    cfg = wl.config(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = nlos.Engine(cfg)
    e.voxelize_stl(stl)
    e.set_probes(wl.probes(e.ground_heights()))
    e.set_source(wl.source())
    e.step(steps)
    e.synchronize()
    p = e.read_probes()
    e.close()
    return p


def test_c2_precision_contract_2000_steps(gpu, city):
    wl, stl = city
    steps = 2000
    tr = {}
    for p in ("f64", "b16", "f32", "f16s"):
        _say(f"[C2] 1024x1024x256 city, {p}, {steps} steps")
        tr[p] = _city_trace(wl, stl, p, steps)
    ref = tr["f64"]
    assert np.abs(ref).max() > 1e-9
    err = {p: rel_l2(tr[p], ref) for p in ("b16", "f32", "f16s")}
    worst = {p: max(rel_l2(a, b) for a, b in zip(tr[p], ref)) for p in err}
    _say(f"[C2] probe rel-L2 vs FP64: {err}; worst probe: {worst}")
    assert err["b16"] <= ERROR_BOUND
    assert err["f32"] <= TOL_F32
    assert err["f16s"] > 0.0


# ---- C3 ------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def city2048():
    wl = workloads.CityWorkload(nx=2048, ny=2048, nz=512, dx=0.5, sponge_cells=64, n_probes=128)
    return wl, wl.stl()


def _c3_run(wl, stl, precision, pattern, steps, source=True):
    cfg = wl.config(precision)
    cfg.stream_pattern = nlos.PATTERNS[pattern]
    e = nlos.Engine(cfg)
    e.voxelize_stl(stl)
    e.set_probes(wl.probes(e.ground_heights()))
    if source:
        e.set_source(wl.source())
    e.step(steps, capture_field=True)
    e.synchronize()
    out = (e.read_probes(), e.read_field())
    e.close()
    return out


@pytest.mark.parametrize("precision", ["f16s", "b16"])
def test_c3_rest_state_is_exact_at_2048x2048x512(gpu, city2048, precision):
    wl, stl = city2048
    _say(f"[C3] 2048x2048x512 rest state, {precision} two-array, 8 steps")
    probes, field = _c3_run(wl, stl, precision, "two_array", 8, source=False)
    assert probes.shape == (128, 8)
    assert not np.any(probes)
    assert not np.any(field)


def test_c3_fp16_inplace_equals_two_array_at_2048x2048x512(gpu, city2048):
    """The configuration names the A-A pattern; at FP16 two arrays also fit, and both
    patterns give the same bits (two kernels and two layouts of the same step)."""
    wl, stl = city2048
    runs = []
    for pattern in ("two_array", "inplace"):
        _say(f"[C3] 2048x2048x512 f16s {pattern}, 12 steps with the moving source")
        runs.append(_c3_run(wl, stl, "f16s", pattern, 12))
    assert np.abs(runs[0][1]).max() > 0
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_c3_subbox_b16_error_4000_steps(gpu):
    sub = workloads.CityWorkload(nx=512, ny=512, nz=256, dx=0.5, sponge_cells=64, n_probes=128)
    stl = sub.stl()
    steps = 4000
    tr = {}
    for p in ("f64", "b16"):
        _say(f"[C3] 512x512x256 sub-box, {p}, {steps} steps")
        tr[p] = _city_trace(sub, stl, p, steps, pattern="auto")
    ref = tr["f64"]
    per = np.array([rel_l2(a, b) for a, b in zip(tr["b16"], ref) if np.abs(b).max() > 0])
    err = rel_l2(tr["b16"], ref)
    _say(f"[C3] B16 sub-box rel-L2 {err:.3e}, per probe median {np.median(per):.3e} max {per.max():.3e}")
    assert len(per) == 128
    assert err <= ERROR_BOUND


# ---- C4 ------------------------------------------------------------------------------------

def test_c4_256_batched_scenes_3000_steps_equal_scenes_alone(gpu):
    wl = workloads.BatchWorkload()
    defs = wl.scenes()
    steps = 3000
    _say(f"[C4] {len(defs)} scenes, {sum(np.prod(d.dims) for d in defs)} cells, {steps} steps batched")
    b = nlos.Batch([d.cfg for d in defs])
    for i, d in enumerate(defs):
        b.scene(i).set_source(d.source)
        b.scene(i).set_probes(d.probes)
    b.autotune(reps=1)
    b.step(steps)
    check = [0, 1, 77, 128, 201, 255]
    got = {i: b.scene(i).read_probes() for i in check}
    b.synchronize()
    b.close()
    for i in check:
        _say(f"[C4] scene {i} {defs[i].dims} stepped alone")
        e = nlos.Engine(defs[i].cfg)
        e.set_source(defs[i].source)
        e.set_probes(defs[i].probes)
        e.step(steps)
        want = e.read_probes()
        e.close()
        assert got[i].shape == want.shape == (4, steps)
        assert np.abs(want).max() > 0
        assert np.array_equal(got[i], want), i


# ---- C5 ------------------------------------------------------------------------------------

def test_c5_canyon_adjoint_against_fp64_fd(gpu):
    steps, every = 600, 100
    wl = workloads.CanyonWorkload(n_samples=steps)
    stl = workloads.stl_bytes(wl.boxes())
    samples = wl.samples()
    probes = wl.probes()

    def engine(precision, src, samp=samples):
        e = nlos.Engine(wl.config(precision))
        e.voxelize_stl(stl)
        e.set_source(src, samp)
        e.set_probes(probes)
        return e

    rng = np.random.Generator(np.random.PCG64(5))
    _say(f"[C5] canyon 256x256x128: observation, {steps} steps")
    e = engine("f32", wl.source(x=0.41 * wl.nx + 4.13, y=wl.ny / 2.0 - 1.37, z=5.21, amp=1.3))
    e.step(steps)
    clean = e.read_probes()
    e.close()
    obs = clean + 0.01 * np.sqrt(np.mean(clean ** 2)) * rng.standard_normal(clean.shape)
    src = wl.source()

    def loss(precision, s=None, samp=samples):
        e = engine(precision, s if s is not None else src, samp)
        e.step(steps)
        r = e.read_probes() - obs
        e.close()
        return float(np.sum(r * r))

    _say("[C5] FP32 recorded forward + backward")
    e = engine("f32", src)
    e.record_forward(steps, every)
    r = e.read_probes() - obs
    g = e.backward(2.0 * r, n_samples=steps)
    e.close()

    def fd4(fun, h):
        return (8.0 * (fun(h) - fun(-h)) - (fun(2 * h) - fun(-2 * h))) / (12.0 * h)

    def moved(d=(0.0, 0.0, 0.0), amp=None):
        return wl.source(x=src.sx + d[0], y=src.sy + d[1], z=src.sz + d[2],
                         amp=src.amp if amp is None else amp)

    _say("[C5] FP64 fourth-order central differences")
    fdp = []
    for ax in range(3):
        def f(x, ax=ax):
            d = [0.0, 0.0, 0.0]
            d[ax] = x
            return loss("f64", moved(d))
        fdp.append(fd4(f, 1e-3))
    fda = fd4(lambda x: loss("f64", moved(amp=src.amp + x)), 1e-2 * src.amp)
    k = int(np.argmax(np.abs(g["d_samples"])))

    def fs(x):
        sx = samples.copy()
        sx[k] += x
        return loss("f64", samp=sx)
    fds = fd4(fs, 1e-2)
    gp, fp = np.asarray(g["d_position"]), np.asarray(fdp)
    errs = {"position": float(np.linalg.norm(gp - fp) / np.linalg.norm(fp)),
            "amp": abs(g["d_amp"] - fda) / abs(fda),
            "sample": abs(g["d_samples"][k] - fds) / abs(fds)}
    _say(f"[C5] adjoint vs FP64 FD relative errors {errs}")
    assert all(v <= TOL_GRAD for v in errs.values()), errs
