"""Field and probe output formats beyond the reference's own: legacy VTK BINARY, the field
dump off the step stream, and the family-A probe_<name>.dat text (01:481-495).

The ASCII VTK text itself is pinned against the reference's frames in test_gpu_parity.py;
here every other encoding is checked against it or against read_field/read_probes.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import nlos_engine as nlos
import scenes

pytestmark = pytest.mark.gpu


def _engine(precision="f32", **kw):
    cfg, solid, src, probes = scenes.box_scene(precision, **kw)
    e = nlos.Engine(cfg)
    e.set_mask(solid)
    e.set_source(src)
    e.set_probes(probes, [f"P{i}" for i in range(len(probes))])
    return e


def _split_vtk(raw: bytes):
    """Header lines (10 of them) and the data bytes of a legacy VTK file."""
    pos = 0
    for _ in range(10):
        pos = raw.index(b"\n", pos) + 1
    return raw[:pos].decode().split("\n")[:-1], raw[pos:]


def test_binary_vtk_holds_the_captured_field(gpu, tmp_path):
    e = _engine()
    e.step(30, capture_field=True)
    e.write_vtk(str(tmp_path / "a.vtk"))
    e.write_vtk(str(tmp_path / "b.vtk"), binary=True)
    ha, _ = _split_vtk((tmp_path / "a.vtk").read_bytes())
    hb, data = _split_vtk((tmp_path / "b.vtk").read_bytes())
    assert ha[2] == "ASCII" and hb[2] == "BINARY"
    assert ha[:2] == hb[:2] and ha[3:] == hb[3:]
    field = e.read_field()
    assert len(data) == field.size * 4 + 1 and data[-1:] == b"\n"
    got = np.frombuffer(data[:-1], ">f4").reshape(field.shape)
    assert np.array_equal(got, field)
    assert np.abs(field).max() > 0


@pytest.mark.parametrize("binary", [False, True])
def test_async_dump_equals_synchronous_write(gpu, tmp_path, binary):
    """A dump runs beside later steps and later captures; its file equals the synchronous
    writer's file of the same captured step."""
    e = _engine()
    e.step(12, capture_field=True)
    e.write_vtk(str(tmp_path / "sync0.vtk"), binary=binary)
    e.dump_field_async(str(tmp_path / "async0.vtk"), binary=binary)
    e.step(9)                          # runs while the dump is in flight
    e.step(1, capture_field=True)      # waits on the device for the dump's copy
    e.write_vtk(str(tmp_path / "sync1.vtk"), binary=binary)
    e.dump_field_async(str(tmp_path / "async1.vtk"), binary=binary)   # waits for dump 0
    e.dump_wait()
    for k in (0, 1):
        assert (tmp_path / f"async{k}.vtk").read_bytes() == (tmp_path / f"sync{k}.vtk").read_bytes()
    assert (tmp_path / "async0.vtk").read_bytes() != (tmp_path / "async1.vtk").read_bytes()


def test_async_dump_errors(gpu, tmp_path):
    e = _engine()
    e.step(3)
    with pytest.raises(nlos.EngineError) as ex:      # nothing captured yet
        e.dump_field_async(str(tmp_path / "x.vtk"))
    assert ex.value.code == 3
    e.step(1, capture_field=True)
    e.dump_field_async(str(tmp_path / "no_such_dir" / "x.vtk"))
    with pytest.raises(nlos.EngineError) as ex:      # the writer's error comes back here
        e.dump_wait()
    assert ex.value.code == 5
    e.dump_wait()                                     # reported once


def test_run_with_async_frames_writes_the_same_files(gpu, tmp_path):
    out = {}
    for mode in ("sync", "async"):
        e = _engine()
        e.run(24 * e.dt, 4, vtk_dir=str(tmp_path / mode), probe_dir=str(tmp_path / mode),
              async_frames=(mode == "async"))
        e.close()
        out[mode] = {n: (tmp_path / mode / n).read_bytes() for n in sorted(os.listdir(tmp_path / mode))}
    assert list(out["sync"]) == list(out["async"])
    assert any(n.endswith(".vtk") for n in out["sync"])
    for n in out["sync"]:
        assert out["sync"][n] == out["async"][n], n


def test_probe_dat_family_a_format(gpu, tmp_path):
    e = _engine()
    e.step(17)
    e.write_probe_dat(str(tmp_path), scale=100.0)
    p = e.read_probes()
    dt = e.dt
    for i, name in enumerate(e.probe_names):
        lines = (tmp_path / f"probe_{name}.dat").read_text().split("\n")
        assert lines[0] == "# Time Pressure" and lines[-1] == ""
        body = lines[1:-1]
        assert len(body) == p.shape[1] == 17
        for s, line in enumerate(body):
            assert line == "%g %g" % ((s + 1) * dt, p[i, s] * 100.0)
