"""CPU checks of the C-ABI library: it loads, exports every entry point the headers
declare, and its default configuration is the reference's (no GPU calls)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

import nlos_engine as nlos

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in ("nlos_engine.h", "nlos_engine_test.h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(nlos_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = nlos.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", nlos.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    for s in decl:
        assert getattr(lib, s) is not None
    assert set(nlos.SIGNATURES) == set(decl)


def test_abi_version_and_struct_sizes():
    lib = nlos.load_library()
    assert lib.nlos_abi_version() == 1
    # nlos_config: 16 doubles + 12 int32; nlos_source: 13 doubles, 4 int32, 1 pointer
    assert C.sizeof(nlos.Config) == 16 * 8 + 12 * 4
    assert C.sizeof(nlos.Source) == 13 * 8 + 4 * 4 + 8


def test_reference_defaults():
    cfg = nlos.default_config()
    assert (cfg.x_min, cfg.x_max, cfg.y_min, cfg.y_max, cfg.z_min, cfg.z_max) == (0, 1650, 0, 1150, 0, 250)
    assert cfg.dx == 2.0 and cfg.c_phys == 340.0 and cfg.tau == 0.55
    assert cfg.sponge_width == 40.0 and cfg.sponge_strength == 0.35 and cfg.sky_z == 120.0
    assert cfg.sponge_faces == nlos.FACES_REFERENCE == cfg.nscbc_faces
    assert cfg.ground == 1 and cfg.use_geometry == 1 and cfg.output_world_coords == 1
    src = nlos.default_source()
    assert (src.t0, src.sigma_t, src.gain, src.stencil_radius, src.stencil_sigma_cells) == (0.05, 0.005, 1.0, 2, 1.25)


def test_grid_and_dt_rules():
    cfg = nlos.reference05_config()
    assert nlos.grid_dims(cfg) == (264, 198, 71)          # 05:332-334 at DX=1.5
    assert abs(nlos.dt_of(cfg) - 0.00254713) < 1e-8        # printed by the reference run
    assert int(round(1.0 / nlos.dt_of(cfg))) == 393


def test_missing_library_fails_loudly(tmp_path):
    saved = nlos._lib
    nlos._lib = None
    try:
        with pytest.raises(RuntimeError):
            nlos.load_library(str(tmp_path / "missing.so"))
    finally:
        nlos._lib = saved
