"""ctypes binding of the engine's C ABI (include/nlos_engine.h).

`Engine` mirrors the reference solver's surface (05_sensor_illusion.cpp:241-1122): the
configuration fields of `LBMSolver` (05:243-268) become `Config`, `DroneSource`
(05:211-239) and the waveform constants become `Source`, and the methods `apply_source`,
`exchange_halos`, `step_collision_bgk`, `write_vtk` and `run` keep their names and
meaning.  The library must be built (`make -C city-scale-nlos-acoustic-engine-code_amd`);
there is no fallback: importing without it raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Iterable, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NLOS_ENGINE_LIB") or os.path.join(HERE, "libnlos_engine.so")

F64, F32, F16S, B16 = 0, 1, 2, 3
VTK_ASCII, VTK_BINARY = 0, 1
VOXEL_HEIGHT, VOXEL_PARITY, VOXEL_WINDING, VOXEL_3AXIS = 0, 1, 2, 4
FAMILY_B, FAMILY_A = 0, 1
PRECISIONS = {"f64": F64, "f32": F32, "f16s": F16S, "b16": B16}
FACE_XM, FACE_XP, FACE_YM, FACE_YP, FACE_ZM, FACE_ZP = 1, 2, 4, 8, 16, 32
FACES_REFERENCE = FACE_XM | FACE_XP | FACE_YM | FACE_YP | FACE_ZP
FACES_ALL = FACES_REFERENCE | FACE_ZM
WAVE_GAUSS_PULSE, WAVE_SINE, WAVE_SAMPLED, WAVE_RICKER = 0, 1, 2, 3
INJECT_GAUSS5, INJECT_TRILINEAR = 0, 1
STEP_CAPTURE_FIELD = 1
PATTERN_AUTO, PATTERN_INPLACE, PATTERN_TWO_ARRAY = 0, 1, 2
PATTERNS = {"auto": PATTERN_AUTO, "inplace": PATTERN_INPLACE, "two_array": PATTERN_TWO_ARRAY}

ERR_NAMES = {1: "NLOS_ERR_ARG", 2: "NLOS_ERR_HIP", 3: "NLOS_ERR_STATE", 4: "NLOS_ERR_NONFINITE",
             5: "NLOS_ERR_IO", 6: "NLOS_ERR_NOMEM", 7: "NLOS_ERR_UNSUPPORTED"}


class Config(C.Structure):
    _fields_ = [("x_min", C.c_double), ("x_max", C.c_double), ("y_min", C.c_double),
                ("y_max", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
                ("dx", C.c_double), ("c_phys", C.c_double), ("tau", C.c_double),
                ("sponge_width", C.c_double), ("sponge_strength", C.c_double),
                ("sky_z", C.c_double), ("wall_reflect", C.c_double),
                ("world_offset", C.c_double * 3),
                ("sponge_faces", C.c_int32), ("nscbc_faces", C.c_int32), ("ground", C.c_int32),
                ("use_geometry", C.c_int32), ("output_world_coords", C.c_int32),
                ("precision", C.c_int32), ("f16_scale_log2", C.c_int32),
                ("stream_pattern", C.c_int32), ("voxelizer", C.c_int32), ("family", C.c_int32),
                ("reserved", C.c_int32 * 2)]


class Source(C.Structure):
    _fields_ = [("sx", C.c_double), ("sy", C.c_double), ("sz", C.c_double),
                ("ex", C.c_double), ("ey", C.c_double), ("ez", C.c_double),
                ("speed", C.c_double), ("amp", C.c_double), ("freq", C.c_double),
                ("t0", C.c_double), ("sigma_t", C.c_double), ("gain", C.c_double),
                ("stencil_sigma_cells", C.c_double), ("stencil_radius", C.c_int32),
                ("waveform", C.c_int32), ("injection", C.c_int32), ("n_samples", C.c_int32),
                ("samples", C.POINTER(C.c_double))]


class Info(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("precision", C.c_int32),
                ("dt", C.c_double), ("stream_pattern", C.c_int32), ("reserved_info", C.c_int32),
                ("step", C.c_int64), ("n_probes", C.c_int64),
                ("n_solid_boundary", C.c_int64), ("bytes_per_update", C.c_double),
                ("device_bytes", C.c_size_t)]


class Timing(C.Structure):
    _fields_ = [("launches", C.c_int64), ("total_ms", C.c_double), ("avg_ms", C.c_double)]


class Grad(C.Structure):
    _fields_ = [("d_position", C.c_double * 3), ("d_amp", C.c_double),
                ("d_samples", C.POINTER(C.c_double)), ("n_samples", C.c_int64)]


class LineSource(C.Structure):
    """nlos_line_source: 04_hpc_scalability's line source (family A)."""
    _fields_ = [("x0", C.c_double), ("y0", C.c_double), ("amp", C.c_double), ("b", C.c_double),
                ("omega", C.c_double), ("r2_max", C.c_double), ("dt", C.c_double), ("step0", C.c_int64)]


def reference04_line_source(step0: int = 0) -> LineSource:
    """04:65-72 and 04:32-37: Amp 0.01, b 2, omega pi/15, centre (0, -6), cut-off 36 m^2,
    dt = 0.577350269 dx / c with dx = 0.125, c = 1."""
    import math
    return LineSource(x0=0.0, y0=-6.0, amp=0.01, b=2.0, omega=math.pi / 15.0, r2_max=36.0,
                      dt=0.577350269 * 0.125 / 1.0, step0=step0)


class Tune(C.Structure):
    _fields_ = [("block_x", C.c_int32), ("block_y", C.c_int32), ("candidates", C.c_int32),
                ("per_scene", C.c_int32), ("best_ms", C.c_double)]


# Every C-ABI entry point of include/nlos_engine.h (plus the test hooks of
# include/nlos_engine_test.h) with its ctypes signature.
_P, _E = C.c_void_p, C.c_void_p
SIGNATURES = {
    "nlos_abi_version": (C.c_int, []),
    "nlos_last_error": (C.c_char_p, []),
    "nlos_config_default": (C.c_int, [C.POINTER(Config)]),
    "nlos_source_default": (C.c_int, [C.POINTER(Source)]),
    "nlos_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "nlos_destroy": (C.c_int, [_E]),
    "nlos_get_info": (C.c_int, [_E, C.POINTER(Info)]),
    "nlos_load_stl": (C.c_int, [_E, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double]),
    "nlos_voxelize_stl": (C.c_int, [_E, _P, C.c_size_t, C.c_double, C.c_double, C.c_double, C.c_double]),
    "nlos_get_ground_heights": (C.c_int, [_E, _P]),
    "nlos_set_mask": (C.c_int, [_E, _P]),
    "nlos_get_mask": (C.c_int, [_E, _P]),
    "nlos_set_density": (C.c_int, [_E, _P]),
    "nlos_set_sponge": (C.c_int, [_E, _P]),
    "nlos_set_line_source": (C.c_int, [_E, C.POINTER(LineSource)]),
    "nlos_set_source": (C.c_int, [_E, C.POINTER(Source)]),
    "nlos_set_probes": (C.c_int, [_E, _P, C.c_int32, _P]),
    "nlos_apply_source": (C.c_int, [_E, C.c_double]),
    "nlos_exchange_halos": (C.c_int, [_E]),
    "nlos_step_collision": (C.c_int, [_E]),
    "nlos_step": (C.c_int, [_E, C.c_int64, C.c_int32]),
    "nlos_step_noalloc": (C.c_int, [_E, C.c_int64, C.c_int32]),
    "nlos_synchronize": (C.c_int, [_E]),
    "nlos_read_probes": (C.c_int, [_E, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "nlos_read_field": (C.c_int, [_E, _P]),
    "nlos_write_probe_files": (C.c_int, [_E, C.c_char_p, C.POINTER(C.c_char_p)]),
    "nlos_write_vtk": (C.c_int, [_E, C.c_char_p]),
    "nlos_write_vtk_binary": (C.c_int, [_E, C.c_char_p]),
    "nlos_dump_field_async": (C.c_int, [_E, C.c_char_p, C.c_int32]),
    "nlos_dump_wait": (C.c_int, [_E]),
    "nlos_write_probe_dat": (C.c_int, [_E, C.c_char_p, C.POINTER(C.c_char_p), C.c_double]),
    "nlos_checkpoint": (C.c_int, [_E, _P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "nlos_restore": (C.c_int, [_E, _P, C.c_size_t]),
    "nlos_timing_enable": (C.c_int, [_E, C.c_int32]),
    "nlos_timing_reset": (C.c_int, [_E]),
    "nlos_timing_get": (C.c_int, [_E, C.POINTER(Timing)]),
    "nlos_record_forward": (C.c_int, [_E, C.c_int64, C.c_int32]),
    "nlos_backward": (C.c_int, [_E, _P, C.c_int64, C.POINTER(Grad)]),
    "nlos_batch_create": (C.c_int, [C.POINTER(Config), C.c_int32, C.POINTER(C.c_void_p)]),
    "nlos_batch_destroy": (C.c_int, [_P]),
    "nlos_batch_size": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "nlos_batch_scene": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_void_p)]),
    "nlos_batch_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(Tune)]),
    "nlos_batch_tile_widths": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32]),
    "nlos_batch_set_tile_widths": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32]),
    "nlos_batch_step": (C.c_int, [_P, C.c_int64]),
    "nlos_batch_synchronize": (C.c_int, [_P]),
    "nlos_batch_timing_enable": (C.c_int, [_P, C.c_int32]),
    "nlos_batch_timing_get": (C.c_int, [_P, C.POINTER(Timing)]),
    "nlos_debug_read_f": (C.c_int, [_E, _P]),
    "nlos_debug_poke": (C.c_int, [_E, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double]),
    "nlos_debug_copy_kernel": (C.c_int, [C.c_size_t, C.c_int32, C.POINTER(C.c_double)]),
    "nlos_debug_pull_pattern": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
}

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Loads the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"engine library missing: {path} (build it with __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != 0:
        raise EngineError(rc, load_library().nlos_last_error().decode())


def default_config(**kw) -> Config:
    """05's member defaults (05:243-268, 455-458) with keyword overrides.  A keyword
    `precision` may be "f64", "f32", "f16s" or "b16"."""
    cfg = Config()
    _check(load_library().nlos_config_default(C.byref(cfg)))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        if k == "stream_pattern" and isinstance(v, str):
            v = PATTERNS[v]
        if k == "world_offset":
            cfg.world_offset[:] = list(v)
            continue
        setattr(cfg, k, v)
    return cfg


def default_source(**kw) -> Source:
    src = Source()
    _check(load_library().nlos_source_default(C.byref(src)))
    for k, v in kw.items():
        setattr(src, k, v)
    return src


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class Engine:
    """One simulation domain on the current GPU (the reference's `LBMSolver`)."""

    def __init__(self, cfg: Config, _handle: Optional[C.c_void_p] = None):
        self.lib = load_library()
        self.owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            _check(self.lib.nlos_create(C.byref(cfg), C.byref(h)))
        else:
            h = _handle
        self.h = h
        self.cfg = cfg
        self._samples = None
        self.probe_names: List[str] = []

    # -- lifetime --------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            if self.owned:
                self.lib.nlos_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> Info:
        i = Info()
        _check(self.lib.nlos_get_info(self.h, C.byref(i)))
        return i

    @property
    def dims(self):
        i = self.info()
        return (i.nx, i.ny, i.nz)

    @property
    def dt(self) -> float:
        return self.info().dt

    # -- geometry ----------------------------------------------------------------------
    def voxelize_stl(self, data: bytes, scale=1.0, sx=0.0, sy=0.0, sz=0.0) -> None:
        buf = C.create_string_buffer(data, len(data))
        _check(self.lib.nlos_voxelize_stl(self.h, C.cast(buf, C.c_void_p), len(data), scale, sx, sy, sz))

    def load_stl(self, path: str, scale=1.0, sx=0.0, sy=0.0, sz=0.0) -> int:
        """GeometryEngine::load; returns the status (NLOS_ERR_IO keeps empty geometry)."""
        return self.lib.nlos_load_stl(self.h, path.encode(), scale, sx, sy, sz)

    def ground_heights(self) -> np.ndarray:
        nx, ny, _ = self.dims
        out = np.empty((ny + 2, nx + 2), np.float64)
        _check(self.lib.nlos_get_ground_heights(self.h, _ptr(out)))
        return out

    def set_mask(self, solid: np.ndarray) -> None:
        """solid: (nz, ny, nx) array, nonzero = SOLID."""
        m = np.ascontiguousarray(solid, dtype=np.uint8)
        assert m.shape == tuple(reversed(self.dims))
        _check(self.lib.nlos_set_mask(self.h, _ptr(m)))

    def get_mask(self) -> np.ndarray:
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx), np.uint8)
        _check(self.lib.nlos_get_mask(self.h, _ptr(out)))
        return out

    # -- state, source, probes ----------------------------------------------------------
    def set_density(self, rho: np.ndarray) -> None:
        r = np.ascontiguousarray(rho, dtype=np.float64)
        assert r.shape == tuple(reversed(self.dims))
        _check(self.lib.nlos_set_density(self.h, _ptr(r)))

    def set_sponge(self, sigma: np.ndarray) -> None:
        """FAMILY_A only: sponge strength per interior cell."""
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        assert s.shape == tuple(reversed(self.dims))
        _check(self.lib.nlos_set_sponge(self.h, _ptr(s)))

    def set_line_source(self, src: LineSource) -> None:
        """FAMILY_A only: 04's line source (nlos_set_line_source)."""
        _check(self.lib.nlos_set_line_source(self.h, C.byref(src)))

    def set_source(self, src: Source, samples: Optional[np.ndarray] = None) -> None:
        if samples is not None:
            self._samples = np.ascontiguousarray(samples, dtype=np.float64)
            src.samples = self._samples.ctypes.data_as(C.POINTER(C.c_double))
            src.n_samples = len(self._samples)
        _check(self.lib.nlos_set_source(self.h, C.byref(src)))

    def set_probes(self, xyz: Sequence[Sequence[float]], names: Optional[Iterable[str]] = None) -> np.ndarray:
        p = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        in_solid = np.zeros(len(p), np.uint8)
        _check(self.lib.nlos_set_probes(self.h, _ptr(p), len(p), _ptr(in_solid)))
        self.probe_names = list(names) if names is not None else [f"P{i}" for i in range(len(p))]
        return in_solid.astype(bool)

    # -- stepping (LBMSolver methods) --------------------------------------------------------
    def apply_source(self, t: float) -> None:
        _check(self.lib.nlos_apply_source(self.h, t))

    def exchange_halos(self) -> None:
        _check(self.lib.nlos_exchange_halos(self.h))

    def step_collision_bgk(self) -> None:
        _check(self.lib.nlos_step_collision(self.h))

    def step(self, n: int, capture_field: bool = False) -> None:
        _check(self.lib.nlos_step(self.h, n, STEP_CAPTURE_FIELD if capture_field else 0))

    def step_noalloc(self, n: int, capture_field: bool = False) -> None:
        """nlos_step without device allocation (nlos_step_noalloc)."""
        _check(self.lib.nlos_step_noalloc(self.h, n, STEP_CAPTURE_FIELD if capture_field else 0))

    def synchronize(self) -> None:
        _check(self.lib.nlos_synchronize(self.h))

    def run(self, T_total: float, Frames: int, vtk_dir: Optional[str] = None,
            probe_dir: Optional[str] = None, async_frames: bool = True,
            binary_frames: bool = False) -> int:
        """LBMSolver::run (05:1064-1121): total_steps = round(T/dt), steps s = 0..total,
        a VTK frame every out_interval steps, probe files at the end.  With async_frames a
        frame is copied and written off the step stream while the next steps run (the files
        are the same); binary_frames writes legacy VTK BINARY instead of the ASCII text."""
        dt = self.dt
        total = int(round(T_total / dt))
        out_interval = max(1, total // Frames)
        frame = 0
        s = self.info().step
        while s <= total:
            f = ((s + out_interval - 1) // out_interval) * out_interval   # next frame step
            if vtk_dir is None or f > total:
                self.step(total - s + 1)
                break
            self.step(f - s + 1, capture_field=True)
            s = f + 1
            os.makedirs(vtk_dir, exist_ok=True)
            path = os.path.join(vtk_dir, f"flight_{frame:05d}.vtk")
            if async_frames:
                self.dump_field_async(path, binary=binary_frames)
            else:
                self.write_vtk(path, binary=binary_frames)
            frame += 1
        if async_frames:
            self.dump_wait()
        if probe_dir is not None and self.probe_names:
            self.write_probe_files(probe_dir)
        return total

    # -- results ----------------------------------------------------------------------------
    def read_probes(self) -> np.ndarray:
        info = self.info()
        n = info.n_probes
        out = np.zeros((n, max(info.step, 1)), np.float64)
        got = C.c_int64()
        _check(self.lib.nlos_read_probes(self.h, _ptr(out), max(info.step, 1), C.byref(got)))
        return out[:, :got.value]

    def read_field(self) -> np.ndarray:
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx), np.float32)
        _check(self.lib.nlos_read_field(self.h, _ptr(out)))
        return out

    def write_probe_files(self, directory: str) -> None:
        names = (C.c_char_p * len(self.probe_names))(*[n.encode() for n in self.probe_names])
        _check(self.lib.nlos_write_probe_files(self.h, directory.encode(), names))

    def write_vtk(self, path: str, binary: bool = False) -> None:
        fn = self.lib.nlos_write_vtk_binary if binary else self.lib.nlos_write_vtk
        _check(fn(self.h, path.encode()))

    def dump_field_async(self, path: str, binary: bool = False) -> None:
        """Write the captured field to path off the step stream (see nlos_dump_field_async)."""
        _check(self.lib.nlos_dump_field_async(self.h, path.encode(), VTK_BINARY if binary else VTK_ASCII))

    def dump_wait(self) -> None:
        _check(self.lib.nlos_dump_wait(self.h))

    def write_probe_dat(self, directory: str, scale: float = 100.0) -> None:
        """probe_<name>.dat in the family-A format (01_acoustic_scattering.cpp:481-495)."""
        names = (C.c_char_p * len(self.probe_names))(*[n.encode() for n in self.probe_names])
        _check(self.lib.nlos_write_probe_dat(self.h, directory.encode(), names, scale))

    def checkpoint(self) -> bytes:
        need = C.c_size_t()
        _check(self.lib.nlos_checkpoint(self.h, None, 0, C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(self.lib.nlos_checkpoint(self.h, C.cast(buf, C.c_void_p), need.value, C.byref(need)))
        return buf.raw

    def restore(self, blob: bytes) -> None:
        buf = C.create_string_buffer(blob, len(blob))
        _check(self.lib.nlos_restore(self.h, C.cast(buf, C.c_void_p), len(blob)))

    def timing(self, on: Optional[bool] = None) -> Timing:
        if on is not None:
            _check(self.lib.nlos_timing_enable(self.h, 1 if on else 0))
        t = Timing()
        _check(self.lib.nlos_timing_get(self.h, C.byref(t)))
        return t

    def timing_reset(self) -> None:
        _check(self.lib.nlos_timing_reset(self.h))

    # -- gradients -------------------------------------------------------------------
    def record_forward(self, n: int, every: int) -> None:
        """n steps like step(n), keeping a checkpoint every `every` steps for backward()."""
        _check(self.lib.nlos_record_forward(self.h, int(n), int(every)))

    def backward(self, dl_dprobe: np.ndarray, n_samples: int = 0) -> dict:
        """Gradient of a loss over the recorded steps from dL/d(rho-1) per probe and step
        (array [probe][step] over all steps so far): d_position (whole-path shift, per
        metre), d_amp, and d_samples for a sampled waveform."""
        dl = np.ascontiguousarray(dl_dprobe, np.float64)
        g = Grad()
        samples = np.zeros(max(n_samples, 1), np.float64)
        if n_samples:
            g.d_samples = samples.ctypes.data_as(C.POINTER(C.c_double))
            g.n_samples = n_samples
        steps = dl.shape[1] if dl.ndim == 2 else 1
        _check(self.lib.nlos_backward(self.h, _ptr(dl), steps, C.byref(g)))
        return {"d_position": np.array(list(g.d_position)), "d_amp": g.d_amp,
                "d_samples": samples[:n_samples].copy()}

    def debug_read_f(self) -> np.ndarray:
        nx, ny, nz = self.dims
        out = np.empty((nz, ny, nx, 19), np.float64)
        _check(self.lib.nlos_debug_read_f(self.h, _ptr(out)))
        return out

    def debug_poke(self, i: int, j: int, k: int, q: int, f: float) -> None:
        _check(self.lib.nlos_debug_poke(self.h, i, j, k, q, f))


def copy_kernel_ms(nbytes: int, reps: int = 5) -> float:
    """Mean time of a plain device copy of nbytes (read nbytes, write nbytes)."""
    ms = C.c_double()
    _check(load_library().nlos_debug_copy_kernel(nbytes, reps, C.byref(ms)))
    return ms.value


def pull_pattern_ms(nx: int, ny: int, nz: int, reps: int = 5) -> float:
    """Mean time of the FP16S two-array sweep's loads and stores alone on an nx*ny*nz grid
    (19 blocked FP16 slots pulled and stored per interior cell; no arithmetic)."""
    ms = C.c_double()
    _check(load_library().nlos_debug_pull_pattern(nx, ny, nz, reps, C.byref(ms)))
    return ms.value


def grid_dims(cfg: Config):
    """(NXg, NYg, NZg) = (int)((max-min)/dx + 0.5) (05:332-334)."""
    return tuple(int((hi - lo) / cfg.dx + 0.5) for lo, hi in
                 ((cfg.x_min, cfg.x_max), (cfg.y_min, cfg.y_max), (cfg.z_min, cfg.z_max)))


def dt_of(cfg: Config) -> float:
    return (0.57735026918962576451 * cfg.dx) / cfg.c_phys


def reference05_config(precision="f64") -> Config:
    """The hard-coded configuration of 05_sensor_illusion.cpp:1128-1166."""
    return default_config(dx=1.5, x_min=668.0, x_max=1064.0, y_min=17.0, y_max=314.0,
                          z_min=0.0, z_max=106.0, sponge_width=20.0 * 1.5,
                          world_offset=(22248.0, 4123.0, 14.0), precision=precision)


def reference05_probes():
    """Probes B1, B2, C1, C2 of 05:1143-1154."""
    sx, sy, sz = -22248.0, -4123.0, -14.0
    B = (23106.0 + sx, 4343.0 + sy, 100.0 + sz)
    Cc = (23026.0 + sx, 4313.0 + sy, 100.0 + sz)
    return (["B1", "B2", "C1", "C2"],
            [B, (B[0] + 1.5, B[1], B[2]), Cc, (Cc[0] + 1.5, Cc[1], Cc[2])])


def reference05_source() -> Source:
    """Static Gaussian pulse of 05:1156-1164."""
    return default_source(sx=912.0, sy=66.6, sz=86.0, ex=912.0, ey=66.6, ez=86.0,
                          speed=0.0, amp=10.0, freq=0.0)


def reference08_config(precision="f64") -> Config:
    """The hard-coded configuration of 08_city_scale_uav.cpp: the member defaults of the box
    (08:250-253), DX = 1.5 and the STL shift of its main (08:1066-1071)."""
    return default_config(dx=1.5, x_min=0.0, x_max=1650.0, y_min=0.0, y_max=1150.0,
                          z_min=0.0, z_max=250.0, sponge_width=20.0 * 1.5,
                          world_offset=(22248.0, 4123.0, 14.0), precision=precision)


def reference08_source() -> Source:
    """The moving harmonic source of 08:1073-1082 (sine waveform, 08:543)."""
    return default_source(sx=912.0, sy=66.6, sz=86.0, ex=912.0, ey=1127.3, ez=86.0,
                          speed=30.0, amp=0.5, freq=20.0, waveform=WAVE_SINE)


class Batch:
    """Independent scenes advanced together (multi-scene launches, replayed HIP graph).
    scene(i) returns an Engine view of scene i for set-up and read-back; stepping goes
    through Batch.step."""

    def __init__(self, cfgs: Sequence[Config]):
        self.lib = load_library()
        arr = (Config * len(cfgs))(*cfgs)
        h = C.c_void_p()
        _check(self.lib.nlos_batch_create(arr, len(cfgs), C.byref(h)))
        self.h = h
        self.scenes = []
        for i, cfg in enumerate(cfgs):
            sh = C.c_void_p()
            _check(self.lib.nlos_batch_scene(self.h, i, C.byref(sh)))
            self.scenes.append(Engine(cfg, _handle=sh))

    def __len__(self) -> int:
        return len(self.scenes)

    def scene(self, i: int) -> "Engine":
        return self.scenes[i]

    def autotune(self, reps: int = 3) -> Tune:
        t = Tune()
        _check(self.lib.nlos_batch_autotune(self.h, reps, C.byref(t)))
        return t

    def tile_widths(self) -> List[int]:
        """Sweep tile width of every scene (tiles of w x 256/w cells)."""
        w = (C.c_int32 * len(self.scenes))()
        _check(self.lib.nlos_batch_tile_widths(self.h, w, len(self.scenes)))
        return list(w)

    def set_tile_widths(self, widths: Sequence[int]) -> None:
        w = (C.c_int32 * len(widths))(*widths)
        _check(self.lib.nlos_batch_set_tile_widths(self.h, w, len(widths)))

    def step(self, n: int) -> None:
        _check(self.lib.nlos_batch_step(self.h, int(n)))

    def synchronize(self) -> None:
        _check(self.lib.nlos_batch_synchronize(self.h))

    def timing(self, on: Optional[bool] = None) -> Timing:
        if on is not None:
            _check(self.lib.nlos_batch_timing_enable(self.h, 1 if on else 0))
        t = Timing()
        _check(self.lib.nlos_batch_timing_get(self.h, C.byref(t)))
        return t

    def close(self) -> None:
        if getattr(self, "h", None):
            for e in self.scenes:
                e.h = None
            self.lib.nlos_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
