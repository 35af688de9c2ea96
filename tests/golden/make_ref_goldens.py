"""Writes golden fixtures by running the compiled, unmodified reference programs.

Runs `oracle/_ref/05_sensor_illusion` (built by `make -C oracle ref` from
/root/reference/05_sensor_illusion.cpp) twice in a scratch directory:

  flat : no `csx_model_binary.stl` present, so the reference falls back to empty geometry
         (05:397-410) and only the ground plane (05:520-525) is solid;
  city : the synthetic city of `city_stl.reference_city_stl()` dropped in as
         `csx_model_binary.stl` (the file name 05:1131 hard-codes).

Everything else is the reference's own hard-coded configuration (05:1128-1170): DX=1.5,
264x198x71 cells, tau=0.55, 393+1 steps, Gaussian pulse of amplitude 10 at (912,66.6,86),
probes B1/B2/C1/C2.  Each run writes `probe_<name>.txt` (05:1106-1118) and 99 ASCII VTK
frames (05:975-1062).  Kept as fixtures:

  ref05_<scene>/probe_<name>.txt   the probe files, byte for byte;
  ref05_<scene>/vtk_frames.json    header lines and a strided z-slice of two frames.

The VTK frames (about 5 GB per run) are deleted after sampling.  This script needs the
reference sources and the image's MPICH; it is run in the build container only.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import city_stl  # noqa: E402

NXG, NYG, NZG = 264, 198, 71
SAMPLE_FRAMES = (50, 98)
SLICES = (57, 5)      # k of the probe/source height and of a near-ground layer
STRIDE = 4


def sample_vtk(path: str):
    with open(path) as fh:
        header = [next(fh).rstrip("\n") for _ in range(10)]
        want = {}
        for k in SLICES:
            for j in range(0, NYG, STRIDE):
                for i in range(0, NXG, STRIDE):
                    want[(k * NYG + j) * NXG + i] = (k, j, i)
        values = {}
        for n, line in enumerate(fh):
            if n in want:
                values[n] = line.strip()
    out = {"header": header, "stride": STRIDE, "slices": {}}
    for k in SLICES:
        vals = []
        for j in range(0, NYG, STRIDE):
            for i in range(0, NXG, STRIDE):
                vals.append(values[(k * NYG + j) * NXG + i])
        out["slices"][str(k)] = vals
    return out


def run_scene(name: str, stl: bytes | None, exe: str, threads: int) -> None:
    dest = os.path.join(HERE, f"ref05_{name}")
    os.makedirs(dest, exist_ok=True)
    work = tempfile.mkdtemp(prefix=f"ref05_{name}_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        if stl is not None:
            with open(os.path.join(work, "csx_model_binary.stl"), "wb") as fh:
                fh.write(stl)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        with open(os.path.join(dest, "stdout.txt"), "w") as log:
            subprocess.run([exe], cwd=work, env=env, check=True, stdout=log,
                           stderr=subprocess.STDOUT)
        for p in ("B1", "B2", "C1", "C2"):
            shutil.copy(os.path.join(work, f"probe_{p}.txt"), os.path.join(dest, f"probe_{p}.txt"))
        frames = {}
        for f in SAMPLE_FRAMES:
            frames[str(f)] = sample_vtk(os.path.join(work, "output_mpi", f"flight_{f:05d}.vtk"))
        with open(os.path.join(dest, "vtk_frames.json"), "w") as fh:
            json.dump({"grid": [NXG, NYG, NZG], "frames": frames}, fh, indent=0)
    finally:
        shutil.rmtree(work, ignore_errors=True)


# 01_acoustic_scattering (family A): 320x320x8 cells, 402 probe records, frames at steps
# 0, 104 and 208 (01:502-524).  Kept: probe_<A|B|C>.dat byte for byte, and of each frame the
# header and every 8th cell in x and y of the z layer 4.
REF01_GRID = (320, 320, 8)
REF01_FRAMES = (0, 104, 208)
REF01_LAYER, REF01_STRIDE = 4, 8


def sample_vtk01(path: str):
    nx, ny, _ = REF01_GRID
    with open(path) as fh:
        header = [next(fh).rstrip("\n") for _ in range(10)]
        want = {(REF01_LAYER * ny + j) * nx + i for j in range(0, ny, REF01_STRIDE)
                for i in range(0, nx, REF01_STRIDE)}
        values = {}
        for n, line in enumerate(fh):
            if n in want:
                values[n] = line.strip()
    vals = [values[(REF01_LAYER * ny + j) * nx + i] for j in range(0, ny, REF01_STRIDE)
            for i in range(0, nx, REF01_STRIDE)]
    return {"header": header, "layer": REF01_LAYER, "stride": REF01_STRIDE, "values": vals}


def run_ref01(exe: str, threads: int) -> None:
    dest = os.path.join(HERE, "ref01")
    os.makedirs(dest, exist_ok=True)
    work = tempfile.mkdtemp(prefix="ref01_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        with open(os.path.join(dest, "stdout.txt"), "w") as log:
            subprocess.run([exe], cwd=work, env=env, check=True, stdout=log, stderr=subprocess.STDOUT)
        for p in ("A", "B", "C"):
            shutil.copy(os.path.join(work, "output_mpi", f"probe_{p}.dat"), os.path.join(dest, f"probe_{p}.dat"))
        frames = {str(f): sample_vtk01(os.path.join(work, "output_mpi", f"cyl_{f:05d}.vtk")) for f in REF01_FRAMES}
        with open(os.path.join(dest, "vtk_frames.json"), "w") as fh:
            json.dump({"grid": list(REF01_GRID), "frames": frames}, fh, indent=0)
    finally:
        shutil.rmtree(work, ignore_errors=True)


def main() -> None:
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scenes = sys.argv[1:] or ["flat", "city", "ref01"]
    for s in scenes:
        prog = "01_acoustic_scattering" if s == "ref01" else "05_sensor_illusion"
        exe = os.path.join(REPO, "oracle", "_ref", prog)
        if not os.path.exists(exe):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
        if s == "ref01":
            run_ref01(exe, threads)
        else:
            run_scene(s, None if s == "flat" else city_stl.reference_city_stl(), exe, threads)
        print("done", s, flush=True)


if __name__ == "__main__":
    main()
