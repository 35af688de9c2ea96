"""Writes the 08_city_scale_uav golden fixture by running the compiled, unmodified program.

Runs `oracle/_ref/08_city_scale_uav` (built by `make -C oracle ref` from
/root/reference/08_city_scale_uav.cpp) in a scratch directory holding the synthetic city of
`city_stl.reference08_city_stl()` as `csx_model_binary.stl` (the name 08:1069 hard-codes).
Everything else is 08's own configuration (08:1062-1092 and the member defaults
08:243-257): DX=1.5, 1100x767x167 cells, tau=0.55, a 20 Hz sine source of amplitude 0.5
moving from (912, 66.6, 86) along +y at 30 m/s, `run(3.0, 3)`: 1178+1 steps (dt = cs*DX/340)
with ASCII VTK frames (08:937-1022) after steps s = 0, 392, 784 and 1176 (08:1026-1058).

The program is stopped once frame 1 (after step s = 392, i.e. 393 steps) is on disk: its
progress line of s = 400 (08:1052-1056) is printed after that frame's file is closed.
Kept as fixtures (ref08_city/vtk_frames.json): the header lines of frames 0 and 1, every
8th cell in x and y of the z layers 5, 57 (the source height) and 100 of frame 1, and of
layer 57 of frame 0.  The frames (1.8 GB each) are deleted.  Needs the reference sources,
the image's MPICH and about 48 GB of host memory (08 keeps two FP64 AoS population arrays);
it is run in the build container only.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import city_stl  # noqa: E402

NXG, NYG, NZG = 1100, 767, 167
STRIDE = 8
LAYERS = {"0": (57,), "1": (5, 57, 100)}


def sample_vtk(path: str, layers):
    want = {}
    for k in layers:
        for j in range(0, NYG, STRIDE):
            for i in range(0, NXG, STRIDE):
                want[(k * NYG + j) * NXG + i] = None
    values = {}
    with open(path) as fh:
        header = [next(fh).rstrip("\n") for _ in range(10)]
        for n, line in enumerate(fh):
            if n in want:
                values[n] = line.strip()
    assert len(values) == len(want), (len(values), len(want))
    out = {"header": header, "stride": STRIDE, "slices": {}}
    for k in layers:
        out["slices"][str(k)] = [values[(k * NYG + j) * NXG + i]
                                 for j in range(0, NYG, STRIDE) for i in range(0, NXG, STRIDE)]
    return out


def main() -> None:
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    exe = os.path.join(REPO, "oracle", "_ref", "08_city_scale_uav")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    dest = os.path.join(HERE, "ref08_city")
    os.makedirs(dest, exist_ok=True)
    work = tempfile.mkdtemp(prefix="ref08_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        with open(os.path.join(work, "csx_model_binary.stl"), "wb") as fh:
            fh.write(city_stl.reference08_city_stl())
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        log_path = os.path.join(work, "stdout.txt")
        with open(log_path, "w") as log:
            proc = subprocess.Popen([exe], cwd=work, env=env, stdout=log, stderr=subprocess.STDOUT)
            try:
                while proc.poll() is None:
                    time.sleep(10)
                    text = open(log_path).read()
                    print(text.replace("\r", "\n").strip().split("\n")[-1], flush=True)
                    if "Step 400/" in text:
                        break
            finally:
                if proc.poll() is None:
                    proc.kill()
                    proc.wait()
        text = open(log_path).read()
        assert "Step 400/" in text, text[-2000:]
        run = [ln for ln in text.split("\n") if ln.startswith("Run: ")][0]
        interval = int(run.split("out_interval=")[1].split()[0])
        assert interval < 400, run
        # the fixture keeps the program's lines up to the stop, progress lines once each
        lines = [ln for ln in text.replace("\r", "\n").split("\n") if ln.strip()]
        with open(os.path.join(dest, "stdout.txt"), "w") as fh:
            fh.write("\n".join(lines) + "\n")
        frames = {f: sample_vtk(os.path.join(work, "output_mpi", f"flight_{int(f):05d}.vtk"), layers)
                  for f, layers in LAYERS.items()}
        with open(os.path.join(dest, "vtk_frames.json"), "w") as fh:
            json.dump({"grid": [NXG, NYG, NZG], "steps": {"0": 1, "1": interval + 1}, "frames": frames}, fh,
                      indent=0)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    print("done ref08_city", flush=True)


if __name__ == "__main__":
    main()
