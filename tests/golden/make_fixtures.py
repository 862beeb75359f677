"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Runs only in a container that has the reference tree: ``make -C oracle ref``
compiles cube_pthread/lbm_ib.c (with -DDUMP_VERIFY) into oracle/_ref/, and
this script runs that binary on small configurations and keeps its outputs:

  <name>_fiber_step<N>.bin   the reference's binary fiber dump, unchanged:
                             {int32 i, int32 j, 12 x f64 Fibernode} records
                             (lbm_ib.c:604-646)
  <name>_fluid_step<N>.npz   the reference's %.24f fluid-cube dump
                             (lbm_ib.c:332-392) parsed back into arrays, kept
                             only over the box the reference names
                             (x 19..22, y 20..25, z 10..33, lbm_ib.c:3511)
  fixtures.json              the command-line flags of every run

Usage:  python tests/golden/make_fixtures.py
"""
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref", "scalable_IB_v4.1_dumpverify")

# (name, grid, sheet rows/cols, width/height, x0/y0/z0, steps, dump, fiber steps, fluid steps)
RUNS = [
    ("refA", 64, 26, 25.0, (16.0, 19.5, 19.5), 20, 10, (1, 10, 20), (20,)),
    ("refB", 48, 20, 19.0, (12.0, 14.5, 14.5), 120, 60, (1, 60, 120), (120,)),
]
BOX = ((19, 22), (20, 25), (10, 33))

LINE = re.compile(r"\((\d+),(\d+),(\d+),\s*(\d+)\):\{([^}]*)\} \|\| \{([^}]*)\} \|\| (\S+) \|\| \{([^}]*)\}")


def parse_fluid(path):
    recs = {}
    with open(path) as fh:
        for line in fh:
            m = LINE.match(line)
            if not m:
                continue
            X, Y, Z, q = map(int, m.group(1, 2, 3, 4))
            if not (BOX[0][0] <= X <= BOX[0][1] and BOX[1][0] <= Y <= BOX[1][1]
                    and BOX[2][0] <= Z <= BOX[2][1]):
                continue
            vel = [float(v) for v in m.group(5).split(",")]
            dfs = [float(v) for v in m.group(6).split(",")]
            rho = float(m.group(7))
            frc = [float(v) for v in m.group(8).split(",")]
            recs.setdefault((X, Y, Z), [vel, rho, frc, [None] * 19])[3][q] = dfs
    keys = sorted(recs)
    xyz = np.array(keys, dtype=np.int16)
    vel = np.array([recs[k][0] for k in keys])
    rho = np.array([recs[k][1] for k in keys])
    frc = np.array([recs[k][2] for k in keys])
    dfs = np.array([recs[k][3] for k in keys])  # [M, 19, 3] = dfeq, df1, df2
    return dict(xyz=xyz, vel=vel, rho=rho, force=frc,
                dfeq=dfs[:, :, 0], df1=dfs[:, :, 1], df2=dfs[:, :, 2])


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])
    meta = {}
    for name, n, rc, wh, x0, steps, dump, fsteps, lsteps in RUNS:
        args = ["-steps", str(steps), "-dump", str(dump), "-fluid_grid_xyz", str(n), str(n), str(n),
                "-cube_size", "4", "-thread_per_task_xyz", "2", "1", "1", "-num_fibersht", "1",
                "-fibersht_width_height", repr(wh), repr(wh), "-fibersht_row_clmn", str(rc), str(rc),
                "-fibersht_xyz_0", repr(x0[0]), repr(x0[1]), repr(x0[2])]
        tmp = tempfile.mkdtemp(prefix="lbmib_ref_")
        try:
            subprocess.check_call([REF_BIN] + args, cwd=tmp, stdout=subprocess.DEVNULL)
            for s in fsteps:
                shutil.copy(os.path.join(tmp, f"Fiber1_dump_step{s}.bin"),
                            os.path.join(HERE, f"{name}_fiber_step{s}.bin"))
            for s in lsteps:
                arrs = parse_fluid(os.path.join(tmp, f"Fluid0_dump_step{s}.dat"))
                np.savez_compressed(os.path.join(HERE, f"{name}_fluid_step{s}.npz"), **arrs)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        meta[name] = dict(args=args, grid=n, rows=rc, cols=rc, width=wh, height=wh, origin=x0,
                          steps=steps, fiber_steps=list(fsteps), fluid_steps=list(lsteps),
                          fluid_box=BOX, threads="2x1x1")
    with open(os.path.join(HERE, "fixtures.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    print("fixtures written:", sorted(meta))


if __name__ == "__main__":
    sys.exit(main())
