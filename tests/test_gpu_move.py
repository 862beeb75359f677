"""The single-instance move (k_move_tiles in csrc/lbmib_ib.hip: the band
velocity and X += dt u(X) in one launch, per tile of 2x2 bins) against the
two-kernel move (k_band_u + k_move, the batched path's kernels, selected
with LBMIB_MOVE_KERNELS=2), bit for bit: populations, fiber positions and
forces after 30 steps, in

- a sheet tilted 65 degrees out of the y-z plane towards x: the nodes of
  one tile (8 cells in z) span about 20 cells in x, wider than one LDS slab
  of an 11x11-cell (y, z) region, so the per-node sums run across slabs;
- two sheets sharing their (y, z) footprint at different x (the two-sheet
  workload's arrangement): one tile holds nodes of both;
- the small inflow case with a jittered sheet, F64, F32 and BF16.

The two-kernel move follows interpolate/move (lbm_ib.c:2614-2699) and is
pinned against the oracle by the batched tests; the tiled one must give the
same bits.  Each mode runs in its own process (the switch is read once per
process).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "distributed-lb-ib_amd")

SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import lbmib
from oracle.oracle import BC_INOUT, BC_PERIODIC, BC_WALL

name, precision, out = sys.argv[2], sys.argv[3], sys.argv[4]
bc = (BC_INOUT, BC_PERIODIC, BC_WALL)
rng = np.random.default_rng(7)
if name == "tilted":
    # rows along y, columns along (sin 65, 0, cos 65) in x-z at unit spacing
    rows, cols = 16, 40
    i, j = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    th = np.deg2rad(65.0)
    xyz = np.stack([20.0 + j * np.sin(th), 8.0 + i + 0.05 * rng.standard_normal(i.shape), 14.0 + j * np.cos(th)], -1)
    sheets = [lbmib.Sheet(rows, cols, cols - 1.0, rows - 1.0, 2e-3, 2e-3, xyz)]
    n = (96, 32, 48)
elif name == "stacked":
    sheets = [lbmib.Sheet.centred(16, 16, 1.0, (xc, 16.0, 16.0), 2e-3, 2e-3, jitter=0.1, seed=k + 1)
              for k, xc in enumerate((24.0, 56.0))]
    n = (96, 32, 32)
else:
    sheets = [lbmib.Sheet.centred(12, 12, 1.0, (18.0, 16.0, 16.0), 2e-3, 2e-3, jitter=0.08, seed=3)]
    n = (64, 32, 32)
kw = dict(nx=n[0], ny=n[1], nz=n[2], bc=bc, precision=precision, tau=0.7, rho0=1.0, u0=(0.02, 0.0, 0.0),
          quirks=True, sheets=sheets)
with lbmib.Engine(**kw) as e:
    e.init()
    e.step(30)
    assert e.status() == 0, e.status()
    fib = np.concatenate([e.fibers(s).reshape(-1) for s in range(len(sheets))])
    np.savez(out, df=e.df(), force=e.force(), fibers=fib)
print("done", name, precision, flush=True)
"""

CASES = [("tilted", "f64"), ("stacked", "f64"), ("small", "f64"), ("small", "f32"), ("small", "bf16")]


def run(tmp_path, name, precision, mode):
    out = tmp_path / f"{name}_{precision}_{mode}.npz"
    env = dict(os.environ)
    env.pop("LBMIB_MOVE_KERNELS", None)
    if mode == "two":
        env["LBMIB_MOVE_KERNELS"] = "2"
    env["PYTHONPATH"] = ROOT + os.pathsep + env.get("PYTHONPATH", "")
    print(f"[move] {name} {precision} {mode}", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT, PKG, name, precision, str(out)], env=env, cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.load(out)


@pytest.mark.parametrize("name,precision", CASES)
def test_tiled_move_matches_two_kernel_move(tmp_path, name, precision):
    a = run(tmp_path, name, precision, "tiles")
    b = run(tmp_path, name, precision, "two")
    for k in ("fibers", "df", "force"):
        assert np.array_equal(a[k], b[k]), k
    # the sheets moved
    assert np.abs(a["fibers"]).max() > 0
