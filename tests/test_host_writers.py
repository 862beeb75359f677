"""The host layer's dump writers (lbm_ib.c:234-692 formats) on a hand-made
grid and sheet, with no device (tests/host_writers_check.c): FieldView grid
and function files, the velocity text dump, the fiber FieldView files and
the stdout printers."""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "distributed-lb-ib_amd")

pytestmark = [pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available"),
              pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "liblbmib_host.so")),
                                 reason="liblbmib_host.so not built")]


def _values(text):
    return [float(v) for v in text.strip().rstrip(",").split(",")]


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    d = tmp_path_factory.mktemp("writers")
    exe = str(d / "host_writers_check")
    subprocess.check_call(["gcc", "-O1", "-I", os.path.join(REPO, "include"), "-o", exe,
                           os.path.join(REPO, "tests", "host_writers_check.c"), "-L", PKG, "-llbmib_host",
                           "-llbmib", "-Wl,-rpath," + PKG, "-lm"])
    r = subprocess.run([exe], cwd=str(d), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return d, r.stdout


def _u(X, Y, Z):
    return X + 10.0 * Y + 100.0 * Z


def test_fieldview_fluid_files(run):
    d, _ = run
    head, body = open(d / "fv_vel.f").read().split("\n", 1)
    assert head == "5,4,3,3"
    Z, Y, X = np.meshgrid(np.arange(3), np.arange(4), np.arange(5), indexing="ij")  # X fastest
    u = _u(X, Y, Z).ravel()
    np.testing.assert_allclose(_values(body), np.concatenate([u, -u / 4, u / 8]), atol=5e-7)
    head, body = open(d / "fv_grid.xyz").read().split("\n", 1)
    assert head == "5,4,3"
    assert _values(body) == list(np.concatenate([X.ravel(), Y.ravel(), Z.ravel()]).astype(float))


def test_velocity_text_dump(run):
    d, _ = run
    rows = [list(map(float, line.split(","))) for line in open(d / "vel.txt").read().split()]
    want = [(_u(x, y, z), -_u(x, y, z) / 4, _u(x, y, z) / 8) for x in range(1, 4) for y in range(1, 3)
            for z in range(1, 3)]  # [start, end), X slowest, Z fastest
    np.testing.assert_allclose(rows, want, atol=5e-7)


def test_fieldview_fiber_files(run):
    d, _ = run
    head, body = open(d / "fv_fiber.xyz").read().split("\n", 1)
    assert head == "3,4,1"
    J, I = np.meshgrid(np.arange(4), np.arange(3), indexing="ij")  # row index fastest
    np.testing.assert_allclose(_values(body), np.concatenate([(I + 0.25).ravel(), (J + 0.5).ravel(),
                                                              (I * J).ravel()]), atol=5e-7)
    head, body = open(d / "fv_fiber_ef.f").read().split("\n", 1)
    assert head == "3,4,1,3"
    np.testing.assert_allclose(_values(body), np.concatenate([I.ravel(), -J.ravel(), np.full(12, 0.5)]),
                               atol=5e-7)


def test_stdout_printers(run):
    _, out = run
    lines = out.splitlines()
    assert lines[0].startswith("(i,j): {cord_x, cord_y, cord_z}")
    assert lines[1].startswith("(0, 1):{0.250000, 1.500000, 0.000000} || 1.000000000000")
    cube = lines.index("(BI,BJ,BK): {vel_x, vel_y, vel_z} || {G0, DF1, DF2}|| rho || {ElasticF_x, y, z}")
    assert lines[cube + 1] == "For cube <0, 0, 0>"
    # cube (0,0,0) of side 2 then cube (0,0,0)..(0,0,0): 8 nodes x (1 + 19) lines
    assert sum(1 for line in lines if line.startswith("For cube")) == 8
    assert lines[cube + 3].startswith("(0, 0, 0, 1):{0.000000000000,")


def build_pthread_caller(d):
    """Compiles tests/host_pthread_caller.c -- a caller written against the
    reference's GV/LV API with its own main() that uses threads, lock_Fluid
    and barr as lbm_ib.c:3836-3955 does -- against include/lbmib_host.h."""
    exe = str(d / "host_pthread_caller")
    subprocess.check_call(["gcc", "-O1", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), "-o", exe,
                           os.path.join(REPO, "tests", "host_pthread_caller.c"), "-L", PKG, "-llbmib_host", "-llbmib",
                           "-Wl,-rpath," + PKG, "-lpthread", "-lm"])
    return exe


def test_reference_style_threaded_caller_compiles(tmp_path):
    exe = build_pthread_caller(tmp_path)
    assert os.path.exists(exe)
    # the structs keep the reference's field order for the fields it has
    src = open(os.path.join(REPO, "include", "lbmib_host.h")).read()
    gv = src[src.index("typedef struct gv_t"):]
    order = [gv.index(f) for f in ("int ib, ie", "pthread_t *threads", "pthread_mutex_t *lock_Fluid",
                                   "int total_threads", "int cube_size", "int P, Q, R", "pthread_barrier_t barr")]
    assert order == sorted(order)
    fg = src[src.index("typedef struct fluid_grid_t"):]
    order = [fg.index(f) for f in ("int x_dim", "Fluidsurface *inlet", "Fluidsurface *outlet",
                                   "Sub_Fluidgrid *sub_fluid_grid")]
    assert order == sorted(order)
