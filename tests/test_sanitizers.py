"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

- The FP64 oracle (oracle/lbmib_oracle.c), driven through whole steps on
  every boundary kind, with two sheets and the fiber-leaves-the-grid error
  path, by oracle/sanitize_check.c.
- The reference-API host layer and CLI (distributed-lb-ib_amd/host/*.c)
  through their flag checks and set-up.  Without a GPU the set-up stops at
  the device context, with the "init failed" error of the CLI.

Both are built with gcc (host code only; the HIP library is linked as built).
A sanitizer finding aborts the program with its report on stderr.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "distributed-lb-ib_amd")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")


def _clean(out):
    return "AddressSanitizer" not in out and "runtime error" not in out and "LeakSanitizer" not in out


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "sanitize_check")
    subprocess.check_call(["gcc", *SAN, "-ffp-contract=off", "-o", exe,
                           os.path.join(REPO, "oracle", "sanitize_check.c"),
                           os.path.join(REPO, "oracle", "lbmib_oracle.c"), "-lm"])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    out = r.stdout + r.stderr
    assert _clean(out), out[-4000:]
    assert r.returncode == 0, out[-2000:]
    assert "0 failing case(s)" in r.stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "liblbmib.so")), reason="liblbmib.so not built")
def test_host_layer_and_cli_under_sanitizers(tmp_path):
    exe = str(tmp_path / "lbmib_cli_san")
    subprocess.check_call(["gcc", *SAN, "-std=c11", "-I", os.path.join(REPO, "include"), "-o", exe,
                           os.path.join(PKG, "host", "lbmib_cli.c"), os.path.join(PKG, "host", "lbmib_host.c"),
                           "-L", PKG, "-llbmib", "-Wl,-rpath," + PKG, "-lpthread", "-lm"])
    full = ["-steps", "2", "-dump", "1000", "-fluid_grid_xyz", "32", "32", "32", "-cube_size", "4",
            "-thread_per_task_xyz", "2", "2", "2", "-num_fibersht", "1", "-fibersht_width_height", "7", "7",
            "-fibersht_row_clmn", "8", "8", "-fibersht_xyz_0", "8", "12", "12"]
    odd = list(full)
    odd[odd.index("-thread_per_task_xyz") + 1] = "3"
    cases = [(["-steps", "0"], {2}), (["-bogus"], {2}), (full[:4], {2}), (odd, {1}), (full, {0, 1})]
    for args, codes in cases:
        # the HIP runtime is not instrumented: leak reports would name its
        # allocations, so only errors are checked
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300, cwd=str(tmp_path),
                           env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"))
        out = r.stdout + r.stderr
        assert _clean(out), (args, out[-4000:])
        assert r.returncode in codes, (args, r.returncode, out[-2000:])
