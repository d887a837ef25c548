"""The oracle's C restatement built with AddressSanitizer and
UndefinedBehaviorSanitizer and driven through its paths
(oracle/sanitize_check.c): the reference step on the reference lattice and
the grid step with terrain and erosion in every precision mode, lane count,
slab count and y/z split.  Any invalid access, leak, overflow or other
undefined behaviour the sanitizers see fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_oracle_paths_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "orc_sanitized"
    cmd = (["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-fopenmp", "-Wall", "-Wno-unknown-pragmas"]
           + SANITIZE + ["-I" + ORACLE, os.path.join(ORACLE, "sph_oracle.c"),
                         os.path.join(ORACLE, "sanitize_check.c"), "-lm", "-o", str(exe)])
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("this gcc has no sanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    # a library preloaded by the environment may come before the ASan runtime
    env = dict(os.environ, OMP_NUM_THREADS="2", UBSAN_OPTIONS="print_stacktrace=1",
               ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0")
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240, env=env)
    assert run.returncode == 0, (run.stdout[-1000:], run.stderr[-3000:])
    assert "oracle paths clean" in run.stdout
