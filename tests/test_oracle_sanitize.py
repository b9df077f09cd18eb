"""The CPU oracle (oracle/sdx_oracle.cpp) built with AddressSanitizer and
UndefinedBehaviorSanitizer (g++, host code only) and driven through every entry
by tests/cxx/oracle_sanitize_main.cpp: two scenes (FP32, FP64) of two 3^3-hex
cubes with pins, per-element materials, sphere and plane contact, friction,
mass and stiffness damping, stepped with every line search, PSD mode (with
unconverged linear solves, so retries and the gradient-descent fallback run)
and forcing term; boundary conditions set and cleared; contact queries,
assembly and objective; the synthetic Newton problems; the CSR PCG; the
element and material routines.  Any sanitizer report fails the run.  No GPU.
"""
import os
import shutil
import subprocess

import numpy as np

from project_superdex_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "tests", "cxx", "oracle_sanitize_main.cpp")
ORACLE = os.path.join(ROOT, "oracle", "sdx_oracle.cpp")


def test_oracle_runs_clean_under_sanitizers(tmp_path):
    cxx = shutil.which("g++") or "g++"
    exe = str(tmp_path / "oracle_sanitize")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include"), MAIN, ORACLE, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    X, tets = mesh.hex_grid_tets(3, 0.04 / 3)
    pins = mesh.nodes_on_plane(X, 1, 0.0)
    path = tmp_path / "cube.txt"
    with open(path, "w") as f:
        f.write(f"{X.shape[0]} {tets.shape[0]}\n")
        f.write(" ".join(repr(float(x)) for x in X.ravel()) + "\n")
        f.write(" ".join(str(int(t)) for t in tets.ravel()) + "\n")
        f.write(f"{len(pins)}\n" + " ".join(str(int(p)) for p in pins) + "\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout[-3000:] + r.stderr[-5000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-5000:]
