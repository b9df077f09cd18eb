"""Host check of the element patches and record maps (csrc/sdx_maps.h), no GPU.

tests/cxx/patch_maps_main.cpp is built by g++ with AddressSanitizer and
UndefinedBehaviorSanitizer and run on several meshes side by side, as a world
holds them: hex grids of 2, 5 and 9 hexes per edge (6 tets per hex, patch
tails of every length), a single tet, and a tet soup of random vertex
quadruples (shared pairs in both orientations, vertices in many tets of a
patch).  Integer-valued contributions are summed through the patch
descriptions the way the element kernel reads them and then over the record
runs the way the gathers do; the result must equal a direct accumulation over
the tets exactly, every contribution must be listed once, every record written.
"""
import os
import shutil
import subprocess

import numpy as np

from project_superdex_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "tests", "cxx", "patch_maps_main.cpp")
CSRC = os.path.join(ROOT, "project_superdex_amd", "csrc")


def meshes():
    out = []
    for n in (2, 5, 9):
        X, tets = mesh.hex_grid_tets(n, 0.01)
        out.append((X.shape[0], tets))
    out.append((4, np.array([[0, 1, 2, 3]])))
    rng = np.random.default_rng(3)
    soup = np.array([rng.choice(40, 4, replace=False) for _ in range(150)])
    out.append((40, soup))
    return out


def test_patch_descriptions_sum_like_a_direct_assembly(tmp_path):
    cxx = shutil.which("g++") or "g++"
    exe = str(tmp_path / "patch_maps")
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I", CSRC, MAIN, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    path = tmp_path / "meshes.txt"
    with open(path, "w") as f:
        for nv, tets in meshes():
            f.write(f"{nv} {len(tets)}\n")
            f.write(" ".join(str(int(x)) for x in np.asarray(tets).ravel()) + "\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(path)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    # fewer records than one per (tet, vertex) and (tet, pair)
    parts = r.stdout.split()
    nv_rec, ne_rec, nt = int(parts[4]), int(parts[6]), int(parts[8])
    assert nv_rec < 4 * nt and ne_rec < 6 * nt
