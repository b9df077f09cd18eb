"""Host check of the row blocks of the fused assembly (csrc/sdx_rowblocks.h), no GPU.

tests/cxx/row_blocks_main.cpp is built by g++ with AddressSanitizer and
UndefinedBehaviorSanitizer and run on several meshes side by side, as a world
holds them: hex grids of 1, 4, 9 and 13 hexes per edge (6 tets per hex) with
their boundary triangles, a single tet, a tet soup of random vertex
quadruples at random points (vertices in many tets, repeated coordinates),
under the FP32 and FP64 slot caps (1024, 512) and a small cap that forces
blocks to split by slots rather than rows.  It checks that every vertex lies
in one block, the block limits, the tets and triangles of each block with
their local rows, and that contributions summed block row by block row (as
the fused kernel sums them) equal a direct accumulation exactly.  On the
grids, a block's tets exceed its share by less than 2.5x (compact blocks).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from project_superdex_amd import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, "tests", "cxx", "row_blocks_main.cpp")
CSRC = os.path.join(ROOT, "project_superdex_amd", "csrc")


def boundary(tets):
    faces = {}
    for t in np.asarray(tets):
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            key = tuple(sorted(int(t[i]) for i in f))
            faces[key] = faces.get(key, 0) + 1
    return np.array([k for k, c in faces.items() if c == 1], dtype=np.int64).reshape(-1, 3)


def meshes(grid_only=False):
    out = []
    for n in (1, 4, 9, 13):
        X, tets = mesh.hex_grid_tets(n, 0.01)
        out.append((X, tets, boundary(tets)))
    if grid_only:
        return out
    out.append((np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]]), np.array([[0, 1, 2, 3]]),
                np.array([[0, 1, 2]])))
    rng = np.random.default_rng(5)
    X = rng.integers(0, 4, size=(60, 3)).astype(np.float64)  # many repeated coordinates
    soup = np.array([rng.choice(60, 4, replace=False) for _ in range(300)])
    out.append((X, soup, boundary(soup)))
    return out


def run(tmp_path, ms, cap):
    cxx = shutil.which("g++") or "g++"
    exe = str(tmp_path / "row_blocks")
    if not os.path.exists(exe):
        cmd = [cxx, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror", "-fsanitize=address,undefined",
               "-fno-sanitize-recover=all", "-I", CSRC, MAIN, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    path = tmp_path / f"meshes_{cap}.txt"
    with open(path, "w") as f:
        for X, tets, tris in ms:
            f.write(f"{X.shape[0]} {len(tets)} {len(tris)}\n")
            f.write(" ".join(repr(float(x)) for x in np.asarray(X).ravel()) + "\n")
            f.write(" ".join(str(int(x)) for x in np.asarray(tets).ravel()) + "\n")
            f.write(" ".join(str(int(x)) for x in np.asarray(tris).ravel()) + "\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(path), str(cap)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
    parts = r.stdout.split()
    return int(parts[2]), float(parts[4]), int(parts[6])


@pytest.mark.parametrize("cap", [1024, 512, 160])
def test_row_blocks_sum_like_a_direct_assembly(tmp_path, cap):
    blocks, redundancy, max_slots = run(tmp_path, meshes(), cap)
    assert blocks > 0 and max_slots <= cap


def test_grid_blocks_are_compact(tmp_path):
    blocks, redundancy, max_slots = run(tmp_path, meshes(grid_only=True), 1024)
    assert redundancy < 2.5, redundancy
    print(f"blocks {blocks}, tet redundancy {redundancy:.2f}, max slots {max_slots}")
