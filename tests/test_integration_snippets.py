"""The C++ binding code of INTEGRATION.md compiles against include/sdx_fem.h
and links libsdxfem.so (CPU), and does what the document says (GPU).

Every ```cpp block of INTEGRATION.md except those marked `// reference side`
is extracted in order into snippets.inc and built with tests/cxx/integration_main.cpp
by g++ (host code only).  On the GPU:
- L1 (SdxSolveBsr3 -> sdx_pcg_bsr3_solve_f32, row-strip float arrays) on the
  reference's golden 100x100 system (krylov_solver_test_data.h) converges to
  rel. residual 1e-6 (recomputed in FP64 on the float-rounded system: within
  1e-5, x itself being rounded to float) and gives the same solution as the
  Python binding of the same entry;
- L2 (SdxIslandCreate / SdxIslandStep / SdxSnapshotRoundTrip) on a 4^3-hex
  cube under a moving plate: every step converges, the snapshot restores an
  equal state, and the positions equal those of the CPU oracle stepped with the
  same parameters within 1e-4 relative.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from project_superdex_amd import _abi, _lib, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOC = os.path.join(ROOT, "INTEGRATION.md")
MAIN = os.path.join(ROOT, "tests", "cxx", "integration_main.cpp")
GOLD = os.path.join(ROOT, "tests", "golden", "krylov_psd100.npz")


def snippets() -> str:
    blocks = re.findall(r"```cpp\n(.*?)```", open(DOC).read(), flags=re.S)
    kept = [b for b in blocks if not b.lstrip().startswith("// reference side")]
    assert len(kept) >= 5 and len(kept) < len(blocks)
    return "\n".join(kept)


def build(tmp) -> str:
    if not os.path.exists(_lib.LIB_PATH):
        from project_superdex_amd import build as b
        b.build()
    with open(os.path.join(tmp, "snippets.inc"), "w") as f:
        f.write(snippets())
    exe = os.path.join(tmp, "integration_main")
    cxx = shutil.which("g++") or "g++"
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), "-I", tmp,
           MAIN, "-o", exe, "-L", libdir, "-lsdxfem", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_snippets_compile_and_link(tmp_path):
    exe = build(str(tmp_path))
    assert os.path.exists(exe)


def golden_strip():
    k = np.load(GOLD)
    A = np.zeros((100, 100))
    A[k["rows"], k["cols"]] = k["vals"]
    Ap = np.eye(102)
    Ap[:100, :100] = A
    ptr, col, strips = [0], [], []
    for bi in range(34):
        row = []
        for bj in range(34):
            blk = Ap[3 * bi:3 * bi + 3, 3 * bj:3 * bj + 3]
            if np.any(blk != 0) or bi == bj:
                col.append(bj)
                row.append(blk)
        strips.append(np.concatenate(row, axis=1).reshape(-1))  # 3 x (3 nnz) row-major
        ptr.append(len(col))
    b = np.concatenate([k["rhs"], [0.0, 0.0]])
    return np.array(ptr, np.int32), np.array(col, np.int32), np.concatenate(strips), b, Ap


@pytest.mark.gpu
def test_l1_snippet_on_the_golden_system(tmp_path):
    exe = build(str(tmp_path))
    ptr, col, strip, b, Ap = golden_strip()
    sysf, xf = tmp_path / "system.bin", tmp_path / "x.bin"
    with open(sysf, "wb") as f:
        f.write(np.array([len(ptr) - 1, len(col)], np.int32).tobytes())
        f.write(ptr.tobytes())
        f.write(col.tobytes())
        f.write(strip.astype(np.float32).tobytes())
        f.write(b.astype(np.float32).tobytes())
    r = subprocess.run([exe, "l1", str(sysf), str(xf)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    iters, conv, rn, rel = r.stdout.split()
    assert int(conv) == _abi.CONV_CONVERGED and float(rel) <= 1e-6
    x = np.fromfile(xf, dtype=np.float32).astype(np.float64)
    A32 = Ap.astype(np.float32).astype(np.float64)
    r32 = b.astype(np.float32).astype(np.float64) - A32 @ x
    assert np.linalg.norm(r32) <= 1e-5 * np.linalg.norm(b)
    # the same entry through the Python binding
    from project_superdex_amd import scene as gs
    kp = _abi.default_krylov_params()
    kp.abs_tol, kp.rel_tol, kp.rel_div_tol, kp.max_iter = 0.0, 1e-6, 1e10, 1000
    kp.norm_type, kp.preconditioner = _abi.NORM_RESIDUAL_L2, _abi.PREC_BLOCK_JACOBI
    kp.initial_guess_zero, kp.precision_mode, kp.refresh_every = 0, _abi.PREC_MODE_MIXED, 10
    xp, sp = gs.pcg_bsr3_solve(ptr, col, strip, b, params=kp, layout=_abi.BSR_REF_ROW_STRIP, dtype=np.float32)
    assert abs(sp["num_iter_done"] - int(iters)) <= 1
    assert np.linalg.norm(xp - x) <= 1e-5 * np.linalg.norm(xp)


@pytest.mark.gpu
def test_l2_snippet_steps_a_pressed_cube_like_the_oracle(tmp_path, oracle):
    exe = build(str(tmp_path))
    spec = mesh.plate_press_scene(n=4, edge=0.04, steps=6, press=1e-3, ramp_steps=6)
    body, plate = spec.bodies[0], spec.colliders[0]
    E, nu, rho, k = 1e5, 0.4, 1000.0, 1e6
    meshf, posf = tmp_path / "mesh.bin", tmp_path / "pos.bin"
    with open(meshf, "wb") as f:
        f.write(np.array([body.X.shape[0], body.tets.shape[0], len(body.pinned)], np.int32).tobytes())
        f.write(np.ascontiguousarray(body.X, np.float64).tobytes())
        f.write(np.ascontiguousarray(body.tets, np.int32).tobytes())
        f.write(np.ascontiguousarray(body.pinned, np.int32).tobytes())
        f.write(np.array([*plate.point, *plate.normal, *plate.velocity, E, nu, rho, k, spec.dt], np.float64).tobytes())
    r = subprocess.run([exe, "l2", str(meshf), str(posf), str(spec.steps)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    steps = [l for l in lines if l.startswith("step")]
    assert len(steps) == spec.steps
    for l in steps:
        f = l.split()
        assert int(f[7]) == _abi.CONV_STOPPED and int(f[9]) == _abi.CONV_CONVERGED and float(f[11]) <= 1e-6
    assert "snapshot equal 1" in lines
    x = np.fromfile(posf, dtype=np.float64).reshape(-1, 3)
    p = _abi.pinned_solver_params(1, lin_rel_tol=1e-6)
    so, _ = oracle.run_scene(spec, p, spec.steps, fp64=False)
    xo = body.X + so.displacements(0)
    err = np.max(np.linalg.norm(x - xo, axis=1) / np.maximum(np.linalg.norm(xo, axis=1), 1e-3))
    assert err <= 1e-4, err
    assert np.abs(x - body.X).max() > 1e-4  # the plate moved the cube
