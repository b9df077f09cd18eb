"""GPU: the allocation-free L1 solve (sdx_pcg_bsr3_solve_noalloc) against
sdx_pcg_bsr3_solve, called through the C ABI.

Two solves in a row on one device, the second on a smaller system (the
config-1 cube's assembled Newton system, 1331 block rows, then the
reference's 100x100 golden system padded to 34 block rows), in every
precision mode: both entry points converge with the same status, their
iteration counts differ by at most 1 (the CG's dot products add their partial
sums in arrival order), and their solutions agree within 1e-5 of |x| (FP32,
mixed) or 1e-10 (FP64).  The row-strip layout gives the same solution.  A
system beyond the scratch array is rejected with SDX_ERR_INVALID_ARGUMENT
before any device work.
"""
import ctypes as C
import os

import numpy as np
import pytest

from project_superdex_amd import _abi, mesh
from project_superdex_amd._lib import lib

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "krylov_psd100.npz")
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def golden_system():
    k = np.load(GOLD)
    A = np.zeros((100, 100))
    A[k["rows"], k["cols"]] = k["vals"]
    Ap = np.eye(102)
    Ap[:100, :100] = A
    indptr, indices, blocks = [0], [], []
    for bi in range(34):
        for bj in range(34):
            blk = Ap[3 * bi:3 * bi + 3, 3 * bj:3 * bj + 3]
            if np.any(blk != 0) or bi == bj:
                indices.append(bj)
                blocks.append(blk)
        indptr.append(len(indices))
    b = np.concatenate([k["rhs"], [0.0, 0.0]])
    return np.array(indptr, np.int32), np.array(indices, np.int32), np.array(blocks), b


def cube_system(oracle):
    """The oracle's assembled Newton system of the 10^3-hex cube at a seeded displacement."""
    spec = mesh.cube_drop_scene(10)
    sc, _ = oracle.build_scene(spec)
    rng = np.random.default_rng(7)
    u = 1e-4 * rng.standard_normal((spec.num_verts, 3))
    u[spec.bodies[0].pinned] = 0
    sc.set_displacements(0, u)
    res, ptr, col, val = sc.assemble(0, 1e-3, _abi.pinned_solver_params())
    return ptr, col, val, res


def call(name, ptr, col, vals, b, kp, layout=_abi.BSR_BLOCK_ROWMAJOR):
    vals = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b)
    st = _abi.LinearStatus()
    rc = getattr(lib(), name)(0, len(ptr) - 1, len(col), ptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                              vals.ctypes.data_as(_dp), layout, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp),
                              C.byref(kp), C.byref(st))
    return rc, x, st


def kparams(mode, rel=1e-6):
    kp = _abi.default_krylov_params()
    kp.rel_tol, kp.abs_tol, kp.max_iter = rel, 0.0, 2000
    kp.norm_type, kp.precision_mode, kp.preconditioner = _abi.NORM_RESIDUAL_L2, mode, _abi.PREC_BLOCK_JACOBI
    kp.refresh_every = 10 if mode == _abi.PREC_MODE_MIXED else 0
    return kp


@pytest.mark.parametrize("mode,tol", [(_abi.PREC_MODE_FP32, 1e-5), (_abi.PREC_MODE_MIXED, 1e-5),
                                      (_abi.PREC_MODE_FP64, 1e-10)])
def test_noalloc_matches_allocating_solve_on_two_calls(oracle, mode, tol):
    systems = [cube_system(oracle), golden_system()]  # the second call is the smaller system
    assert len(systems[1][0]) < len(systems[0][0])
    for ptr, col, vals, b in systems:
        kp = kparams(mode, 1e-10 if mode == _abi.PREC_MODE_FP64 else 1e-6)
        rc0, x0, s0 = call("sdx_pcg_bsr3_solve", ptr, col, vals, b, kp)
        rc1, x1, s1 = call("sdx_pcg_bsr3_solve_noalloc", ptr, col, vals, b, kp)
        print(f"mode {mode}: {len(ptr) - 1} rows, iterations {s0.num_iter_done} / {s1.num_iter_done}")
        assert rc0 == 0 and rc1 == 0, lib().sdx_last_error()
        assert s0.convergence == s1.convergence == _abi.CONV_CONVERGED
        assert abs(s0.num_iter_done - s1.num_iter_done) <= 1
        assert np.linalg.norm(x1 - x0) <= tol * np.linalg.norm(x0)


def test_noalloc_row_strip_layout_and_repeat(oracle):
    ptr, col, vals, b = cube_system(oracle)
    strip = np.concatenate([np.concatenate(list(vals[ptr[i]:ptr[i + 1]]), axis=1).reshape(-1)
                            for i in range(len(ptr) - 1)])
    kp = kparams(_abi.PREC_MODE_FP64, 1e-10)
    rc0, x0, s0 = call("sdx_pcg_bsr3_solve_noalloc", ptr, col, vals, b, kp)
    rc1, x1, s1 = call("sdx_pcg_bsr3_solve_noalloc", ptr, col, strip, b, kp, layout=_abi.BSR_REF_ROW_STRIP)
    assert rc0 == 0 and rc1 == 0, lib().sdx_last_error()
    assert s0.convergence == s1.convergence == _abi.CONV_CONVERGED
    assert np.linalg.norm(x1 - x0) <= 1e-10 * np.linalg.norm(x0)


def test_noalloc_rejects_a_system_beyond_its_scratch():
    n = 400_000  # diagonal blocks only: about 80 MB of device scratch in mixed precision
    ptr = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    vals = np.tile(np.eye(3), (n, 1, 1))
    b = np.ones(3 * n)
    rc, x, st = call("sdx_pcg_bsr3_solve_noalloc", ptr, col, vals, b, kparams(_abi.PREC_MODE_MIXED))
    assert rc == _abi.SDX_ERR_INVALID_ARGUMENT
    assert b"too large" in lib().sdx_last_error()
    assert np.all(x == 0)
