"""Python mirror of the reference's soft-body call surface, bound to libsdxfem.

Mirrors the shape of the reference API (paths under
superdex_physics/libraries/mochi/mochi_physics/include/mochi_physics/cpp_api/):
  Scene::Step(dt) ......................... Scene.step                 (mochi_scene.h:63)
  Scene::CreateSoftActor(SoftActorParams) . Scene.create_soft_actor    (mochi_scene.h:99)
  Actor::AddBoundaryConditionNodesWorldPermanent
                                             SoftActor.add_boundary_condition_nodes (mochi_actor.h:202)
  CreateRigidActor(plane / sphere shape) .. Scene.create_rigid_plane / create_rigid_sphere
  Actor::SetRootTransform (translation) ... RigidActor.set_position     (mochi_actor.h:79)
  Actor::GetDofValues / velocities ........ SoftActor.displacements / velocities
Every call goes through the C ABI of include/sdx_fem.h; there is no
Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _abi
from ._lib import check, lib

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(_ip)


class SoftActor:
    def __init__(self, scene: "Scene", index: int, n_verts: int, n_tets: int):
        self.scene, self.index, self.n_verts, self.n_tets = scene, index, n_verts, n_tets

    def add_boundary_condition_nodes(self, nodes):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        check(lib().sdx_pin_nodes(self.scene._w, self.index, _i(nodes), len(nodes)))

    def set_element_materials(self, E, nu):
        E = np.ascontiguousarray(E, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        check(lib().sdx_set_element_materials(self.scene._w, self.index, _d(E), _d(nu), len(E)))

    def _get(self, fn):
        out = np.zeros(3 * self.n_verts)
        check(fn(self.scene._w, self.index, out.ctypes.data_as(_dp), out.size))
        return out.reshape(-1, 3)

    def displacements(self):
        return self._get(lib().sdx_get_displacements)

    def velocities(self):
        return self._get(lib().sdx_get_velocities)

    def positions_world(self):
        return self._get(lib().sdx_get_positions_world)

    def set_displacements(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        check(lib().sdx_set_displacements(self.scene._w, self.index, u.ctypes.data_as(_dp), u.size))

    def set_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        check(lib().sdx_set_velocities(self.scene._w, self.index, v.ctypes.data_as(_dp), v.size))

    def set_boundary_conditions(self, dofs, values_world, clear: bool = True):
        """AddBoundaryConditionDofsWorld (after ClearBoundaryConditions with clear): hold
        DoFs 3 node + axis at world coordinates from the next step on."""
        dofs = np.ascontiguousarray(dofs, dtype=np.int32).reshape(-1)
        vals = np.ascontiguousarray(values_world, dtype=np.float64).reshape(-1)
        if dofs.size != vals.size:
            raise ValueError("dofs and values must have the same length")
        check(lib().sdx_set_boundary_conditions(self.scene._w, self.index, _i(dofs), _d(vals), dofs.size,
                                                int(clear)))

    def bsr_size(self):
        nr, nb = C.c_int32(), C.c_int32()
        check(lib().sdx_actor_bsr_size(self.scene._w, self.index, C.byref(nr), C.byref(nb)))
        return nr.value, nb.value

    def assemble(self, dt: float, params: _abi.SolverParams):
        """Residual and Dirichlet-applied BSR (row-major 3x3 blocks) at the current displacement."""
        nr, nb = self.bsr_size()
        res = np.zeros(3 * nr)
        ptr = np.zeros(nr + 1, dtype=np.int32)
        col = np.zeros(nb, dtype=np.int32)
        val = np.zeros(9 * nb)
        check(lib().sdx_assemble(self.scene._w, self.index, dt, C.byref(params), res.ctypes.data_as(_dp),
                                 ptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip), val.ctypes.data_as(_dp)))
        return res, ptr, col, val.reshape(nb, 3, 3)

    def objective(self, dt: float, params: _abi.SolverParams) -> float:
        """Objective (merit) at the current displacement, stage start as in assemble()."""
        out = C.c_double()
        check(lib().sdx_objective(self.scene._w, self.index, dt, C.byref(params), C.byref(out)))
        return out.value


class RigidActor:
    def __init__(self, scene: "Scene", index: int, kind: str = "plane", position=(0.0, 0.0, 0.0), actor: int = -1):
        self.scene, self.index, self.kind, self.actor = scene, index, kind, actor
        scene._put_collider_position(index, position)

    @property
    def position(self) -> np.ndarray:
        """The current target (a copy)."""
        return self.scene._coll_pos[self.index].copy()

    def set_position(self, pos):
        pos = np.asarray(pos, float)
        check(lib().sdx_set_rigid_position(self.scene._w, self.index, _d(pos)))
        self.scene._coll_pos[self.index] = pos


class EnvScene:
    """One environment of a batched Scene: its soft actor and the colliders attached to it."""

    def __init__(self, scene: "Scene", index: int):
        self.scene, self.index = scene, index

    @property
    def soft_actor(self) -> SoftActor:
        return self.scene.actors[self.index]

    @property
    def colliders(self) -> list:
        return [c for c in self.scene.colliders if c.actor == self.index]

    def __eq__(self, other):
        return isinstance(other, EnvScene) and other.scene is self.scene and other.index == self.index

    def __hash__(self):
        return hash((id(self.scene), self.index))


class Scene:
    """One world on one HIP device: soft actors, kinematic colliders, step()."""

    def __init__(self, device: int = 0, precision: int = _abi.PREC_MODE_FP32):
        w = C.c_void_p()
        check(lib().sdx_world_create(device, precision, C.byref(w)))
        self._w = w
        self.precision = precision
        self.actors: list[SoftActor] = []
        self.colliders: list[RigidActor] = []
        # collider targets, row k for collider k (capacity doubles as colliders are added)
        self._coll_pos_buf = np.zeros((0, 3))
        self.solver_params: _abi.SolverParams | None = None  # what SceneBatchExecutorV3 steps with

    @property
    def _coll_pos(self) -> np.ndarray:
        return self._coll_pos_buf[:len(self.colliders)]

    def _put_collider_position(self, index: int, position):
        if index >= len(self._coll_pos_buf):
            grown = np.zeros((max(16, 2 * len(self._coll_pos_buf), index + 1), 3))
            grown[:len(self._coll_pos_buf)] = self._coll_pos_buf
            self._coll_pos_buf = grown
        self._coll_pos_buf[index] = np.asarray(position, dtype=np.float64)

    def close(self):
        if getattr(self, "_w", None):
            lib().sdx_world_destroy(self._w)
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def create_soft_actor(self, X, tets, params: _abi.SoftParams | None = None) -> SoftActor:
        X = np.ascontiguousarray(X, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        params = params or _abi.default_soft_params()
        idx = C.c_int32()
        check(lib().sdx_add_soft_tet_mesh(self._w, _d(X), X.shape[0], _i(tets), tets.shape[0], C.byref(params),
                                          C.byref(idx)))
        a = SoftActor(self, idx.value, X.shape[0], tets.shape[0])
        self.actors.append(a)
        return a

    def create_rigid_plane(self, point, normal, actor: int = -1) -> RigidActor:
        idx = C.c_int32()
        check(lib().sdx_add_rigid_plane(self._w, _d(np.asarray(point, float)), _d(np.asarray(normal, float)),
                                        actor, C.byref(idx)))
        r = RigidActor(self, idx.value, "plane", point, actor)
        self.colliders.append(r)
        return r

    def create_rigid_sphere(self, center, radius: float, actor: int = -1) -> RigidActor:
        idx = C.c_int32()
        check(lib().sdx_add_rigid_sphere(self._w, _d(np.asarray(center, float)), float(radius), actor,
                                         C.byref(idx)))
        r = RigidActor(self, idx.value, "sphere", center, actor)
        self.colliders.append(r)
        return r

    def set_rigid_positions(self, colliders, positions):
        """Move colliders[k] (RigidActor or index) to positions[k] in one call."""
        if isinstance(colliders, np.ndarray):
            idx = np.ascontiguousarray(colliders, dtype=np.int32)
        else:
            idx = np.ascontiguousarray([c.index if isinstance(c, RigidActor) else int(c) for c in colliders],
                                       dtype=np.int32)
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(len(idx), 3)
        check(lib().sdx_set_rigid_positions(self._w, len(idx), idx.ctypes.data_as(_ip), pos.ctypes.data_as(_dp)))
        self._coll_pos_buf[idx] = pos  # (the library checked every index)

    def env(self, index: int) -> "EnvScene":
        """Environment `index` (soft actor `index` and the colliders attached to it) as a scene
        of its own, the unit of SceneBatchExecutorV3."""
        if not 0 <= index < len(self.actors):
            raise IndexError("no such environment")
        return EnvScene(self, index)

    def register_contact_query(self, enable: bool = True):
        """CQueryActorContactForces: contact wrenches available after the next step."""
        check(lib().sdx_register_contact_query(self._w, int(enable)))

    def contact_wrench(self, actor: int, collider: int) -> np.ndarray:
        """Force on the actor from the collider (3), torque of the reaction on the collider about
        its position (3), least signed distance of the contact points (inf: none), last step."""
        out = np.zeros(7)
        check(lib().sdx_get_contact_wrench(self._w, int(actor), int(collider), out.ctypes.data_as(_dp)))
        return out

    def step(self, dt: float, params: _abi.SolverParams | None = None, as_array: bool = False):
        """One step; per-actor statistics as dicts, or with as_array as one
        numpy structured array (fields of sdx_step_stats, no per-actor copies)."""
        params = params or _abi.default_solver_params()
        n = len(self.actors)
        if as_array:  # written in place: no per-step ctypes conversion
            out = np.zeros(max(1, n), dtype=_stats_dtype())
            check(lib().sdx_step(self._w, dt, C.byref(params), out.ctypes.data_as(C.POINTER(_abi.StepStats)), n))
            return out[:n]
        stats = (_abi.StepStats * max(1, n))()
        check(lib().sdx_step(self._w, dt, C.byref(params), stats, n))
        return [s.as_dict() for s in stats[:n]]

    def info(self):
        out = (C.c_int64 * 8)()
        check(lib().sdx_world_info(self._w, out))
        return {"n_verts": out[0], "n_tets": out[1], "nnzb": out[2], "n_tris": out[3], "n_envs": out[4],
                "matvec_tiles": out[5], "vector_tiles": out[6], "cg_groups": out[7]}

    # State snapshots (Scene::CaptureState / RestoreState / ReleaseState /
    # IsEqualState / CaptureStateToBytes / RestoreStateFromBytes /
    # CaptureStateToFile, mochi_physics/include/mochi_physics/cpp_api/mochi_scene.h:79-95;
    # capture::RestorePartialState / CaptureStateToJson, mochi_physics/src/mochi_capture.h:38-56)
    def capture_state(self) -> int:
        """Snapshot of the dynamic state kept in device memory; returns its handle."""
        h = C.c_int32()
        check(lib().sdx_capture_state(self._w, C.byref(h)))
        return h.value

    def _pull_collider_positions(self):
        """Re-read every collider target from the world (a restore sets them), in one call."""
        n = len(self.colliders)
        if n == 0:
            return
        ids = np.arange(n, dtype=np.int32)
        out = np.zeros((n, 3))
        check(lib().sdx_get_rigid_positions(self._w, n, _i(ids), out.ctypes.data_as(C.POINTER(C.c_double))))
        self._coll_pos_buf[:n] = out

    def restore_state(self, handle: int, release_immediately: bool = False):
        check(lib().sdx_restore_state(self._w, int(handle), int(bool(release_immediately))))
        self._pull_collider_positions()

    def release_state(self, handle: int):
        check(lib().sdx_release_state(self._w, int(handle)))

    def release_all_states(self):
        check(lib().sdx_release_state(self._w, 0))

    def is_equal_state(self, a: int, b: int) -> bool:
        eq = C.c_int32()
        check(lib().sdx_is_equal_state(self._w, int(a), int(b), C.byref(eq)))
        return bool(eq.value)

    def capture_state_to_bytes(self) -> bytes:
        n = C.c_int64()
        check(lib().sdx_state_size(self._w, C.byref(n)))
        out = bytearray(n.value)
        check(lib().sdx_capture_state_to_bytes(self._w, (C.c_uint8 * n.value).from_buffer(out), n.value))
        return bytes(out)

    def restore_state_from_bytes(self, data: bytes):
        self.restore_partial_state_from_bytes(data, 0)

    def restore_partial_state_from_bytes(self, data: bytes, excluded: int):
        """Restore every component except the kinds in `excluded` (_abi.STATE_* bits)."""
        data = bytes(data)  # a bytes object passes its own buffer, uncopied
        check(lib().sdx_restore_partial_state_from_bytes(self._w, data, len(data), int(excluded)))
        self._pull_collider_positions()

    def capture_state_to_file(self, path: str):
        check(lib().sdx_capture_state_to_file(self._w, os.fsencode(path)))

    def capture_state_to_json(self, pretty: bool = False) -> str:
        n = C.c_int64()
        check(lib().sdx_capture_state_to_json(self._w, int(pretty), None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().sdx_capture_state_to_json(self._w, int(pretty), buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def set_cg_burst(self, burst: int):
        check(lib().sdx_set_cg_burst(self._w, burst))

    def schedule(self) -> dict:
        sc = _abi.Schedule()
        check(lib().sdx_get_schedule(self._w, C.byref(sc)))
        return sc.as_dict()

    def set_schedule(self, **kw):
        sc = _abi.Schedule()
        check(lib().sdx_get_schedule(self._w, C.byref(sc)))
        for k, v in kw.items():
            setattr(sc, k, int(v))
        check(lib().sdx_set_schedule(self._w, C.byref(sc)))

    def set_warm_start_model(self, model):
        """Snapshot `model`'s parameters as the learned CG warm start (None removes it)."""
        check(lib().sdx_world_set_warm_start_model(self._w, model._m if model is not None else None))

    def collect(self, capacity: int, per_env: int = 64):
        """Collect (feature row, Newton step / scale) pairs during the next steps."""
        check(lib().sdx_world_collect(self._w, int(capacity), int(per_env)))

    def collected(self, copy: bool = True):
        """Rows collected so far: (X[n][16], T[n][3]) or their count."""
        n = C.c_int64()
        check(lib().sdx_world_collected(self._w, C.byref(n), None, None, 0))
        if not copy:
            return n.value
        X = np.zeros((n.value, _abi.MLP_ROW), dtype=np.float32)
        T = np.zeros((n.value, _abi.MLP_OUT), dtype=np.float32)
        check(lib().sdx_world_collected(self._w, C.byref(n), X.ctypes.data_as(C.POINTER(C.c_float)),
                                        T.ctypes.data_as(C.POINTER(C.c_float)), X.shape[0]))
        return X, T

    def autotune(self, dt: float, params: _abi.SolverParams, trials: int = 2):
        """Time candidate schedules on the current state (restored afterwards);
        returns (chosen schedule, its step time in ms)."""
        sc = _abi.Schedule()
        ms = C.c_double()
        check(lib().sdx_autotune(self._w, dt, C.byref(params), trials, C.byref(sc), C.byref(ms)))
        return sc.as_dict(), ms.value

    def kernel_timing(self, kernel: int, enable: int = -1):
        """enable 1 clears and starts timing kernel class `kernel`; returns (launches, total ms)."""
        out = (C.c_double * 2)()
        check(lib().sdx_kernel_timing(self._w, kernel, enable, out))
        return out[0], out[1]


_STATS_DTYPE = None


def is_equal_state_bytes(a: bytes, b: bytes) -> bool:
    """capture::IsEqualState on two byte forms: equal values, whatever the bytes."""
    ba = (C.c_uint8 * max(1, len(a))).from_buffer_copy(a or b"\0")
    bb = (C.c_uint8 * max(1, len(b))).from_buffer_copy(b or b"\0")
    eq = C.c_int32()
    check(lib().sdx_is_equal_state_bytes(ba, len(a), bb, len(b), C.byref(eq)))
    return bool(eq.value)


def state_bytes_to_json(data: bytes, pretty: bool = False) -> str:
    """The JSON document of a byte form (no world or device needed)."""
    buf = (C.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
    n = C.c_int64()
    check(lib().sdx_state_bytes_to_json(buf, len(data), int(pretty), None, 0, C.byref(n)))
    out = C.create_string_buffer(n.value + 1)
    check(lib().sdx_state_bytes_to_json(buf, len(data), int(pretty), out, n.value + 1, C.byref(n)))
    return out.value.decode()



def _stats_dtype() -> np.dtype:
    """numpy layout of sdx_step_stats (built once: deriving it from the ctypes
    structure costs tens of microseconds)."""
    global _STATS_DTYPE
    if _STATS_DTYPE is None:
        _STATS_DTYPE = np.dtype(_abi.StepStats)
    return _STATS_DTYPE


def pcg_bsr3_solve(indptr, indices, blocks, b, x0=None, params: _abi.KrylovParams | None = None,
                   layout: int = _abi.BSR_BLOCK_ROWMAJOR, device: int = 0, dtype=np.float64):
    """L1: solve A x = b on the GPU; blocks (nnzb, 3, 3) row-major (or strip layout values).
    dtype float64 calls sdx_pcg_bsr3_solve, float32 sdx_pcg_bsr3_solve_f32 (the arrays of the
    reference's default real = float build); x comes back in that type."""
    dtype = np.dtype(dtype)
    if dtype == np.float64:
        fn, ptr_t = lib().sdx_pcg_bsr3_solve, _dp
    elif dtype == np.float32:
        fn, ptr_t = lib().sdx_pcg_bsr3_solve_f32, C.POINTER(C.c_float)
    else:
        raise ValueError("dtype must be float64 or float32")
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    vals = np.ascontiguousarray(blocks, dtype=dtype).reshape(-1)
    b = np.ascontiguousarray(b, dtype=dtype)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=dtype)
    params = params or _abi.default_krylov_params()
    st = _abi.LinearStatus()
    check(fn(device, len(indptr) - 1, len(indices), _i(indptr), _i(indices), vals.ctypes.data_as(ptr_t), layout,
             b.ctypes.data_as(ptr_t), x.ctypes.data_as(ptr_t), C.byref(params), C.byref(st)))
    return x, {"num_iter_done": st.num_iter_done, "convergence": st.convergence,
               "residual_norm": st.residual_norm, "relative_residual_norm": st.relative_residual_norm}


def build_scene(spec, device: int = 0, precision: int = _abi.PREC_MODE_FP32):
    """Scene of a mesh.SceneSpec; returns (scene, rigid actors in spec order)."""
    sc = Scene(device, precision)
    for b in spec.bodies:
        a = sc.create_soft_actor(b.X, b.tets, b.soft_params(_abi))
        if np.ndim(b.E):
            a.set_element_materials(b.E, b.nu)
        if len(b.pinned):
            a.add_boundary_condition_nodes(b.pinned)
    rigid = []
    for c in spec.colliders:
        if c.kind == "plane":
            rigid.append(sc.create_rigid_plane(c.point, c.normal, c.actor))
        else:
            rigid.append(sc.create_rigid_sphere(c.point, c.radius, c.actor))
    return sc, rigid
