"""SceneBatchExecutorV3 for the soft environments of one batched world.

The reference's executor (mochi_physics/pybind/src/scene_batch_executor_v2.cpp:55-1317, ABI
version 3) steps N scenes on a thread pool and exchanges state with the caller through flat
[num_scenes, total_dofs] arrays.  Here every scene is one environment of a batched
``scene.Scene`` (``Scene.env(i)``): a single ``sdx_step`` advances them all on the GPU, so
``num_workers`` only sets the reported worker count.  The call surface (constructor arguments,
properties, ``write_state``, ``write_boundary_conditions``, ``step``, ``step_control``,
``close``, readback mask bits, array shapes, argument checks and their messages) follows the
reference; the state it exchanges follows this path's actors:

- a soft actor slot has 3 n_verts DoFs, node by node (x, y, z): qpos the world positions of its
  nodes, qvel their velocities, boundary-condition values world coordinates held by the next
  step (AddBoundaryConditionDofsWorld after ClearBoundaryConditions, as the reference's
  write_boundary_conditions does per selected actor);
- a kinematic collider slot has 3 DoFs: qpos and boundary-condition values its position target
  for the next step, qvel the velocity of its last step ((target - previous target) / dt);
  qvel cannot be written (it follows the targets);
- neither kind has links or actuators (their counts must be 0, as for the reference's zero-DoF
  actors), and neither takes external forces: generalized_forces must be zero, the reference's
  "This actor does not support external forces." for soft actors
  (mochi_physics/src/mochi_actor.cpp:2226-2266);
- contacts (contact_kinds 0 touch flag, 1 force, 2 torque) read the world's contact query
  (sdx_get_contact_wrench): kind 0 is 1 when the source's contact points with `other` (any
  collider acting on the environment when None) include one at a signed distance <= the
  contact's distance; kind 1 the total contact force on the source (a soft actor: from every
  collider acting on it; a collider: from the environment's soft actor); kind 2 the torque on
  a collider source about its position (a soft source is rejected, "GetContactTorqueWorld only
  supported for rigid actors.", mochi_actor.cpp:2339-2345);
- diverged: the environment's step ended Diverged.
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi
from .scene import EnvScene, RigidActor, SoftActor

SCENE_BATCH_EXECUTOR_ABI_VERSION = 3


class SceneBatchExecutorV3:
    READ_QPOS = 1 << 0
    READ_QVEL = 1 << 1
    READ_LINKS = 1 << 2
    READ_CONTACTS = 1 << 3
    READ_DIVERGED = 1 << 4
    READ_ALL = READ_QPOS | READ_QVEL | READ_LINKS | READ_CONTACTS | READ_DIVERGED

    def __init__(self, scenes, actors, links, actuator_counts, contact_sources, contact_others, contact_kinds,
                 contact_distances, num_workers):
        scenes, actors, links = list(scenes), list(actors), list(links)
        contact_sources, contact_others = list(contact_sources), list(contact_others)
        if num_workers == 0:
            raise ValueError("SceneBatchExecutorV3 requires num_workers >= 1")
        if (len(scenes) == 0 or len(scenes) != len(actors) or len(scenes) != len(links)
                or len(scenes) != len(contact_sources) or len(scenes) != len(contact_others)):
            raise ValueError("scenes, actors, links, contact_sources, and contact_others must be "
                             "non-empty and have equal length")
        self._num_workers = min(int(num_workers), len(scenes))
        contact_kinds, contact_distances = list(contact_kinds), list(contact_distances)
        if len(contact_kinds) != len(contact_distances):
            raise ValueError("contact_kinds and contact_distances must have equal length")
        for k, d in zip(contact_kinds, contact_distances):
            if int(k) < 0 or int(k) > 2 or not math.isfinite(float(d)) or float(d) < 0:
                raise ValueError("contact kinds must be 0..2 and distances finite/non-negative")
        self._kinds = [int(k) for k in contact_kinds]
        self._distances = [float(d) for d in contact_distances]

        world = None
        seen = set()
        for sc in scenes:
            if not isinstance(sc, EnvScene) or sc in seen or (world is not None and sc.scene is not world):
                raise ValueError("SceneBatchExecutorV3 requires unique live scenes of one batched world")
            world = sc.scene
            seen.add(sc)
        self._world, self._scenes = world, scenes
        slot_count = -1
        used = set()
        self._slots = []  # per scene: the actor objects of its slots
        for i, sc in enumerate(scenes):
            scene_actors, scene_links = list(actors[i]), list(links[i])
            if len(scene_actors) == 0 or len(scene_actors) != len(scene_links):
                raise ValueError("each actors/links entry must be a non-empty equal-length pair")
            if slot_count < 0:
                slot_count = len(scene_actors)
            elif len(scene_actors) != slot_count:
                raise ValueError("all scenes must use the same actor slot count")
            row = []
            for a, lk in zip(scene_actors, scene_links):
                ok = (isinstance(a, SoftActor) and a.scene is world and a.index == sc.index) or \
                     (isinstance(a, RigidActor) and a.scene is world and a.actor == sc.index)
                if not ok or id(a) in used:
                    raise ValueError("each actor slot must contain a unique live actor in its scene")
                used.add(id(a))
                if len(list(lk)) != 0:
                    raise ValueError("link actors must be nested links of their actor slot")
                row.append(a)
            self._slots.append(row)
        actuator_counts = list(actuator_counts)
        if slot_count < 1 or len(actuator_counts) != slot_count:
            raise ValueError("actuator_counts must have one entry per actor slot")
        self._dof_counts, self._dof_offsets = [], []
        off, has_soft = 0, False
        for k in range(slot_count):
            n = self._dofs_of(self._slots[0][k])
            if int(actuator_counts[k]) < 0:
                raise ValueError("actor DoF, link, and actuator counts must be non-negative")
            if int(actuator_counts[k]) != 0:
                raise ValueError("soft actors and kinematic colliders have no actuators")
            for row in self._slots:
                if self._dofs_of(row[k]) != n or type(row[k]) is not type(self._slots[0][k]):
                    raise ValueError("all scenes must use the same actor slot DoF/link layout")
            has_soft = has_soft or isinstance(self._slots[0][k], SoftActor)
            self._dof_counts.append(n)
            self._dof_offsets.append(off)
            off += n
        if not has_soft:
            raise ValueError("SceneBatchExecutorV3 requires at least one soft actor slot")
        self._n_dofs = off
        self._slot_count = slot_count

        self._sources, self._others = [], []
        for i, sc in enumerate(scenes):
            srcs, oths = list(contact_sources[i]), list(contact_others[i])
            if len(srcs) != len(self._kinds) or len(srcs) != len(oths):
                raise ValueError("each contact source/other list must match contact_kinds")
            for c, (s, o) in enumerate(zip(srcs, oths)):
                if not self._in_scene(s, sc):
                    raise ValueError("contact sources must belong to their scene")
                if o is not None and not self._in_scene(o, sc):
                    raise ValueError("contact actors must belong to their scene")
                if self._kinds[c] == 2 and isinstance(s, SoftActor):
                    raise ValueError("GetContactTorqueWorld only supported for rigid actors.")
            self._sources.append(srcs)
            self._others.append(oths)
        if self._kinds:
            world.register_contact_query(True)
        self._prev_targets = {c.index: c.position.copy() for c in world.colliders}
        self._coll_velocity = {c.index: np.zeros(3) for c in world.colliders}
        self._closed = False

    # ---------------------------------------------------------------- layout
    @staticmethod
    def _dofs_of(a):
        return 3 * a.n_verts if isinstance(a, SoftActor) else 3

    @staticmethod
    def _in_scene(a, sc):
        if isinstance(a, SoftActor):
            return a.scene is sc.scene and a.index == sc.index
        if isinstance(a, RigidActor):
            return a.scene is sc.scene and a.actor in (-1, sc.index)
        return False

    abi_version = SCENE_BATCH_EXECUTOR_ABI_VERSION

    @property
    def num_workers(self):
        return self._num_workers

    @property
    def num_scenes(self):
        return len(self._scenes)

    @property
    def num_actors(self):
        return self._slot_count

    @property
    def num_dofs(self):
        return self._n_dofs

    @property
    def num_links(self):
        return 0

    @property
    def num_actuators(self):
        return 0

    @property
    def num_contacts(self):
        return len(self._kinds)

    @property
    def actor_dof_counts(self):
        return list(self._dof_counts)

    @property
    def actor_dof_offsets(self):
        return list(self._dof_offsets)

    actor_qpos_offsets = actor_dof_offsets
    actor_qvel_offsets = actor_dof_offsets

    @property
    def actor_link_counts(self):
        return [0] * self._slot_count

    @property
    def actor_link_offsets(self):
        return [0] * self._slot_count

    @property
    def actor_actuator_counts(self):
        return [0] * self._slot_count

    @property
    def actor_actuator_offsets(self):
        return [0] * self._slot_count

    @property
    def closed(self):
        return self._closed

    # ---------------------------------------------------------------- checks
    def _check_open(self):
        if self._closed:
            raise RuntimeError("SceneBatchExecutorV3 is closed")

    def _state_array(self, name, a, dtype=np.float64):
        a = np.asarray(a)
        if a.ndim != 2 or a.shape[0] != len(self._scenes) or a.shape[1] != self._n_dofs:
            raise ValueError(f"{name} must have shape [num_scenes, total_dofs]")
        return np.ascontiguousarray(a, dtype=dtype)

    @staticmethod
    def _finite(name, a):
        if not np.all(np.isfinite(a)):
            raise ValueError(f"{name} values must be finite")

    def _write_payload(self, name, values, mask):
        if mask is None:
            if values is not None:
                raise ValueError(f"{name} must be None when its mask is None")
            return None
        m = np.asarray(mask)
        if m.ndim != 2 or m.shape[0] != len(self._scenes) or m.shape[1] != self._n_dofs:
            raise ValueError(f"{name}_write_mask must have shape [num_scenes, total_dofs]")
        if values is None:
            raise ValueError(f"{name} is required when its mask is supplied")
        v = self._state_array(name, values)
        self._finite(name, v)
        if np.any(m > 1) or np.any(m < 0):
            raise ValueError(f"{name}_write_mask values must be 0 or 1")
        if not np.any(m):
            return None
        return v, m.astype(bool)

    def _readback(self, mask, qpos_out, qvel_out, link_state_out, contact_out, diverged_out):
        if mask & ~self.READ_ALL:
            raise ValueError("readback_mask contains unknown fields")

        def state(name, obj, on):
            if not on:
                if obj is not None:
                    raise ValueError(f"{name} must be None when readback is disabled")
                return None
            if obj is None:
                raise ValueError(f"{name} is required by readback_mask")
            a = np.asarray(obj)
            if a.ndim != 2 or a.shape[0] != len(self._scenes) or a.shape[1] != self._n_dofs:
                raise ValueError(f"{name} must have shape [num_scenes, total_dofs]")
            return obj

        out = {"qpos": state("qpos_out", qpos_out, mask & self.READ_QPOS),
               "qvel": state("qvel_out", qvel_out, mask & self.READ_QVEL)}
        for name, obj, bit, shape, msg in (
                ("links", link_state_out, self.READ_LINKS, (len(self._scenes), 0, 16),
                 "link_state_out must have shape [num_scenes, total_links, 16]"),
                ("contacts", contact_out, self.READ_CONTACTS, (len(self._scenes), len(self._kinds), 3),
                 "contact_out must have shape [num_scenes, num_contacts, 3]"),
                ("diverged", diverged_out, self.READ_DIVERGED, (len(self._scenes),),
                 "diverged_out must have shape [num_scenes]")):
            label = {"links": "link_state_out", "contacts": "contact_out", "diverged": "diverged_out"}[name]
            if not mask & bit:
                if obj is not None:
                    raise ValueError(f"{label} must be None when readback is disabled")
                out[name] = None
                continue
            if obj is None:
                raise ValueError(f"{label} is required by readback_mask")
            if np.asarray(obj).shape != shape:
                raise ValueError(msg)
            out[name] = obj
        return out

    # ---------------------------------------------------------------- state
    def write_state(self, qpos=None, qvel=None, qpos_write_mask=None, qvel_write_mask=None):
        qp = self._write_payload("qpos", qpos, qpos_write_mask)
        qv = self._write_payload("qvel", qvel, qvel_write_mask)
        if qp is None and qv is None:
            return
        self._check_open()
        try:
            for i, row in enumerate(self._slots):
                for k, a in enumerate(row):
                    sl = slice(self._dof_offsets[k], self._dof_offsets[k] + self._dof_counts[k])
                    if isinstance(a, SoftActor):
                        if qp is not None and qp[1][i, sl].any():
                            x = a.positions_world().reshape(-1)
                            m = qp[1][i, sl]
                            x[m] = qp[0][i, sl][m]
                            a.set_displacements(x - self._rest(a))
                        if qv is not None and qv[1][i, sl].any():
                            v = a.velocities().reshape(-1)
                            m = qv[1][i, sl]
                            v[m] = qv[0][i, sl][m]
                            a.set_velocities(v)
                    else:
                        if qv is not None and qv[1][i, sl].any():
                            raise ValueError("a kinematic collider's velocity follows its position targets")
                        if qp is not None and qp[1][i, sl].any():
                            p = a.position.copy()
                            m = qp[1][i, sl]
                            p[m] = qp[0][i, sl][m]
                            a.set_position(p)
        except Exception:
            self.close()
            raise

    def write_boundary_conditions(self, values=None, write_mask=None):
        payload = self._write_payload("boundary_condition_values", values, write_mask)
        if payload is None:
            return
        self._check_open()
        vals, mask = payload
        try:
            for i, row in enumerate(self._slots):
                for k, a in enumerate(row):
                    sl = slice(self._dof_offsets[k], self._dof_offsets[k] + self._dof_counts[k])
                    m = mask[i, sl]
                    if not m.any():
                        continue
                    if isinstance(a, SoftActor):
                        dofs = np.nonzero(m)[0].astype(np.int32)
                        a.set_boundary_conditions(dofs, vals[i, sl][m], clear=True)
                    else:
                        p = a.position.copy()
                        p[m] = vals[i, sl][m]
                        a.set_position(p)
        except Exception:
            self.close()
            raise

    def _rest(self, a):
        if not hasattr(self, "_rest_cache"):
            self._rest_cache = {}
        if a.index not in self._rest_cache:
            self._rest_cache[a.index] = (a.positions_world() - a.displacements()).reshape(-1)
        return self._rest_cache[a.index]

    # ---------------------------------------------------------------- stepping
    @property
    def solver_params(self) -> _abi.SolverParams:
        return getattr(self._world, "solver_params", None) or _abi.default_solver_params()

    def _advance(self, dt):
        stats = self._world.step(dt, self.solver_params, as_array=True)
        for c in self._world.colliders:
            self._coll_velocity[c.index] = (c.position - self._prev_targets[c.index]) / dt
            self._prev_targets[c.index] = c.position.copy()
        return stats

    def _fill(self, stats, out, dt):
        if out["diverged"] is not None:
            div = out["diverged"]
            for i, sc in enumerate(self._scenes):
                div[i] = 1 if stats[sc.index]["convergence"] == _abi.CONV_DIVERGED else 0
        if out["qpos"] is not None or out["qvel"] is not None:
            for i, row in enumerate(self._slots):
                for k, a in enumerate(row):
                    sl = slice(self._dof_offsets[k], self._dof_offsets[k] + self._dof_counts[k])
                    if isinstance(a, SoftActor):
                        if out["qpos"] is not None:
                            out["qpos"][i, sl] = a.positions_world().reshape(-1)
                        if out["qvel"] is not None:
                            out["qvel"][i, sl] = a.velocities().reshape(-1)
                    else:
                        if out["qpos"] is not None:
                            out["qpos"][i, sl] = a.position
                        if out["qvel"] is not None:
                            out["qvel"][i, sl] = self._coll_velocity[a.index]
        if out["contacts"] is not None:
            co = out["contacts"]
            for i, sc in enumerate(self._scenes):
                acting = [c for c in self._world.colliders if c.actor in (-1, sc.index)]
                for c, kind in enumerate(self._kinds):
                    src, oth = self._sources[i][c], self._others[i][c]
                    co[i, c, :] = 0.0
                    if kind == 0:
                        if isinstance(src, SoftActor):
                            cands = [oth] if isinstance(oth, RigidActor) else (acting if oth is None else [])
                        else:
                            cands = [src] if (oth is None or isinstance(oth, SoftActor)) else []
                        for col in cands:
                            if self._world.contact_wrench(sc.index, col.index)[6] <= self._distances[c]:
                                co[i, c, 0] = 1.0
                                break
                    elif kind == 1:
                        if isinstance(src, SoftActor):
                            co[i, c, :] = sum(self._world.contact_wrench(sc.index, col.index)[:3] for col in acting) \
                                if acting else 0.0
                        else:
                            co[i, c, :] = -self._world.contact_wrench(sc.index, src.index)[:3]
                    else:
                        co[i, c, :] = self._world.contact_wrench(sc.index, src.index)[3:6]

    def step(self, time_step_sec, generalized_forces, qpos_out=None, qvel_out=None, link_state_out=None,
             contact_out=None, diverged_out=None, readback_mask=READ_ALL):
        dt = float(time_step_sec)
        if not math.isfinite(dt) or dt < 0:
            raise ValueError("step requires a finite, non-negative time_step_sec")
        if dt == 0:
            raise ValueError("step needs a positive time_step_sec on this path")
        f = self._state_array("generalized_forces", generalized_forces)
        self._finite("generalized_forces", f)
        if np.any(f != 0):
            raise ValueError("This actor does not support external forces.")
        self._check_open()
        out = self._readback(int(readback_mask), qpos_out, qvel_out, link_state_out, contact_out, diverged_out)
        try:
            stats = self._advance(dt)
            self._fill(stats, out, dt)
        except Exception:
            self.close()
            raise

    def step_control(self, time_step_sec, controls, qpos_indices, qvel_indices, kp, kd, gear, force_ranges,
                     num_steps, qpos_out=None, qvel_out=None, link_state_out=None, contact_out=None,
                     diverged_out=None, readback_mask=READ_ALL):
        if int(num_steps) < 1:
            raise ValueError("step_control requires positive num_steps")
        dt = float(time_step_sec)
        if not math.isfinite(dt) or dt < 0:
            raise ValueError("step_control requires a finite, non-negative time_step_sec")
        if dt == 0:
            raise ValueError("step_control needs a positive time_step_sec on this path")
        c = np.asarray(controls)
        if c.ndim != 2 or c.shape[0] != len(self._scenes) or c.shape[1] != 0:
            raise ValueError("controls must have shape [num_scenes, num_actuators]")
        for a in (qpos_indices, qvel_indices, kp, kd, gear):
            if np.asarray(a).shape != (0,):
                raise ValueError("control metadata must match the flattened actuator layout")
        if np.asarray(force_ranges).shape != (0, 2):
            raise ValueError("control metadata must match the flattened actuator layout")
        out = self._readback(int(readback_mask), qpos_out, qvel_out, link_state_out, contact_out, diverged_out)
        self._check_open()
        try:
            stats = None
            for _ in range(int(num_steps)):
                stats = self._advance(dt)
            self._fill(stats, out, dt)
        except Exception:
            self.close()
            raise

    def close(self):
        self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
