"""ctypes binding of the CPU oracle (oracle/liboracle_sdx.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from project_superdex_amd import _abi
from project_superdex_amd.mesh import SceneSpec

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sdx.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.sdxo_scene_create.restype = C.c_void_p
        L.sdxo_scene_create.argtypes = [C.c_int]
        L.sdxo_scene_destroy.argtypes = [C.c_void_p]
        L.sdxo_set_threads.argtypes = [C.c_int]
        L.sdxo_add_soft_tet_mesh.argtypes = [C.c_void_p, _dp, C.c_int, _ip, C.c_int,
                                             C.POINTER(_abi.SoftParams)]
        L.sdxo_set_element_materials.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_int]
        L.sdxo_pin_nodes.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        L.sdxo_add_collider.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_int]
        L.sdxo_set_collider_position.argtypes = [C.c_void_p, C.c_int, _dp]
        L.sdxo_step.argtypes = [C.c_void_p, C.c_double, C.POINTER(_abi.SolverParams),
                                C.POINTER(_abi.StepStats), C.c_int]
        L.sdxo_get_state.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.sdxo_set_state.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.sdxo_actor_info.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        L.sdxo_assemble.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(_abi.SolverParams),
                                    _dp, _dp, _dp, _ip, _ip, _dp]
        L.sdxo_boundary.argtypes = [C.c_void_p, C.c_int, _ip, _dp]
        L.sdxo_set_contact_query.argtypes = [C.c_void_p, C.c_int]
        L.sdxo_get_contact_wrench.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.sdxo_set_boundary_conditions.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int, C.c_int]
        L.sdxo_objective.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(_abi.SolverParams),
                                     _dp, _dp, _dp]
        L.sdxo_newton_synthetic.argtypes = [C.c_int, C.c_int, _dp, C.POINTER(_abi.SolverParams), _dp,
                                            C.POINTER(_abi.StepStats), _dp]
        L.sdxo_forcing_term.restype = C.c_double
        L.sdxo_forcing_term.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double]
        L.sdxo_smith_nh.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, C.c_int]
        L.sdxo_polyrelu.argtypes = [C.c_double, C.c_double, C.c_double, _dp]
        L.sdxo_stress_element.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp]
        L.sdxo_stress_damping_element.restype = C.c_double
        L.sdxo_stress_damping_element.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int,
                                                  C.c_int, _dp, _dp]
        L.sdxo_pcg_csr.argtypes = [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_int,
                                   C.POINTER(_abi.KrylovParams), _ip, _ip, _dp]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(_ip)


def set_threads(n: int) -> int:
    return lib().sdxo_set_threads(n)


def smith_nh(F, mu, lam, fp64=True, project=False):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(9)
    e = C.c_double()
    P = np.zeros(9)
    Cm = np.zeros(81)
    lib().sdxo_smith_nh(int(fp64), mu, lam, _d(F), C.byref(e), P.ctypes.data_as(_dp),
                        Cm.ctypes.data_as(_dp), int(project))
    return e.value, P.reshape(3, 3), Cm.reshape(3, 3, 3, 3)


def polyrelu(x, t, shift):
    out = np.zeros(3)
    lib().sdxo_polyrelu(x, t, shift, out.ctypes.data_as(_dp))
    return tuple(out)


def stress_element(X, u, mu, lam, project=True):
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(12)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(12)
    r = np.zeros(12)
    H = np.zeros(144)
    lib().sdxo_stress_element(_d(X), _d(u), mu, lam, int(project), r.ctypes.data_as(_dp),
                              H.ctypes.data_as(_dp))
    return r, H.reshape(12, 12)


def newton_synthetic(kind: int, x0, params: _abi.SolverParams):
    """The Newton driver on a synthetic problem of the reference's Newton tests:
    kind 0 the scalar problem, 1 the tridiagonal Broyden problem; returns (x, stats, merit)."""
    x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
    x = np.zeros_like(x0)
    st = _abi.StepStats()
    m = C.c_double()
    rc = lib().sdxo_newton_synthetic(kind, x0.size, _d(x0), C.byref(params), x.ctypes.data_as(_dp), C.byref(st),
                                     C.byref(m))
    assert rc == 0, rc
    return x, st, m.value


def forcing_term(strategy, k, res_norm, prev_res_norm, lin_res_norm, prev_eta, rel_tol):
    return lib().sdxo_forcing_term(strategy, k, res_norm, prev_res_norm, lin_res_norm, prev_eta, rel_tol)


def stress_damping_element(X, u, us, kappa, mu, lam, geo=False, project=False):
    """Stiffness damping of one tet: (energy, residual[12], tangent[12, 12])."""
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(12)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(12)
    us = np.ascontiguousarray(us, dtype=np.float64).reshape(12)
    r = np.zeros(12)
    H = np.zeros(144)
    e = lib().sdxo_stress_damping_element(_d(X), _d(u), _d(us), kappa, mu, lam, int(geo), int(project),
                                          r.ctypes.data_as(_dp), H.ctypes.data_as(_dp))
    return e, r, H.reshape(12, 12)


def pcg_csr(indptr, indices, data, b, x0=None, bs=3, params=None, fp64=True):
    n = len(b)
    kp = params or _abi.default_krylov_params()
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it, conv, rn = C.c_int32(), C.c_int32(), C.c_double()
    data = np.ascontiguousarray(data, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    lib().sdxo_pcg_csr(int(fp64), n, _i(indptr), _i(indices), data.ctypes.data_as(_dp),
                       b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), bs, C.byref(kp),
                       C.byref(it), C.byref(conv), C.byref(rn))
    return x, it.value, conv.value, rn.value


class OracleScene:
    """The oracle's Scene: soft tet actors, pinned nodes, plane/sphere colliders, Step."""

    def __init__(self, fp64: bool = False):
        self._h = lib().sdxo_scene_create(int(fp64))
        self.fp64 = fp64
        self.n_verts = []
        self._keep = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().sdxo_scene_destroy(h)
            self._h = None

    def add_soft_tet_mesh(self, X, tets, params: _abi.SoftParams) -> int:
        X = np.ascontiguousarray(X, dtype=np.float64)
        tets = np.ascontiguousarray(tets, dtype=np.int32)
        a = lib().sdxo_add_soft_tet_mesh(self._h, _d(X), X.shape[0], _i(tets), tets.shape[0],
                                         C.byref(params))
        self.n_verts.append(X.shape[0])
        return a

    def set_element_materials(self, actor, E, nu):
        E = np.ascontiguousarray(E, dtype=np.float64)
        nu = np.ascontiguousarray(nu, dtype=np.float64)
        rc = lib().sdxo_set_element_materials(self._h, actor, _d(E), _d(nu), len(E))
        assert rc == 0, rc

    def pin_nodes(self, actor, nodes):
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        lib().sdxo_pin_nodes(self._h, actor, _i(nodes), len(nodes))

    def add_plane(self, p, n, actor=-1) -> int:
        return lib().sdxo_add_collider(self._h, 0, _d(np.array(p, float)), _d(np.array(n, float)), 0.0, actor)

    def add_sphere(self, c, r, actor=-1) -> int:
        return lib().sdxo_add_collider(self._h, 1, _d(np.array(c, float)), None, float(r), actor)

    def set_collider_position(self, cid, p):
        lib().sdxo_set_collider_position(self._h, cid, _d(np.array(p, float)))

    def step(self, dt, params: _abi.SolverParams):
        stats = (_abi.StepStats * len(self.n_verts))()
        lib().sdxo_step(self._h, dt, C.byref(params), stats, len(self.n_verts))
        return [s.as_dict() for s in stats]

    def displacements(self, actor):
        out = np.zeros(3 * self.n_verts[actor])
        lib().sdxo_get_state(self._h, actor, 0, out.ctypes.data_as(_dp))
        return out.reshape(-1, 3)

    def velocities(self, actor):
        out = np.zeros(3 * self.n_verts[actor])
        lib().sdxo_get_state(self._h, actor, 1, out.ctypes.data_as(_dp))
        return out.reshape(-1, 3)

    def set_displacements(self, actor, u):
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        lib().sdxo_set_state(self._h, actor, 0, u.ctypes.data_as(_dp))

    def set_velocities(self, actor, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        lib().sdxo_set_state(self._h, actor, 1, v.ctypes.data_as(_dp))

    def info(self, actor):
        out = (C.c_int64 * 4)()
        lib().sdxo_actor_info(self._h, actor, out)
        return {"n_verts": out[0], "n_tets": out[1], "nnzb": out[2], "n_tris": out[3]}

    def assemble(self, actor, dt, params, u0=None, v0=None):
        inf = self.info(actor)
        nV, nnzb = inf["n_verts"], inf["nnzb"]
        res = np.zeros(3 * nV)
        ptr = np.zeros(nV + 1, dtype=np.int32)
        col = np.zeros(nnzb, dtype=np.int32)
        val = np.zeros(9 * nnzb)
        lib().sdxo_assemble(self._h, actor, dt, C.byref(params),
                            None if u0 is None else _d(u0), None if v0 is None else _d(v0),
                            res.ctypes.data_as(_dp), ptr.ctypes.data_as(_ip), col.ctypes.data_as(_ip),
                            val.ctypes.data_as(_dp))
        return res, ptr, col, val.reshape(nnzb, 3, 3)

    def set_contact_query(self, on=True):
        lib().sdxo_set_contact_query(self._h, int(on))

    def contact_wrench(self, actor, collider):
        out = np.zeros(7)
        rc = lib().sdxo_get_contact_wrench(self._h, actor, collider, out.ctypes.data_as(_dp))
        assert rc == 0, rc
        return out

    def set_boundary_conditions(self, actor, dofs, values, clear=True):
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        values = np.ascontiguousarray(values, dtype=np.float64)
        lib().sdxo_set_boundary_conditions(self._h, actor, dofs.ctypes.data_as(_ip), values.ctypes.data_as(_dp),
                                           dofs.size, int(clear))

    def objective(self, actor, dt, params, u0=None, v0=None) -> float:
        out = C.c_double()
        lib().sdxo_objective(self._h, actor, dt, C.byref(params), None if u0 is None else _d(u0),
                             None if v0 is None else _d(v0), C.byref(out))
        return out.value

    def boundary(self, actor):
        n = self.info(actor)["n_tris"]
        tris = np.zeros((n, 3), dtype=np.int32)
        area = np.zeros(n)
        lib().sdxo_boundary(self._h, actor, tris.ctypes.data_as(_ip), area.ctypes.data_as(_dp))
        return tris, area


def build_scene(spec: SceneSpec, fp64: bool = False) -> tuple[OracleScene, list]:
    """Create the oracle scene of a SceneSpec; returns (scene, collider ids)."""
    sc = OracleScene(fp64=fp64)
    for b in spec.bodies:
        a = sc.add_soft_tet_mesh(b.X, b.tets, b.soft_params(_abi))
        if np.ndim(b.E):
            sc.set_element_materials(a, b.E, b.nu)
        if len(b.pinned):
            sc.pin_nodes(a, b.pinned)
    ids = []
    for c in spec.colliders:
        if c.kind == "plane":
            ids.append(sc.add_plane(c.point, c.normal, c.actor))
        else:
            ids.append(sc.add_sphere(c.point, c.radius, c.actor))
    return sc, ids


def run_scene(spec: SceneSpec, params: _abi.SolverParams, steps: int | None = None, fp64: bool = False):
    """Step the oracle through `steps` steps of `spec`; returns (scene, per-step stats)."""
    sc, ids = build_scene(spec, fp64)
    all_stats = []
    for s in range(spec.steps if steps is None else steps):
        for cid, c in zip(ids, spec.colliders):
            sc.set_collider_position(cid, c.position(s, spec.dt))
        all_stats.append(sc.step(spec.dt, params))
    return sc, all_stats
