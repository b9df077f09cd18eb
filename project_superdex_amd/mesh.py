"""Seeded tetrahedral meshes and scenes for the soft-body FEM step.

The reference builds its soft actors from a tet mesh handed to
``Context::CreateTetMeshShape(coords, connectivity)``
(superdex_physics/libraries/mochi/mochi_physics/include/mochi_physics/cpp_api/mochi_context.h:142)
and pins nodes with ``Actor::AddBoundaryConditionNodesWorldPermanent``
(.../cpp_api/mochi_actor.h:202).  This module produces those inputs: a regular
hex grid split 6 tets per hex (Kuhn / Freudenthal split, every tet of a hex
shares the hex diagonal (0,0,0)->(1,1,1)), the node list of the bottom face,
and the scenes (material, colliders, motion) used by the tests and the bench.

Everything here is set-up code that runs once per topology on the host.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# Y is up, gravity points down -Y (the reference's default scene gravity).
GRAVITY = (0.0, -9.8, 0.0)


def hex_grid_tets(n: int, h: float, origin=(0.0, 0.0, 0.0)):
    """Return (X, tets) for an n x n x n hex grid of spacing h split 6 tets/hex.

    Node (i, j, k) sits at origin + h*(i, j, k) and has index
    i + (n+1)*(j + (n+1)*k).  Each hex is cut into the 6 tets
    [c000, c000+e_a, c000+e_a+e_b, c111] over the permutations (a, b, c) of
    the axes; odd permutations swap two nodes so every tet has positive
    signed volume h^3/6.
    """
    if n < 1:
        raise ValueError("grid needs at least one hex per side")
    m = n + 1
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    # node index i + m*(j + m*k): build with k slowest
    idx = lambda i, j, k: i + m * (j + m * k)
    X = np.zeros((m * m * m, 3), dtype=np.float64)
    X[idx(ii, jj, kk).ravel(), 0] = (origin[0] + h * ii).ravel()
    X[idx(ii, jj, kk).ravel(), 1] = (origin[1] + h * jj).ravel()
    X[idx(ii, jj, kk).ravel(), 2] = (origin[2] + h * kk).ravel()

    unit = np.eye(3, dtype=np.int64)
    tet_offsets = []
    for perm in itertools.permutations(range(3)):
        a, b, _ = perm
        o0 = np.zeros(3, dtype=np.int64)
        o1 = unit[a]
        o2 = unit[a] + unit[b]
        o3 = np.ones(3, dtype=np.int64)
        # sign of the permutation decides the orientation of (o1, o2, o3)
        det = np.linalg.det(np.stack([o1, o2, o3]).astype(np.float64))
        if det < 0:
            o1, o2 = o2, o1
        tet_offsets.append((o0, o1, o2, o3))

    hi, hj, hk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    hi, hj, hk = hi.ravel(), hj.ravel(), hk.ravel()
    # hexes ordered with i fastest, k slowest to keep neighbouring tets close in memory
    order = np.lexsort((hi, hj, hk))
    hi, hj, hk = hi[order], hj[order], hk[order]
    tets = np.empty((hi.size, 6, 4), dtype=np.int32)
    for t, offs in enumerate(tet_offsets):
        for c, o in enumerate(offs):
            tets[:, t, c] = idx(hi + o[0], hj + o[1], hk + o[2])
    return X, tets.reshape(-1, 4)


def tet_volumes(X: np.ndarray, tets: np.ndarray) -> np.ndarray:
    e1 = X[tets[:, 1]] - X[tets[:, 0]]
    e2 = X[tets[:, 2]] - X[tets[:, 0]]
    e3 = X[tets[:, 3]] - X[tets[:, 0]]
    return np.einsum("ij,ij->i", e1, np.cross(e2, e3)) / 6.0


def nodes_on_plane(X: np.ndarray, axis: int, value: float, tol: float = 1e-9) -> np.ndarray:
    return np.nonzero(np.abs(X[:, axis] - value) <= tol)[0].astype(np.int32)


def lame(E: float, nu: float):
    """Lame constants (lambda, mu) from Young's modulus and Poisson ratio."""
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


@dataclass
class Collider:
    """An analytic rigid collider that moves kinematically.

    kind 'plane': point p on the plane and normal n pointing to the free side.
    kind 'sphere': centre c and radius r.
    position(step) gives p (plane) or c (sphere) at the end of step `step`.
    """

    kind: str
    point: tuple
    normal: tuple = (0.0, 1.0, 0.0)
    radius: float = 0.0
    velocity: tuple = (0.0, 0.0, 0.0)
    ramp_steps: int = 0  # >0: move with `velocity` for ramp_steps steps then hold
    actor: int = -1  # -1: collides with every actor

    def position(self, step: int, dt: float):
        s = step + 1
        if self.ramp_steps > 0:
            s = min(s, self.ramp_steps)
        return tuple(self.point[i] + self.velocity[i] * dt * s for i in range(3))


class ColliderPaths:
    """Collider.position of a list of colliders, all at once: at(step, dt) is
    the (n, 3) array of their positions."""

    def __init__(self, colliders):
        self.P = np.array([c.point for c in colliders], dtype=np.float64).reshape(-1, 3)
        self.V = np.array([c.velocity for c in colliders], dtype=np.float64).reshape(-1, 3)
        self.ramp = np.array([c.ramp_steps for c in colliders], dtype=np.int64)

    def at(self, step: int, dt: float) -> np.ndarray:
        if not self.ramp.any():
            return self.P + self.V * (dt * (step + 1))
        s = np.full(len(self.ramp), step + 1)
        s = np.where(self.ramp > 0, np.minimum(s, self.ramp), s)
        return self.P + self.V * (dt * s)[:, None]


@dataclass
class SoftBody:
    X: np.ndarray
    tets: np.ndarray
    E: float | np.ndarray = 1e5
    nu: float | np.ndarray = 0.4
    density: float = 1000.0
    pinned: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    penalty: float = 1e6
    mass_damping: float = 0.0
    coulomb: float = 0.0
    viscous: float = 0.0
    falloff_vel: float = 0.01
    normal_damping: float = 0.0
    stiffness_damping: float = 0.0
    stiffness_damping_geometric: int = 0

    def soft_params(self, abi):
        """The sdx_soft_params of this body (homogeneous part; per-element fields apart)."""
        E0 = float(np.mean(self.E)) if np.ndim(self.E) else float(self.E)
        nu0 = float(np.mean(self.nu)) if np.ndim(self.nu) else float(self.nu)
        return abi.default_soft_params(E0, nu0, self.density, self.penalty, self.mass_damping, self.coulomb,
                                       self.viscous, self.falloff_vel, self.normal_damping,
                                       self.stiffness_damping, self.stiffness_damping_geometric)


@dataclass
class SceneSpec:
    bodies: list
    colliders: list
    dt: float = 1e-3
    steps: int = 100
    newton_iters: int = 1
    name: str = ""

    @property
    def num_tets(self) -> int:
        return int(sum(b.tets.shape[0] for b in self.bodies))

    @property
    def num_verts(self) -> int:
        return int(sum(b.X.shape[0] for b in self.bodies))

    def subset(self, envs) -> "SceneSpec":
        """The scene restricted to the soft bodies `envs` (with their own
        colliders and the global ones), actors renumbered in that order."""
        envs = list(envs)
        remap = {e: i for i, e in enumerate(envs)}
        cols = []
        for c in self.colliders:
            if c.actor == -1:
                cols.append(c)
            elif c.actor in remap:
                cols.append(Collider(kind=c.kind, point=c.point, normal=c.normal, radius=c.radius,
                                     velocity=c.velocity, ramp_steps=c.ramp_steps, actor=remap[c.actor]))
        return SceneSpec(bodies=[self.bodies[e] for e in envs], colliders=cols, dt=self.dt, steps=self.steps,
                         newton_iters=self.newton_iters, name=f"{self.name}[{len(envs)}]")


def soft_cube(n: int, edge: float, E=1e5, nu=0.4, density=1000.0, penalty=1e6, origin=(0.0, 0.0, 0.0)):
    X, tets = hex_grid_tets(n, edge / n, origin)
    pinned = nodes_on_plane(X, 1, origin[1])
    return SoftBody(X=X, tets=tets, E=E, nu=nu, density=density, pinned=pinned, penalty=penalty)


def cube_drop_scene(n: int = 10, edge: float = 0.1, steps: int = 100) -> SceneSpec:
    """One soft cube, bottom face pinned, gravity only (the 10^3-hex CPU case)."""
    return SceneSpec(bodies=[soft_cube(n, edge)], colliders=[], steps=steps, name=f"cube{n}")


def plate_press_scene(n: int = 40, edge: float = 0.1, steps: int = 50, press: float = 2e-3,
                      ramp_steps: int = 20, dt: float = 1e-3) -> SceneSpec:
    """Soft cube under a rigid plate that moves down `press` metres over `ramp_steps` steps.

    The plate is a plane with normal -Y whose point starts on the top face
    (y = edge) and is then held at y = edge - press.
    """
    body = soft_cube(n, edge)
    vy = -press / (ramp_steps * dt)
    plate = Collider(kind="plane", point=(0.0, edge, 0.0), normal=(0.0, -1.0, 0.0),
                     velocity=(0.0, vy, 0.0), ramp_steps=ramp_steps, actor=0)
    return SceneSpec(bodies=[body], colliders=[plate], dt=dt, steps=steps, name=f"press{n}")


def batched_press_scene(num_envs: int, seed: int, n_min: int, n_max: int, h: float = 0.01, steps: int = 50,
                        press=(1e-3, 3e-3), ramp=(10, 30), newton_iters: int = 1, dt: float = 1e-3) -> SceneSpec:
    """Many independent plate presses of uneven size (the plate-press case, batched).

    Per env: a cube of n ~ U{n_min..n_max} hexes per side of spacing h, E
    log-uniform in [1e4, 1e6] Pa, nu ~ U(0.3, 0.45), bottom face pinned, under
    a rigid plate (plane, normal -Y) that starts on the top face and moves
    down a depth ~ U(press) over ~ U{ramp} steps, then holds.
    """
    rng = np.random.default_rng(seed)
    bodies, colliders = [], []
    for e in range(num_envs):
        n = int(rng.integers(n_min, n_max + 1))
        E = float(10.0 ** rng.uniform(4.0, 6.0))
        nu = float(rng.uniform(0.3, 0.45))
        edge = n * h
        origin = (e * 0.5, 0.0, 0.0)
        bodies.append(soft_cube(n, edge, E=E, nu=nu, origin=origin))
        depth = float(rng.uniform(*press))
        ramp_steps = int(rng.integers(ramp[0], ramp[1] + 1))
        colliders.append(Collider(kind="plane", point=(origin[0], edge, 0.0), normal=(0.0, -1.0, 0.0),
                                  velocity=(0.0, -depth / (ramp_steps * dt), 0.0), ramp_steps=ramp_steps, actor=e))
    return SceneSpec(bodies=bodies, colliders=colliders, dt=dt, steps=steps, newton_iters=newton_iters,
                     name=f"press_batch{num_envs}")


def batched_indent_scene(num_envs: int, seed: int, n_min: int, n_max: int, h: float = 0.01,
                         radius: float = 0.02, speed: float = 0.05, steps: int = 20,
                         indenters=(1, 1), newton_iters: int = 1, dt: float = 1e-3) -> SceneSpec:
    """Many independent soft cubes of uneven size, each pressed by sphere indenters.

    Per env: n ~ U{n_min..n_max} hexes per side of spacing h, E log-uniform
    in [1e4, 1e6] Pa, nu ~ U(0.3, 0.45), bottom face pinned.  Each indenter
    is a sphere of radius `radius` that starts touching the top face at a
    random point and moves with a random velocity of magnitude <= speed
    (downward component always present).
    """
    rng = np.random.default_rng(seed)
    bodies, colliders = [], []
    for e in range(num_envs):
        n = int(rng.integers(n_min, n_max + 1))
        E = float(10.0 ** rng.uniform(4.0, 6.0))
        nu = float(rng.uniform(0.3, 0.45))
        edge = n * h
        # envs are laid out on a line so that their world boxes never overlap
        origin = (e * 0.5, 0.0, 0.0)
        body = soft_cube(n, edge, E=E, nu=nu, origin=origin)
        bodies.append(body)
        k = int(rng.integers(indenters[0], indenters[1] + 1))
        for _ in range(k):
            px = origin[0] + rng.uniform(0.2, 0.8) * edge
            pz = origin[2] + rng.uniform(0.2, 0.8) * edge
            if indenters[1] > 1:
                d = rng.normal(size=3)
                d[1] = -abs(d[1]) - 0.5
                d = d / np.linalg.norm(d) * rng.uniform(0.3, 1.0) * speed
                vel = tuple(float(v) for v in d)
            else:
                vel = (0.0, -speed, 0.0)
            colliders.append(Collider(kind="sphere", point=(px, edge + radius, pz), radius=radius,
                                      velocity=vel, actor=e))
    return SceneSpec(bodies=bodies, colliders=colliders, dt=dt, steps=steps,
                     newton_iters=newton_iters, name=f"batch{num_envs}")
