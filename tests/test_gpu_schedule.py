"""GPU: CG convergence across residual refreshes, launch schedules, graph replay and the autotuner.

- mixed precision on a press large enough to pass several FP64 residual
  refreshes: every solve converges, its FP64 true residual |b - A x| / |b| is
  at or below 1e-6, iteration counts stay within 15% of the oracle's
  (standard PCG) and positions match the oracle as in the parity tests;
- every launch schedule (matvec lanes, element workgroup, burst length,
  graph replay, persistent kernels over work lists of active tiles, the
  first iterations run environment group by group, tail bursts as one
  cooperative launch, the whole solve as one launch) steps to the same state: displacements within 1e-3 of the
  largest displacement (FP32 solves at rel. residual 1e-6 differ at the
  1e-4 level when only the summation order changes; a wrong result is off
  by O(1)), iteration counts within 2;
- graph replays are checked step after step against eager launches, with
  host work and synchronisations between the replays;
- the autotuner leaves the state bit-for-bit as it found it.
"""
import numpy as np
import pytest

from project_superdex_amd import _abi, mesh
from project_superdex_amd import scene as gs

from test_gpu_parity import compare, run_gpu

pytestmark = pytest.mark.gpu

MIXED = _abi.PREC_MODE_MIXED


def mixed_params(newton=1):
    return _abi.pinned_solver_params(newton, precision_mode=MIXED, refresh_every=10)


def test_press_mixed_converges_through_refreshes(oracle):
    spec = mesh.plate_press_scene(n=16, steps=14, ramp_steps=8)
    p = mixed_params()
    sg, stg = run_gpu(spec, p, 14, precision=MIXED)
    so, sto = oracle.run_scene(spec, _abi.pinned_solver_params(1), 14, fp64=False)
    compare(spec, sg, so, 2e-3)
    for a, b in zip(stg, sto):
        a, b = a[0], b[0]
        assert a["linear_convergence"] == _abi.CONV_CONVERGED, a
        assert a["linear_true_rel_res"] <= 1e-6, a
        assert a["linear_iters"] >= 20  # several refreshes happened
        assert abs(a["linear_iters"] - b["linear_iters"]) <= max(3, 0.15 * b["linear_iters"]), (a, b)


def batched_spec(steps=5):
    return mesh.batched_indent_scene(num_envs=6, seed=11, n_min=3, n_max=7, steps=steps, speed=0.2)


def state(sc):
    return [(a.displacements(), a.velocities()) for a in sc.actors]


@pytest.mark.parametrize("sched", [
    # the CG as bursts of two kernels per iteration (env_cg 0)
    dict(env_cg=0, use_graphs=0),
    dict(env_cg=0, spmv_lanes=4),
    dict(env_cg=0, spmv_lanes=8, cg_burst=3),
    dict(env_cg=0, element_block=64, cg_burst=32),
    dict(env_cg=0, element_block=256, use_graphs=0, cg_burst=1),
    dict(env_cg=0, active_lists=0),
    dict(env_cg=0, active_lists=0, use_graphs=0, spmv_lanes=8),
    dict(env_cg=0, cg_groups=3, group_iters=8, cg_burst=4),
    dict(env_cg=0, cg_groups=6, group_iters=5, cg_burst=2, use_graphs=0),  # 3 bursts per group: list parity flips
    dict(env_cg=0, cg_groups=4, group_iters=64, cg_burst=8, spmv_lanes=8),  # every iteration grouped
    dict(env_cg=0, fused_tail=1),  # tail bursts as one cooperative launch each
    dict(env_cg=0, fused_tail=1, cg_burst=16, spmv_lanes=16),
    dict(fused_asm=0, asm_groups=1),  # the record path: every environment's records kept
    dict(fused_asm=0, asm_groups=2),  # assembly groups share the record buffers
    dict(fused_asm=0, asm_groups=6, use_graphs=0),  # one environment per assembly group
    dict(env_cg=0, fused_asm=0, asm_groups=4, cg_groups=3, group_iters=8, cg_burst=4),
    dict(fused_asm=1, asm_groups=3),  # fused assembly (groups apply to residual-only assemblies)
    # the whole solve as one launch, one workgroup per environment at a time
    dict(env_cg=1, spmv_lanes=4),
    dict(env_cg=2, spmv_lanes=8, cg_burst=3),
    dict(env_cg=1, spmv_lanes=16, cg_groups=3, group_iters=8),  # burst fields do not apply
])
def test_schedules_step_to_the_same_state(sched):
    # one Newton iteration per step: a second one starts from a residual at
    # the FP32 noise floor, where the line search's accept/reject flips with
    # rounding (as in any FP32 build)
    spec = batched_spec()
    p = mixed_params(1)
    base, st0 = run_gpu(spec, p, 5, precision=MIXED)
    sc, rigid = gs.build_scene(spec, precision=MIXED)
    sc.set_schedule(**sched)
    assert {k: sc.schedule()[k] for k in sched} == sched
    # groups of about equal size: at most the requested number
    g = sc.info()["cg_groups"]
    assert g == 1 if sched.get("cg_groups", 1) == 1 else 1 < g <= sched["cg_groups"]
    st1 = []
    for s in range(5):
        for r, c in zip(rigid, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        st1.append(sc.step(spec.dt, p))
        # host work and a synchronisation between replays
        _ = sc.actors[0].displacements()
    for a, b in zip(base.actors, sc.actors):
        ua, ub = a.displacements(), b.displacements()
        assert np.abs(ua - ub).max() <= 1e-3 * np.abs(ua).max()
    for sa, sb in zip(st0, st1):
        for ea, eb in zip(sa, sb):
            assert abs(ea["linear_iters"] - eb["linear_iters"]) <= 2
            assert eb["linear_convergence"] == _abi.CONV_CONVERGED


def test_graph_replays_match_eager_step_by_step():
    spec = batched_spec(steps=8)
    p = mixed_params(1)
    scenes = []
    for graphs in (0, 1):
        sc, rigid = gs.build_scene(spec, precision=MIXED)
        sc.set_schedule(env_cg=0, use_graphs=graphs, cg_burst=4)
        scenes.append((sc, rigid))
    for s in range(8):
        snaps = []
        for sc, rigid in scenes:
            for r, c in zip(rigid, spec.colliders):
                r.set_position(c.position(s, spec.dt))
            sc.step(spec.dt, p)
            snaps.append(np.concatenate([a.displacements().ravel() for a in sc.actors]))
        assert np.abs(snaps[0] - snaps[1]).max() <= 1e-3 * np.abs(snaps[0]).max(), s


def test_autotune_restores_state_and_picks_a_valid_schedule():
    spec = batched_spec(steps=4)
    p = mixed_params(1)
    sc, rigid = gs.build_scene(spec, precision=MIXED)
    for s in range(2):
        for r, c in zip(rigid, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        sc.step(spec.dt, p)
    before = state(sc)
    sched, ms = sc.autotune(spec.dt, p, trials=1)
    after = state(sc)
    for (u0, v0), (u1, v1) in zip(before, after):
        np.testing.assert_array_equal(u0, u1)
        np.testing.assert_array_equal(v0, v1)
    assert ms > 0
    assert sched["spmv_lanes"] in (4, 8, 16) and sched["element_block"] in (64, 128, 256)
    assert sched["env_cg"] in (0, 1, 2)
    assert sched == sc.schedule()
    # stepping on with the tuned schedule matches a scene that never tuned
    ref, rrig = gs.build_scene(spec, precision=MIXED)
    for s in range(4):
        for r, c in zip(rrig, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        ref.step(spec.dt, p)
    for s in range(2, 4):
        for r, c in zip(rigid, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        sc.step(spec.dt, p)
    for a, b in zip(ref.actors, sc.actors):
        assert np.abs(a.displacements() - b.displacements()).max() <= 1e-3 * np.abs(a.displacements()).max()


def test_autotune_on_first_step_via_params():
    spec = batched_spec(steps=2)
    p = mixed_params(1)
    p.schedule = -1
    sg, st = run_gpu(spec, p, 2, precision=MIXED)
    q = mixed_params(1)
    ref, _ = run_gpu(spec, q, 2, precision=MIXED)
    for a, b in zip(ref.actors, sg.actors):
        assert np.abs(a.displacements() - b.displacements()).max() <= 1e-3 * np.abs(a.displacements()).max()


def test_batched_collider_positions_match_one_by_one():
    """sdx_set_rigid_positions moves every collider of a step in one call and
    steps to the state of one sdx_set_rigid_position call per collider; a
    batch with a bad index is rejected before any collider moves."""
    spec = batched_spec(steps=4)
    p = mixed_params()
    s1, r1 = gs.build_scene(spec, precision=MIXED)
    s2, r2 = gs.build_scene(spec, precision=MIXED)
    paths = mesh.ColliderPaths(spec.colliders)
    ids = np.array([r.index for r in r2], dtype=np.int32)
    for s in range(spec.steps):
        for r, c in zip(r1, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        s2.set_rigid_positions(ids, paths.at(s, spec.dt))
        a = s1.step(spec.dt, p)
        b = s2.step(spec.dt, p, as_array=True)
        for ea, eb in zip(a, b):
            assert abs(ea["linear_iters"] - int(eb["linear_iters"])) <= 2
    for (ua, _), (ub, _) in zip(state(s1), state(s2)):
        assert np.abs(ua - ub).max() <= 1e-3 * np.abs(ua).max()
    # a batch naming a collider that does not exist moves nothing
    before = paths.at(spec.steps - 1, spec.dt)
    with pytest.raises(RuntimeError):
        s2.set_rigid_positions(np.array([ids[0], 10_000], dtype=np.int32), before[:2] + 1.0)
    s1.step(spec.dt, p)
    s2.step(spec.dt, p)
    for (ua, _), (ub, _) in zip(state(s1), state(s2)):
        assert np.abs(ua - ub).max() <= 1e-3 * np.abs(ua).max()
