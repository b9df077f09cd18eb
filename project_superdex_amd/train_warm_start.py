"""Train the learned CG warm start on the device.

Data: (feature row, Newton step / scale) pairs collected from the plain
solver (zero initial guess) on the batched indentation generator: by default
8 seeds x 512 environments x 50 steps = 204,800 linear solves, 16 vertices
sampled per environment and solve.  Target: the Newton step / s, or with
--target change its change since the previous step (dx / s minus the feature
columns 9..11, dprev / s); the warm start combines the network's guess with
the previous step itself (least residual over the two), so either target
spans dprev + change.  Training: Adam on the device (FP32 forward / backward
kernels).  The held-out seed reports the prediction error and the CG
iterations from zero, from the previous step alone (a network that outputs
zero) and with the network.  Measured (held-out seed, 128 environments x 20
steps): zero 83,948, previous step alone 70,809, with the network 71,526
(step target) or 71,568 (change target): the gain comes from the previous
step; the network's guesses add none.  The history-only start
(SDX_WARM_START_HISTORY) is that combination without the network: on the
4096-environment workload it runs 3.8% faster than the learned start (the
inference it skips) at the same CG iteration count.  With --generator press
(8 seeds x 512 batched plate presses x 50 steps = 204,800 solves; weights in
weights/warm_start_mlp_press.npz) the held-out presses (128 environments x 40
steps) took 158,397 CG iterations from zero, 119,323 from the previous steps
alone and 119,473 with the network (step target; 119,513 with the change
target): again the previous steps carry the whole gain.

    python -m project_superdex_amd.train_warm_start [--generator indent|press] [--target step|change]
                                                    [--out weights/warm_start_mlp.npz]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import _abi, mesh
from . import scene as gs
from .learned import DEFAULT_WEIGHTS, WarmStartModel


def generator(seed: int, envs: int, steps: int, kind: str = "indent"):
    """The batched indentation case (sphere indenters), or the plate-press case batched."""
    if kind == "press":
        return mesh.batched_press_scene(envs, seed, 6, 16, h=0.01, steps=steps)
    return mesh.batched_indent_scene(envs, seed, 6, 16, h=0.01, radius=0.02, speed=0.05, steps=steps,
                                     indenters=(1, 1), newton_iters=1)


def params(learned: bool = False):
    p = _abi.pinned_solver_params(1, lin_rel_tol=1e-6, precision_mode=_abi.PREC_MODE_MIXED, refresh_every=10)
    p.warm_start = _abi.WARM_START_LEARNED if learned else _abi.WARM_START_ZERO
    return p


def run(spec, p, model=None, collect=None):
    sc, rigid = gs.build_scene(spec, precision=_abi.PREC_MODE_MIXED)
    if model is not None:
        sc.set_warm_start_model(model)
    if collect:
        sc.collect(*collect)
    stats = []
    paths = mesh.ColliderPaths(spec.colliders)
    ids = np.array([r.index for r in rigid], dtype=np.int32)
    for s in range(spec.steps):
        if len(ids):
            sc.set_rigid_positions(ids, paths.at(s, spec.dt))
        stats.append(sc.step(spec.dt, p))
    return sc, stats


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=8, help="training seeds 1..N")
    ap.add_argument("--envs", type=int, default=512)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--per-env", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--holdout-seed", type=int, default=1000)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--target", choices=["change", "step"], default="step",
                    help="learn the change of the Newton step since the previous one, or the step itself")
    ap.add_argument("--generator", choices=["indent", "press"], default="indent",
                    help="training and held-out scenes: batched sphere indentation or batched plate presses")
    args = ap.parse_args(argv)

    t0 = time.time()
    # every training seed collects into its own scene; the model trains on the
    # concatenation, copied once through the host
    X, T = [], []
    solves = 0
    for seed in range(1, args.seeds + 1):
        spec = generator(seed, args.envs, args.steps, args.generator)
        sc, _ = run(spec, params(), collect=(args.envs * args.steps * args.per_env, args.per_env))
        x, t = sc.collected()
        X.append(x)
        T.append(t)
        solves += args.envs * args.steps
        sc.close()
        print(f"[train] seed {seed}: {x.shape[0]} rows", file=sys.stderr, flush=True)
    X, T = np.concatenate(X), np.concatenate(T)
    if args.target == "change":
        T = np.ascontiguousarray(T - X[:, 9:12])
    t_data = time.time() - t0
    model = WarmStartModel(seed=0)
    t1 = time.time()
    losses = model.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_train = time.time() - t1
    model.save(args.out)
    print(f"[train] {solves} solves, {X.shape[0]} rows, loss {losses[0]:.4f} -> {losses[-1]:.4f}",
          file=sys.stderr, flush=True)

    # held-out seed: prediction error and CG iterations with / without the warm start
    spec = generator(args.holdout_seed, min(args.envs, 128), 20 if args.generator == "indent" else 40,
                     args.generator)
    s0, st0 = run(spec, params(), collect=(spec.steps * len(spec.bodies) * 64, 64))
    Xh, Th = s0.collected()
    if args.target == "change":
        Th = Th - Xh[:, 9:12]
    err = float(np.linalg.norm(model.forward(Xh) - Th) / np.linalg.norm(Th))
    s1, st1 = run(spec, params(True), model=model)
    null = WarmStartModel(seed=0)
    null.params = np.zeros_like(null.params)
    s2, st2 = run(spec, params(True), model=null)
    it0 = sum(e["linear_iters"] for st in st0 for e in st)
    it1 = sum(e["linear_iters"] for st in st1 for e in st)
    it2 = sum(e["linear_iters"] for st in st2 for e in st)
    kept = sum(e["warm_start_accepted"] for st in st1 for e in st)
    tried = sum(e["warm_start_tried"] for st in st1 for e in st)
    print(json.dumps({"solves": solves, "rows": int(X.shape[0]), "epochs": args.epochs, "loss_first": losses[0],
                      "loss_last": losses[-1], "data_seconds": t_data, "train_seconds": t_train,
                      "target": args.target, "generator": args.generator, "holdout_rel_error": err, "holdout_cg_iters_zero": it0,
                      "holdout_cg_iters_previous_step": it2, "holdout_cg_iters_learned": it1,
                      "holdout_kept": kept, "holdout_tried": tried, "weights": os.path.relpath(args.out)}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
