"""Throughput of the implicit soft-body FEM step on one MI355X.

Workload (default): one soft cube of 40^3 hexes split 6 tets/hex (384,000
tets, 68,921 vertices), Neo-Hookean E=1e5 Pa, nu=0.4, rho=1000, bottom face
pinned, gravity -Y, rigid plate pressing 2 mm down over 20 steps (penalty
k=1e6 Pa/m); backward Euler dt=1e-3, 1 Newton iteration per step,
block-Jacobi PCG to relative residual 1e-6 (ResidualL2) with FP32 matvecs,
FP64 dot products and an FP64 residual refresh every 10 iterations.

A "step" is one Scene::Step of that workload; everything it needs is
resident in HBM before the timed region.  The JSON line reports
tet-steps/s, the roofline of the dominant kernel (the CG matvec, timed with
HIP events on the world's stream during the timed region) and the CPU
oracle timed on a bounded sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "tet-steps/sec at rel. residual 1e-6, learned+batched vs plain PCG, 1 GPU"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cg_spmv_press40_traffic.txt")


def matvec_bytes(nB: int, nV: int, s: int = 4) -> int:
    """Algorithmic HBM bytes of one CG matvec launch (k_cg_spmv, w = A z):
    block values and column indices once ((9s+4) B), row pointers (4(V+1)),
    z read once (3sV), w written once (3sV)."""
    return (9 * s + 4) * nB + 4 * (nV + 1) + 6 * s * nV


def step_bytes(nT, nB, nV, cg_iters, ls_trials, s=4):
    """Reference-equivalent algorithmic bytes of one step with 1 Newton
    iteration: full assembly + k_CG CG iterations + (1 + L + e) residual
    assemblies + final residual (e = 1, explosion control on)."""
    bytes_cg = (9 * s + 4) * nB + 4 * (nV + 1) + 9 * s * nV + 30 * s * nV
    bytes_asm = nT * (16 + 15 * s) + 9 * s * nB + 6 * s * nV
    bytes_res = nT * (16 + 15 * s) + 6 * s * nV
    return bytes_asm + cg_iters * bytes_cg + (1 + ls_trials + 1) * bytes_res + bytes_res


def read_traffic():
    if not os.path.exists(TRAFFIC_FILE):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(TRAFFIC_FILE).read())
    return float(m.group(1)) if m else None


def cpu_baseline(spec, params, seconds_budget: float = 20.0):
    """The oracle (C++ restatement, OpenMP) on the same workload, bounded to
    about `seconds_budget` of stepping; returns a cpu_baseline object."""
    from oracle import oracle_api

    threads = min(16, os.cpu_count() or 1)
    oracle_api.set_threads(threads)
    sc, ids = oracle_api.build_scene(spec, fp64=False)
    steps, t_used = 0, 0.0
    while t_used < seconds_budget and steps < spec.steps:
        for cid, c in zip(ids, spec.colliders):
            sc.set_collider_position(cid, c.position(steps, spec.dt))
        t0 = time.perf_counter()
        sc.step(spec.dt, params)
        t_used += time.perf_counter() - t0
        steps += 1
    rate = spec.num_tets * steps / t_used
    return {"value": rate, "unit": "tet-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {steps} steps of the same workload (FP32 oracle, {threads} OpenMP threads), "
                      f"{t_used:.1f} s of stepping"}


def copy_probe(mib: int) -> int:
    """Three launches of one elementwise kernel that reads `mib` MiB and writes
    `mib` MiB: the known-size reference for FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4}))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=40, help="hexes per cube side")
    ap.add_argument("--precision", choices=["fp32", "mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--burst", type=int, default=0, help="CG iterations between host checks (0: default)")
    ap.add_argument("--time-every", type=int, default=16,
                    help="time every N-th CG matvec launch with HIP events (0: no kernel timing)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 elementwise copies of this many MiB (read + write), the "
                         "calibration kernel for the HBM byte counters, and exit")
    args = ap.parse_args(argv)

    if args.copy_probe_mib:
        return copy_probe(args.copy_probe_mib)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every GPU box here has one GPU; run with --gpus 1",
              file=sys.stderr)
        return 2

    from project_superdex_amd import _abi, mesh
    from project_superdex_amd import scene as gs

    mode = {"fp32": _abi.PREC_MODE_FP32, "mixed": _abi.PREC_MODE_MIXED, "fp64": _abi.PREC_MODE_FP64}[args.precision]
    total = args.warmup + args.steps
    spec = mesh.plate_press_scene(n=args.n, steps=total)
    params = _abi.pinned_solver_params(1, lin_rel_tol=1e-6, precision_mode=mode,
                                       refresh_every=10 if mode == _abi.PREC_MODE_MIXED else 0)
    sc, rigid = gs.build_scene(spec, device=0, precision=mode)
    if args.burst:
        sc.set_cg_burst(args.burst)
    info = sc.info()

    def do_step(s):
        for r, c in zip(rigid, spec.colliders):
            r.set_position(c.position(s, spec.dt))
        return sc.step(spec.dt, params)[0]

    for s in range(args.warmup):
        do_step(s)
    if args.time_every > 0:
        sc.kernel_timing(0, enable=args.time_every)
    stats = []
    t0 = time.perf_counter()
    for s in range(args.warmup, total):
        stats.append(do_step(s))
    elapsed = time.perf_counter() - t0
    launches, ms = sc.kernel_timing(0) if args.time_every > 0 else (0, 0.0)
    sc.kernel_timing(0, enable=0)

    nT, nV, nB = info["n_tets"], info["n_verts"], info["nnzb"]
    cg = float(np.mean([s["linear_iters"] for s in stats]))
    ls = float(np.mean([s["line_search_iters"] for s in stats]))
    ok = all(s["linear_convergence"] == _abi.CONV_CONVERGED and s["linear_rel_res"] <= 1e-6
             and s["linear_true_rel_res"] <= 1e-6 for s in stats)
    true_rel = max(s["linear_true_rel_res"] for s in stats)
    value = nT * args.steps / elapsed
    s_bytes = 8 if mode == _abi.PREC_MODE_FP64 else 4
    per_launch = matvec_bytes(nB, nV, s_bytes)
    avg_ms = ms / launches if launches else None
    achieved = per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms else None
    traffic = read_traffic()
    step_gbs = step_bytes(nT, nB, nV, cg, ls, s_bytes) / (elapsed / args.steps) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "mixed": "f32 (f64 dots + refresh)", "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"{args.n}^3-hex soft cube ({nT} tets, {nV} verts, {nB} BSR blocks) under a 2 mm plate "
                        f"press, 1 Newton iter/step, block-Jacobi PCG rel 1e-6, precision {args.precision}",
            "tets": nT, "verts": nV, "nnzb": nB, "dt": spec.dt,
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_cg_spmv (CG matvec w = A z + z.w)",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic, "bytes_per_launch": per_launch, "avg_launch_us": 1e3 * avg_ms if avg_ms else None,
            "timed_launches": int(launches), "timed_every": args.time_every,
        },
        "step": {"cg_iters": cg, "ls_iters": ls, "all_converged": ok, "max_true_rel_res": true_rel,
                 "reference_equivalent_GBps": step_gbs},
    }
    if not args.no_cpu_baseline:
        cpu_spec = mesh.plate_press_scene(n=args.n, steps=total)
        cpu_params = _abi.pinned_solver_params(1, lin_rel_tol=1e-6)
        out["cpu_baseline"] = cpu_baseline(cpu_spec, cpu_params, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
