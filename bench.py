"""Benchmark of the MI355X ideal-gas LES step (one JSON line on stdout).

Workload: the FP64 drop-in case on a periodic Voronoi mesh of a 128^3 lattice
with every generator jittered uniformly by +-0.3h (seeded), 2,097,152 cells,
Morton-ordered; Taylor-Green vortex at Re 1600, Ma 0.1, Vreman SGS (c = 0.07),
SSP-RK3 at CFL 0.8. One "step" is one RK3 step (3 stages) of the whole mesh;
the metric counts cell-stage updates (3 * ncv per step).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]

Timing: W untimed steps, then exactly K steps bracketed by stream
synchronisation and HIP events on the solver's stream (state resident in HBM).
A second pass of K steps records HIP events around every kernel launch for the
per-kernel durations used by the roofline figure.

`cpu_baseline`: the oracle (the FP64 C restatement of the reference step,
single thread) timed on a bounded sample of the same workload: one RK3 step
of the same 2.1M-cell mesh and initial field.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-stage updates/s FP64 vs mixed @2.1M cells; ensemble solves/s + p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(Ge, Gc, Gg):
    """Algorithmic HBM bytes per cell-stage (FP64) for the whole step and for
    the two flux kernels of a stage.

    Whole step: extended-face n,c + 2 ids 56*Ge, compact geometry 40*Gc,
    gradient CSR 28*Gg, inv_vol 8, state read+write 80, RK registers 80,
    lagged coefficients 48, per-step pass (72 + 8*Ge)/3.
    Compact-face flux kernel: 56*Gc + 40*Gc + 48 (its pairs, their compact
    geometry, the lagged coefficients).
    Extended-face flux + RK3 kernel: 56*(Ge - Gc) + 8 + 80 + 80 (the other
    pairs, inv_vol, state read+write, RK registers).
    """
    step = 56.0 * Ge + 40.0 * Gc + 28.0 * Gg + 8 + 80 + 80 + 48 + (72.0 + 8.0 * Ge) / 3.0
    compact = 56.0 * Gc + 40.0 * Gc + 48
    ext = 56.0 * (Ge - Gc) + 8 + 80 + 80
    return step, compact, ext


def copy_calibration(nbytes):
    """Plain device copy of known size, for scaling rocprof FETCH_SIZE/WRITE_SIZE."""
    import torch

    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128, help="lattice side of the jittered mesh")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--copy-calibration", type=int, default=0, help="run only a device copy of this many bytes")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU: no sharding, no collectives)", file=sys.stderr)
        return 2

    from nextgen_amd.ic import tgv_params, tgv_state
    from nextgen_amd.ideal_gas import IdealGasSolver, Mesh, parse_params

    params = parse_params(tgv_params(Re=1600.0, Ma=0.1, sgs="VREMAN", cfl=0.8) + "CHECK_INTERVAL 1000000\n")
    t0 = time.time()
    mesh = Mesh.jittered(args.n, jitter=0.3, seed=args.seed)
    log(f"mesh: ncv={mesh.ncv} nfa={mesh.nfa} nef={mesh.nef} nnz={mesh.nnz} built in {time.time() - t0:.1f}s")
    ncv = mesh.ncv
    Ge, Gc, Gg = mesh.nef / ncv, mesh.nfa / ncv, mesh.nnz / ncv
    x_cv = mesh.get("x_cv")
    rho, u, rhoE = tgv_state(x_cv)

    t0 = time.time()
    g = IdealGasSolver(mesh, params)
    g.set_state(rho, u, rhoE)
    g.init_complete()
    log(f"solver ready in {time.time() - t0:.1f}s")
    g.advance_solution(args.warmup)
    d0 = g.report()
    log(f"warmup done: step {d0.step} dt {d0.dt:.6e} ke {d0.ke:.12e}")

    ms, _ = g.time_steps(args.steps)
    d1 = g.report()
    if d1.nan_flag:
        raise RuntimeError("NaN during the timed steps")
    log(f"timed {args.steps} steps: {ms:.3f} ms ({ms / args.steps:.4f} ms/step)")
    ms2, kms = g.time_steps(args.steps, per_kernel=True)
    log(f"per-kernel ms: flux_compact {kms[0]:.4f} flux_ext+rk {kms[1]:.4f} grad {kms[2]:.4f} "
        f"props {kms[3]:.4f} per-step sum {kms[4]:.4f}")
    d2 = g.report()

    cell_stage_per_s = 3.0 * ncv * args.steps / (ms * 1e-3)
    b_step, b_compact, b_ext = algorithmic_bytes(Ge, Gc, Gg)
    ext_gbs = b_ext * ncv / (kms[1] * 1e-3) / 1e9
    compact_gbs = b_compact * ncv / (kms[0] * 1e-3) / 1e9
    step_gbs = b_step * cell_stage_per_s / 1e9

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import OracleSolver

        t0 = time.time()
        o = OracleSolver(mesh, params)
        o.set_state(rho, u, rhoE)
        o.init_complete()
        log(f"oracle ready in {time.time() - t0:.1f}s")
        t0 = time.time()
        o.step(1)
        dt_cpu = time.time() - t0
        cpu = {
            "value": 3.0 * ncv / dt_cpu,
            "unit": "cell-stage updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"1 RK3 step of the same {ncv}-cell mesh and TGV field, FP64 C restatement "
                      f"(gcc -O2 -ffp-contract=off), single thread: {dt_cpu:.2f} s",
        }
        log(f"cpu baseline: {dt_cpu:.2f}s per step")
        o.close()

    out = {
        "metric": METRIC,
        "value": cell_stage_per_s,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded jittered-lattice Voronoi mesh, analytic TGV initial field",
        "config": {
            "workload": f"FP64 drop-in: periodic Voronoi mesh of a {args.n}^3 lattice jittered +-0.3h "
                        f"(seed {args.seed}), {ncv} cells, TGV Re 1600 Ma 0.1, Vreman SGS, RK3 CFL 0.8",
            "ncv": ncv, "nfa": mesh.nfa, "nef": mesh.nef, "nnz": mesh.nnz,
            "G_e": Ge, "G_c": Gc, "G_g": Gg, "precision": "fp64",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": ext_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ext_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_flux<STAGE,1>: extended-face flux gather + RK3 update + EOS recovery (average of its 3 stage launches)",
            "algorithmic_bytes_per_launch": b_ext * ncv,
            "avg_launch_ms": kms[1],
        },
        "compact_kernel_roofline": {"kernel": "k_flux_compact", "achieved": compact_gbs,
                                    "frac": compact_gbs / HBM_PEAK_GBS, "avg_launch_ms": kms[0],
                                    "algorithmic_bytes_per_launch": b_compact * ncv},
        "step_roofline": {"achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_cell_stage": b_step},
        "kernel_ms": {"flux_compact": kms[0], "flux_ext_rk": kms[1], "grad": kms[2], "props": kms[3],
                      "per_step_sum": kms[4]},
        "diag": {"step": d2.step, "time": d2.time, "ke": d2.ke, "enstrophy": d2.enstrophy},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
