"""Benchmark of the MI355X ideal-gas LES step (one JSON line on stdout).

Workload: the FP64 drop-in case on a periodic Voronoi mesh of a 128^3 lattice
with every generator jittered uniformly by +-0.3h (seeded), 2,097,152 cells,
Morton-ordered; Taylor-Green vortex at Re 1600, Ma 0.1, Vreman SGS (c = 0.07),
SSP-RK3 at CFL 0.8. One "step" is one RK3 step (3 stages) of the whole mesh;
the metric counts cell-stage updates (3 * ncv per step). `value` is the FP64
rate; the `mixed` object is the mixed-precision step (PRECISION MIXED) on the
same mesh and initial field, timed the same way.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--ic tgv|k4]
                  [--precision both|fp64|mixed] [--no-cpu]
  python bench.py --accuracy N [--ic k4] [--every E]   (mixed vs FP64 over N steps)
  python bench.py --ensemble 256 [--ens-steps 300]       (batched ensemble vs one at a time)
  python bench.py --serve 24 [--loads 0.5 1 2]            (serving loop vs sequential FP64)

The default run also appends bounded `ensemble` (48 solves, 100 steps) and
`serving` (10 requests at 4x the sequential FP64 service rate) objects
(--no-extras skips them).

Timing: W untimed steps, then exactly K steps bracketed by stream
synchronisation and HIP events on the solver's stream (state resident in HBM).
A second pass of K steps records HIP events around every kernel launch for the
per-kernel durations used by the roofline figures.

`roofline`: the flux kernel with the largest time per step; its algorithmic
bytes per launch over its HIP-event launch time. `traffic`: HBM bytes per
launch of that kernel measured with rocprofv3 FETCH_SIZE and WRITE_SIZE passes
(scaled by the plain-copy calibration) and recorded in profiles/traffic.json
for this workload; null when no measurement of this workload is recorded.

`cpu_baseline`: the oracle (the FP64 C restatement of the reference step) on
the host's cores (OpenMP threads, at most 16), timed on a bounded sample of
the same workload (one RK3 step of the same 2.1M-cell mesh and field after a
warm-up step) in two builds: the oracle's (-O3 -ffp-contract=off) and one for
the host CPU (-O3 -march=native); also in the reference's own unit,
core-s/Mcv/step.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cell-stage updates/s FP64 vs mixed @2.1M cells; ensemble solves/s + p99 latency"
KERNEL_NAMES = {
    "flux_compact": "k_cpair: compact-pair flux, thread per pair (skew + symmetric + viscous/LF/a_sgs closure)",
    "flux_ext_rk": "k_flux_lds: compact-pair flux gather + extended-face flux + RK3 update + EOS recovery",
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(Ge, Gc, Gg, mixed=False):
    """Algorithmic HBM bytes per cell-stage for the whole step and for the two
    flux kernels of a stage (each pair's data counted once).

    FP64, whole step: extended-face n,c + 2 ids 56*Ge, compact geometry 40*Gc,
    gradient CSR 28*Gg, inv_vol 8, state read+write 80, RK registers 80,
    lagged coefficients 48, per-step pass (72 + 8*Ge)/3.
    Compact-face flux kernel: 56*Gc + 40*Gc + 48 (its pairs, their compact
    geometry, the lagged coefficients).
    Extended-face flux + RK3 kernel: 56*(Ge - Gc) + 8 + 80 + 80 (the other
    pairs, inv_vol, state read+write, RK registers).
    Mixed precision: geometry, coefficients and RK registers at 4 bytes
    (32*Ge, 20*Gc, 16*Gg, 24, RK 40, per-step (36 + 8*Ge)/3); state and
    inv_vol stay FP64.
    """
    if not mixed:
        step = 56.0 * Ge + 40.0 * Gc + 28.0 * Gg + 8 + 80 + 80 + 48 + (72.0 + 8.0 * Ge) / 3.0
        compact = 56.0 * Gc + 40.0 * Gc + 48
        ext = 56.0 * (Ge - Gc) + 8 + 80 + 80
    else:
        step = 32.0 * Ge + 20.0 * Gc + 16.0 * Gg + 8 + 80 + 40 + 24 + (36.0 + 8.0 * Ge) / 3.0
        compact = 32.0 * Gc + 20.0 * Gc + 24
        ext = 32.0 * (Ge - Gc) + 8 + 80 + 40
    return step, compact, ext


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """HBM bytes/s (read + write) of a plain device-to-device copy of nbytes
    (hipMemcpy through the HIP runtime, timed with HIP events): the practical
    bandwidth the flux kernels are also compared with."""
    import ctypes as C

    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    vp = C.c_void_p
    src, dst = vp(), vp()
    e0, e1 = vp(), vp()
    ms = C.c_float()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)), "hipMalloc")
    try:
        ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)), "hipMalloc")
        ok(hip.hipMemset(src, 1, C.c_size_t(nbytes)), "hipMemset")
        ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        kind = 3  # hipMemcpyDeviceToDevice
        for _ in range(3):
            ok(hip.hipMemcpy(dst, src, C.c_size_t(nbytes), kind), "hipMemcpy")
        ok(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(reps):
            ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), kind, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, None), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        if dst:
            hip.hipFree(dst)
        hip.hipFree(src)
    return 2.0 * nbytes / (ms.value * 1e-3 / reps)


def copy_calibration(nbytes):
    """Plain device copy of known size, for scaling rocprof FETCH_SIZE/WRITE_SIZE."""
    bw = copy_bandwidth(nbytes, reps=3)
    print(json.dumps({"copy_bytes_each_way": nbytes, "copy_GBps": bw / 1e9}))


SGS_NAMES = {"NONE": "no", "VREMAN": "Vreman", "SIGMA": "Sigma", "DSM_LOCAL": "DSM_LOCAL"}


def workload(args):
    from nextgen_amd.ic import k4_params, tgv_params
    from nextgen_amd.ideal_gas import Mesh

    t0 = time.time()
    mesh = Mesh.jittered(args.n, jitter=0.3, seed=args.seed)
    log(f"mesh: ncv={mesh.ncv} nfa={mesh.nfa} nef={mesh.nef} nnz={mesh.nnz} built in {time.time() - t0:.1f}s")
    if args.ic == "tgv":
        text = tgv_params(Re=1600.0, Ma=0.1, sgs=args.sgs, cfl=0.8)
        state = mesh.ic_tgv(0.1, 1.4)
        desc = "TGV Re 1600 Ma 0.1"
    else:
        text = k4_params(sgs=args.sgs)
        state = mesh.ic_k4(4.0, 0.3, 1.4, seed=args.seed)
        desc = f"k^4 spectrum k0 4 Mt 0.3 (seed {args.seed}), Re 1600"
    if args.scalars:
        text += "REGISTER_SCALAR " + " ".join(args.scalars) + "\n"
        desc += f", {len(args.scalars)} passive scalars"
    return mesh, text + "CHECK_INTERVAL 1000000\n", state, desc


def make_solver(mesh, text, state, precision):
    from nextgen_amd.ideal_gas import IdealGasSolver, parse_params

    t0 = time.time()
    g = IdealGasSolver(mesh, parse_params(text + f"PRECISION {precision}\n"))
    g.set_state(*state)
    g.init_complete()
    log(f"{precision} solver ready in {time.time() - t0:.1f}s")
    return g


def time_solver(g, args, label):
    g.advance_solution(args.warmup)
    d0 = g.report()
    log(f"{label} warmup done: step {d0.step} dt {d0.dt:.6e} ke {d0.ke:.12e}")
    ms, _ = g.time_steps(args.steps)
    if g.report().nan_flag:
        raise RuntimeError("NaN during the timed steps")
    log(f"{label} timed {args.steps} steps: {ms:.3f} ms ({ms / args.steps:.4f} ms/step)")
    _, kms = g.time_steps(args.steps, per_kernel=True)
    log(f"{label} per-kernel ms: flux_compact {kms[0]:.4f} flux_ext+rk {kms[1]:.4f} grad {kms[2]:.4f} "
        f"props {kms[3]:.4f} per-step sum {kms[4]:.4f}")
    return ms, kms, g.report()


def kernel_rooflines(ncv, Ge, Gc, Gg, kms, mixed):
    b_step, b_compact, b_ext = algorithmic_bytes(Ge, Gc, Gg, mixed)
    out = {}
    for key, b, t in (("flux_compact", b_compact, kms[0]), ("flux_ext_rk", b_ext, kms[1])):
        gbs = b * ncv / (t * 1e-3) / 1e9
        out[key] = {"kernel": KERNEL_NAMES[key], "achieved": gbs, "frac": gbs / HBM_PEAK_GBS, "avg_launch_ms": t,
                    "algorithmic_bytes_per_launch": b * ncv, "ms_per_step": 3 * t}
    return b_step, out


def traffic_key(key, precision, n, ic, sgs, nsc):
    """The workload a recorded traffic figure belongs to: kernel role,
    precision, lattice side, initial field, SGS model and scalar count."""
    return f"{key}/{precision}/n{n}/{ic}/{sgs.lower()}/sc{nsc}"


def recorded_traffic(key, precision, n, ic, sgs="VREMAN", nsc=0):
    """HBM bytes per launch recorded from rocprofv3 passes of exactly this
    workload (None when no pass covered it)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get(traffic_key(key, precision, n, ic, sgs, nsc), {}).get("bytes_per_launch")


def host_cores():
    """Host threads this process may use (the GPU box grants 16 per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle_build():
    """ig_oracle.c compiled for this host CPU (-O3 -march=native, the
    production-like build; contraction left at gcc's default), or None."""
    import shutil
    import subprocess
    import tempfile

    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        return None, "no C compiler on this host"
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle_native.so")
    cmd = [cc, "-O3", "-march=native", "-std=c11", "-fPIC", "-fopenmp", "-shared", "-o", out,
           os.path.join(ROOT, "oracle", "ig_oracle.c"), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, r.stderr[-300:]
    return out, " ".join(cmd[1:6])


def cpu_baseline(mesh, text, state, steps=1):
    """The oracle (FP64 C restatement of the reference step) on the host's
    cores: one warm-up step, then `steps` timed RK3 steps of the same mesh and
    field, in the oracle build (-O3 -ffp-contract=off) and in a build for this
    host CPU (-O3 -march=native). Reports cell-stage updates/s of the faster
    build and core-s/Mcv/step (FlowSolver.hpp:2596-2605) of both."""
    from nextgen_amd.ideal_gas import parse_params
    from oracle.oracle import ORACLE_SO, OracleSolver

    cores = host_cores()
    ncv = mesh.ncv
    builds = [("oracle", ORACLE_SO, "gcc -O3 -ffp-contract=off -fopenmp")]
    native, flags = native_oracle_build()
    if native:
        builds.append(("native", native, "gcc " + flags + " -fopenmp"))
    else:
        log(f"cpu baseline: no native build ({flags})")
    res = {}
    for name, path, how in builds:
        t0 = time.time()
        o = OracleSolver(mesh, parse_params(text), threads=cores, libpath=path)
        o.set_state(*state)
        o.init_complete()
        o.step(1)
        t1 = time.time()
        o.step(steps)
        sec = (time.time() - t1) / steps
        o.close()
        res[name] = {"s_per_step": sec, "cell_stage_per_s": 3.0 * ncv / sec,
                     "core_s_per_Mcv_per_step": cores * sec / (ncv * 1e-6), "compiler": how}
        log(f"cpu baseline ({name}, {cores} threads): {sec:.2f} s/step, "
            f"{cores * sec / (ncv * 1e-6):.2f} core-s/Mcv/step (ready in {t1 - t0:.1f}s)")
    best = min(res, key=lambda k: res[k]["s_per_step"])
    b = res[best]
    return {
        "value": b["cell_stage_per_s"],
        "unit": "cell-stage updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{steps} RK3 step(s) after one warm-up step, the same {ncv}-cell mesh and initial field, "
                  f"FP64 C restatement of the reference step on {cores} OpenMP threads ({cpu_model()}); "
                  f"value from the {best} build ({b['compiler']})",
        "cpu_model": cpu_model(),
        "builds": res,
        "core_s_per_Mcv_per_step": b["core_s_per_Mcv_per_step"],
    }


def accuracy(args):
    """Mixed precision against FP64 on the same mesh and field over N steps."""
    mesh, text, state, desc = workload(args)
    g64 = make_solver(mesh, text, state, "FP64")
    gmx = make_solver(mesh, text, state, "MIXED")
    rows = []
    worst = {"ke": 0.0, "enstrophy": 0.0, "mass": 0.0, "energy": 0.0}
    done = 0
    d64, dmx = g64.report(), gmx.report()
    m0, e0 = d64.sums[0], d64.sums[4]
    while done < args.accuracy:
        k = min(args.every, args.accuracy - done)
        g64.advance_solution(k)
        gmx.advance_solution(k)
        done += k
        d64, dmx = g64.report(), gmx.report()
        err = {"ke": abs(dmx.ke - d64.ke) / d64.ke, "enstrophy": abs(dmx.enstrophy - d64.enstrophy) / d64.enstrophy,
               "mass": abs(dmx.sums[0] - d64.sums[0]) / m0, "energy": abs(dmx.sums[4] - d64.sums[4]) / e0}
        for key in worst:
            worst[key] = max(worst[key], err[key])
        rows.append({"step": d64.step, "time_fp64": d64.time, "time_mixed": dmx.time, "ke_fp64": d64.ke,
                     "ke_mixed": dmx.ke, "enstrophy_fp64": d64.enstrophy, "enstrophy_mixed": dmx.enstrophy, **err})
        log(f"step {d64.step}: ke {d64.ke:.10e} / {dmx.ke:.10e} err {err['ke']:.2e}, enstrophy err "
            f"{err['enstrophy']:.2e}, mass {err['mass']:.2e}, energy {err['energy']:.2e}")
    out = {"mode": "accuracy", "workload": f"{mesh.ncv} cells, {desc}", "steps": args.accuracy,
           "reference": "FP64 GPU step (parity-tested against the FP64 oracle)",
           "max_rel_err": worst,
           "conservation_drift": {"fp64_mass": abs(d64.sums[0] - m0) / m0, "mixed_mass": abs(dmx.sums[0] - m0) / m0,
                                  "fp64_energy": abs(d64.sums[4] - e0) / e0,
                                  "mixed_energy": abs(dmx.sums[4] - e0) / e0},
           "history": rows}
    print(json.dumps(out), flush=True)
    return 0


def ensemble_leg(count, steps, seed, tune_steps=2, sample_count=16):
    """Batched ensemble: `count` solves of uneven size (n ~ U{16..48}, jitter
    0.3h, Re, Ma, TGV or k^4 as the ensemble draws them) and uneven horizon
    (U{steps/2..steps} steps) in one batch (one fused launch per kernel, graph
    replay, autotuned variant), FP64 and mixed. All solves start together; the
    batch is stepped to the next horizon, where the solves that reach it are
    read back and retired, so each solve's latency is its own completion time
    (p50 / p99 over the solves). Against the same solves run one at a time in
    FP64 (a uniform sample, scaled)."""
    from nextgen_amd.batch import Batch
    from nextgen_amd.ensemble import ensemble_specs, run_alone
    from nextgen_amd.ideal_gas import CG_PREC_FP64, CG_PREC_MIXED
    import numpy as np

    specs = ensemble_specs(count, seed=seed)
    rng = np.random.default_rng(seed + 1)
    horizons = [int(h) for h in rng.integers(max(1, steps // 2), steps + 1, size=len(specs))]
    t0 = time.time()
    meshes = [sp.mesh() for sp in specs]
    cells = sum(m.ncv for m in meshes)
    cell_steps = sum(m.ncv * h for m, h in zip(meshes, horizons))
    log(f"ensemble: {len(specs)} meshes, {cells} cells, built in {time.time() - t0:.1f}s")
    states = [sp.state(m) for sp, m in zip(specs, meshes)]
    out = {"instances": len(specs), "cells": cells, "steps": [min(horizons), max(horizons)],
           "spec": f"n ~ U{{16..48}} (jitter 0.3h), Re ~ U[400,3200], Ma ~ U[0.05,0.3], TGV or k^4 by coin flip, "
                   f"horizon ~ U{{{max(1, steps // 2)}..{steps}}} steps, seed {seed}"}
    kes = {}
    for prec, name in ((CG_PREC_FP64, "fp64"), (CG_PREC_MIXED, "mixed")):
        t0 = time.time()
        b = Batch(prec)
        ids = [b.add(m, sp.deck(prec, check_interval=100), *st) for sp, m, st in zip(specs, meshes, states)]
        setup = time.time() - t0
        cnt = b.counts()
        log(f"{name}: {cnt['instances']} instances on the device ({cnt['device_bytes'] / 1e9:.1f} GB) in {setup:.1f}s")
        # the autotuner picks the launch variant (LDS or global gather, extended
        # slots per batch) for this mix; every variant advances the solves by
        # tune_steps, the same in both precisions
        report = b.tune(tune_steps)
        chosen = [ln for ln in report.splitlines() if ln.startswith("chosen")]
        log(f"{name}: tuner {chosen[0] if chosen else report}")
        left = dict(zip(ids, horizons))
        base = {iid: b.diag(iid).step for iid in ids}  # the horizon counts net steps after tuning
        done, kes[name] = {}, {}
        tstart = time.perf_counter()
        while left:
            k = min(left.values())
            b.step(k)
            t = time.perf_counter() - tstart
            for iid in list(left):
                left[iid] -= k
                if left[iid] > 0:
                    continue
                d = b.diag(iid)
                target = base[iid] + horizons[ids.index(iid)]
                if d.step < target:
                    left[iid] = target - d.step  # a NaN rollback: net steps, as the reference run loop
                    continue
                kes[name][iid] = d
                done[iid] = t
                b.remove(iid)
                del left[iid]
        sec = time.perf_counter() - tstart
        b.close()
        lat = [done[i] for i in ids]
        out[name] = {"seconds": sec, "solves_per_s": len(specs) / sec,
                     "cell_stage_per_s": 3.0 * cell_steps / sec,
                     "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
                     "setup_s": setup, "device_bytes": cnt["device_bytes"],
                     "tuner": chosen[0] if chosen else report}
        log(f"{name}: all solves to their horizons in {sec:.2f}s -> {len(specs) / sec:.2f} solves/s, "
            f"latency p50 {out[name]['latency_p50_s']:.3f}s p99 {out[name]['latency_p99_s']:.3f}s")
    errs = [abs(kes["mixed"][i].ke - kes["fp64"][i].ke) / kes["fp64"][i].ke for i in kes["fp64"]]
    out["mixed"]["ke_rel_err_vs_fp64_max"] = max(errs)
    out["tune_steps_per_variant"] = tune_steps
    # one at a time, FP64 (a uniform sample, scaled to the whole set)
    sample = list(range(0, len(specs), max(1, len(specs) // sample_count)))
    t_alone = 0.0
    for i in sample:
        sec, _ = run_alone(specs[i], CG_PREC_FP64, horizons[i])
        t_alone += sec
    seq = t_alone * len(specs) / len(sample)
    out["sequential_fp64"] = {"seconds_estimate": seq, "solves_per_s": len(specs) / seq,
                              "sample": f"{len(sample)} of {len(specs)} solves timed alone (GPU stepping only), scaled"}
    out["speedup_vs_sequential_fp64"] = {k: out[k]["solves_per_s"] / out["sequential_fp64"]["solves_per_s"]
                                         for k in ("fp64", "mixed")}
    return out


def ensemble_bench(args):
    out = {"mode": "ensemble", **ensemble_leg(args.ensemble, args.ens_steps, args.seed)}
    print(json.dumps(out), flush=True)
    return 0


def serve_leg(count, loads, seed, hbm_budget=96e9, workers=3, seq_mixed=True, rate=None):
    """Serving loop (continuous batching, mixed precision) against the same
    requests one at a time (FP64, and in the server's own precision), over
    arrival rates given as multiples of the sequential FP64 service rate. Each
    request has its own mesh and initial field (built before the clock, as a
    request's files would be read); its setup (device layout, state upload,
    initComplete) counts in both executions. With `rate`
    (requests/s) only the server runs, at that rate (for a kernel trace of the
    serving loop alone)."""
    from nextgen_amd.ideal_gas import CG_PREC_FP64, CG_PREC_MIXED
    from nextgen_amd.serve import MeshPool, Server, requests_poisson, sequential, summarize

    pool = MeshPool()
    base = requests_poisson(1.0, count, None, seed=seed)
    t0 = time.time()
    pool.build([r.spec for r in base])
    cells = [r.spec.n ** 3 for r in base]
    log(f"serving: {len(base)} request meshes and fields ({min(cells)}..{max(cells)} cells) built in "
        f"{time.time() - t0:.1f}s")
    if rate is not None:
        reqs = [type(r)(r.rid, r.arrival / rate, r.spec) for r in base]
        srv = Server(pool, CG_PREC_MIXED, hbm_budget=hbm_budget, workers=workers)
        res, info = srv.run(reqs)
        return {"requests": len(base), "cells": cells, "rate_per_s": rate,
                "served_mixed": summarize(res, info["makespan"]), **info,
                "timeline": [{"rid": r.rid, "cells": r.cells, "steps": r.steps, "arrival": r.arrival,
                              "attached": r.start, "done": r.done, "prep_s": srv.prep_seconds.get(r.rid)}
                             for r in sorted(res, key=lambda q: q.rid)]}
    seq = {}
    for prec, name in ((CG_PREC_FP64, "fp64"), (CG_PREC_MIXED, "mixed")):
        if name == "mixed" and not seq_mixed:
            continue
        t0 = time.time()
        seq[name] = sequential(pool, base, prec)
        log(f"serving: sequential {name} over all requests in {time.time() - t0:.1f}s")
    service = {k: {r.rid: r.done - r.start for r in v} for k, v in seq.items()}
    mean_service = sum(service["fp64"].values()) / len(base)
    ke64 = {r.rid: r.ke for r in seq["fp64"]}
    out = {"requests": len(base), "cells": cells, "horizons": [r.spec.steps for r in base],
           "spec": "cells log-uniform [2e4, 2e6] (own mesh and initial field per request), horizon U{200..2000} steps, "
                   "Re U[400,3200], Ma U[0.05,0.3], TGV or k^4; Poisson arrivals; setup counted in both",
           "sequential_fp64_mean_service_s": mean_service, "workers": workers, "sweep": []}

    def queue(reqs, svc):
        clock, res = 0.0, []
        for r in sorted(reqs, key=lambda q: q.arrival):
            start = max(clock, r.arrival)
            clock = start + svc[r.rid]
            res.append(type(seq["fp64"][0])(r.rid, r.arrival, start, clock, 0, 0, 0.0, 0.0, 0.0))
        return res

    for load in loads:
        rate = load / mean_service
        reqs = [type(r)(r.rid, r.arrival / rate, r.spec) for r in base]
        srv = Server(pool, CG_PREC_MIXED, hbm_budget=hbm_budget, workers=workers)
        res, info = srv.run(reqs)
        errs = [abs(r.ke - ke64[r.rid]) / ke64[r.rid] for r in res]
        row = {"load": load, "rate_per_s": rate, "served_mixed": summarize(res, info["makespan"]),
               "sequential_fp64": summarize(queue(reqs, service["fp64"])),
               "ke_rel_err_vs_fp64_max": max(errs), "peak_device_bytes": srv.peak_bytes,
               **{k: v for k, v in info.items() if k != "makespan"}}
        if "mixed" in service:
            row["sequential_mixed"] = summarize(queue(reqs, service["mixed"]))
            row["solves_per_s_ratio_vs_sequential_mixed"] = (row["served_mixed"]["solves_per_s"] /
                                                             row["sequential_mixed"]["solves_per_s"])
        row["solves_per_s_ratio"] = row["served_mixed"]["solves_per_s"] / row["sequential_fp64"]["solves_per_s"]
        out["sweep"].append(row)
        log(f"load {load}: served {row['served_mixed']['solves_per_s']:.3f}/s p50 {row['served_mixed']['tts_p50_s']:.1f}s "
            f"p99 {row['served_mixed']['tts_p99_s']:.1f}s vs sequential FP64 {row['sequential_fp64']['solves_per_s']:.3f}/s "
            f"p99 {row['sequential_fp64']['tts_p99_s']:.1f}s ({row['solves_per_s_ratio']:.2f}x)")
    return out


def serve_bench(args):
    out = {"mode": "serving", **serve_leg(args.serve, args.loads, args.seed, args.hbm_budget * 1e9,
                                          rate=args.serve_rate)}
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=128, help="lattice side of the jittered mesh")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--ic", choices=["tgv", "k4"], default="tgv")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both")
    ap.add_argument("--sgs", choices=["NONE", "VREMAN", "SIGMA", "DSM_LOCAL"], default="VREMAN",
                    help="SGS_MODEL of the step workload")
    ap.add_argument("--scalars", nargs="*", default=[], help="passive scalars to register (names)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--accuracy", type=int, default=0, help="compare mixed with FP64 over this many steps")
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--copy-calibration", type=int, default=0, help="run only a device copy of this many bytes")
    ap.add_argument("--hbm-probe", type=float, default=0.0,
                    help="run only the HBM reference probe (streaming reads, copy, gather) on buffers of this many GiB")
    ap.add_argument("--ensemble", type=int, default=0, help="batched ensemble of this many solves")
    ap.add_argument("--ens-steps", type=int, default=300)
    ap.add_argument("--serve", type=int, default=0, help="serving loop over this many requests")
    ap.add_argument("--loads", type=float, nargs="+", default=[0.5, 1.0, 2.0],
                    help="arrival rates as multiples of the sequential FP64 service rate")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="run only the server, at this many requests/s (no sequential baselines)")
    ap.add_argument("--hbm-budget", type=float, default=96.0, help="GB of device memory for admitted solves")
    ap.add_argument("--no-extras", action="store_true", help="skip the bounded ensemble and serving legs")
    ap.add_argument("--extra-ensemble", type=int, default=48, help="solves of the default run's ensemble leg")
    ap.add_argument("--extra-ens-steps", type=int, default=100)
    ap.add_argument("--extra-serve", type=int, default=10, help="requests of the default run's serving leg")
    ap.add_argument("--extra-load", type=float, nargs="+", default=[4.0, 8.0],
                    help="arrival rates of the default run's serving leg, as multiples of the sequential FP64 rate")
    args = ap.parse_args()

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return 0
    if args.hbm_probe:
        # a child process (built with the library): this one has not touched the GPU
        probe = os.path.join(ROOT, "nextgen_amd", "lib", "hbm_probe")
        if not os.path.exists(probe):
            print(f"{probe} is not built (make -C nextgen_amd/csrc)", file=sys.stderr)
            return 1
        return subprocess.run([probe, str(args.hbm_probe)]).returncode
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported by this build "
              "(the path is single-GPU: no sharding, no collectives)", file=sys.stderr)
        return 2
    if args.accuracy:
        return accuracy(args)
    if args.ensemble:
        return ensemble_bench(args)
    if args.serve:
        return serve_bench(args)

    mesh, text, state, desc = workload(args)
    ncv = mesh.ncv
    Ge, Gc, Gg = mesh.nef / ncv, mesh.nfa / ncv, mesh.nnz / ncv
    res = {}
    for prec in (["fp64", "mixed"] if args.precision == "both" else [args.precision]):
        g = make_solver(mesh, text, state, prec.upper())
        ms, kms, d = time_solver(g, args, prec)
        g.close()
        b_step, kr = kernel_rooflines(ncv, Ge, Gc, Gg, kms, prec == "mixed")
        rate = 3.0 * ncv * args.steps / (ms * 1e-3)
        res[prec] = {"value": rate, "ms_per_step": ms / args.steps, "kms": kms, "kr": kr, "b_step": b_step, "diag": d}
    main_prec = "fp64" if "fp64" in res else "mixed"
    r = res[main_prec]

    cpu = None if args.no_cpu else cpu_baseline(mesh, text, state)

    dom = max(r["kr"], key=lambda k: r["kr"][k]["ms_per_step"])
    kd = r["kr"][dom]
    traffic = recorded_traffic(dom, main_prec, args.n, args.ic, args.sgs, len(args.scalars))
    copy_gbs = copy_bandwidth() / 1e9
    log(f"device copy: {copy_gbs:.0f} GB/s")
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic: seeded jittered-lattice Voronoi mesh, analytic initial field",
        "config": {
            "workload": f"{'FP64 drop-in' if main_prec == 'fp64' else 'mixed precision'}: periodic Voronoi mesh "
                        f"of a {args.n}^3 lattice jittered +-0.3h (seed {args.seed}), {ncv} cells, {desc}, "
                        f"{SGS_NAMES[args.sgs]} SGS, RK3 CFL 0.8",
            "ncv": ncv, "nfa": mesh.nfa, "nef": mesh.nef, "nnz": mesh.nnz,
            "G_e": Ge, "G_c": Gc, "G_g": Gg, "precision": main_prec,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": kd["achieved"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kd["frac"],
            "traffic": traffic,
            "kernel": kd["kernel"],
            "algorithmic_bytes_per_launch": kd["algorithmic_bytes_per_launch"],
            "avg_launch_ms": kd["avg_launch_ms"],
            "measured_GBps": traffic / (kd["avg_launch_ms"] * 1e-3) / 1e9 if traffic else None,
            "copy_GBps": copy_gbs,
        },
        "kernel_rooflines": r["kr"],
        "step_roofline": {"achieved": r["b_step"] * r["value"] / 1e9, "frac": r["b_step"] * r["value"] / 1e9 / HBM_PEAK_GBS,
                          "algorithmic_bytes_per_cell_stage": r["b_step"]},
        "kernel_ms": dict(zip(["flux_compact", "flux_ext_rk", "grad", "props", "per_step_sum"], r["kms"])),
        "diag": {"step": r["diag"].step, "time": r["diag"].time, "ke": r["diag"].ke,
                 "enstrophy": r["diag"].enstrophy},
        "cpu_baseline": cpu,
    }
    if "mixed" in res and main_prec == "fp64":
        mx = res["mixed"]
        out["mixed"] = {
            "value": mx["value"], "ms_per_step": mx["ms_per_step"], "speedup_vs_fp64": mx["value"] / r["value"],
            "kernel_ms": dict(zip(["flux_compact", "flux_ext_rk", "grad", "props", "per_step_sum"], mx["kms"])),
            "kernel_rooflines": mx["kr"],
            "step_roofline": {"achieved": mx["b_step"] * mx["value"] / 1e9,
                              "frac": mx["b_step"] * mx["value"] / 1e9 / HBM_PEAK_GBS,
                              "algorithmic_bytes_per_cell_stage": mx["b_step"]},
            "diag": {"step": mx["diag"].step, "ke": mx["diag"].ke, "enstrophy": mx["diag"].enstrophy,
                     "ke_rel_err_vs_fp64": abs(mx["diag"].ke - r["diag"].ke) / r["diag"].ke},
        }
    if not args.no_extras:
        # bounded forms of the ensemble and serving workloads (the metric's
        # second half); full runs: --ensemble 256, --serve N
        out["ensemble"] = ensemble_leg(args.extra_ensemble, args.extra_ens_steps, args.seed, sample_count=8)
        out["serving"] = serve_leg(args.extra_serve, args.extra_load, args.seed, args.hbm_budget * 1e9,
                                   seq_mixed=True)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
