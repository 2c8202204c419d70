"""The batched-ensemble and serving workloads at their own sizes.

* Ensemble: 32 solves drawn as the 256-solve ensemble draws them (lattice side
  n uniform in [16, 48], jitter +-0.3h, Re in [400, 3200], Ma in [0.05, 0.3],
  TGV or k^4 by a seeded coin flip), in one batch (fused launch per kernel,
  HIP graph, autotuned variant). Every batched solve is bit-identical to the
  same solve alone, in FP64 and in mixed precision; two sampled solves go
  against the FP64 oracle for 10 steps (FP64: KE / enstrophy / sums 1e-10;
  mixed: 1e-5 with mass and energy 1e-9).
* Serving: Poisson arrivals of requests of 2e4 to 3e5 cells (log-uniform),
  continuous batching in mixed precision; every served result equals the
  same request run alone, and one served request goes against the FP64 oracle
  (KE and enstrophy 1e-5).
"""
import numpy as np
import pytest

from nextgen_amd.ensemble import ensemble_specs
from nextgen_amd.ideal_gas import CG_PREC_FP64, CG_PREC_MIXED, IdealGasSolver

pytestmark = pytest.mark.gpu

N_ENSEMBLE = 32
ORACLE_THREADS = 16


def solo(sp, precision, steps, mesh=None):
    m = mesh or sp.mesh()
    g = IdealGasSolver(m, sp.deck(precision))
    g.set_state(*sp.state(m))
    g.init_complete()
    g.advance_solution(steps)
    out = g.get_state(), g.report()
    g.close()
    return out


def oracle_history(sp, mesh, steps):
    from oracle.oracle import OracleSolver
    o = OracleSolver(mesh, sp.deck(CG_PREC_FP64), threads=ORACLE_THREADS)
    o.set_state(*sp.state(mesh))
    o.init_complete()
    hist = []
    for _ in range(steps):
        assert o.step() == 0
        hist.append(o.diag())
    return hist


def close_to(d, od, tol, cons_tol):
    assert d.step == od["step"]
    assert abs(d.ke - od["ke"]) <= tol * od["ke"], (d.ke, od["ke"])
    assert abs(d.enstrophy - od["enstrophy"]) <= tol * od["enstrophy"]
    for q in range(5):
        assert abs(d.sums[q] - od["sums"][q]) <= cons_tol * max(abs(od["sums"][q]), od["sums"][0]), q


@pytest.fixture(scope="module")
def ensemble():
    specs = ensemble_specs(N_ENSEMBLE, seed=2024)
    print(f"ensemble: {N_ENSEMBLE} solves, n {min(s.n for s in specs)}..{max(s.n for s in specs)}, "
          f"{sum(s.n ** 3 for s in specs)} cells", flush=True)
    return specs, {i: sp.mesh() for i, sp in enumerate(specs)}


@pytest.mark.parametrize("precision", [CG_PREC_FP64, CG_PREC_MIXED])
def test_ensemble_solves_equal_solo_runs_and_oracle(ensemble, precision):
    from nextgen_amd.batch import Batch
    specs, meshes = ensemble
    b = Batch(precision)
    ids = [b.add(meshes[i], sp.deck(precision), *sp.state(meshes[i])) for i, sp in enumerate(specs)]
    # two sampled solves of the batch (the largest TGV and the largest k^4)
    # against the FP64 oracle, step by step for 10 steps
    sample = [max((k for k, sp in enumerate(specs) if sp.ic == ic), key=lambda k: specs[k].n)
              for ic in ("tgv", "k4")]
    hist = {i: [] for i in sample}
    for _ in range(10):
        b.step(1)
        for i in sample:
            hist[i].append(b.diag(ids[i]))
    tol, cons = (1e-10, 1e-10) if precision == CG_PREC_FP64 else (1e-5, 1e-9)
    for i in sample:
        for d, od in zip(hist[i], oracle_history(specs[i], meshes[i], 10)):
            close_to(d, od, tol, cons)
        print(f"batched solve {i} (n {specs[i].n}, {specs[i].ic}) matches the oracle over 10 steps", flush=True)
    # the autotuner's variants (each advances the batch), then graph replays
    report = b.tune(2)
    print(report.splitlines()[-1] if report else "", flush=True)
    b.step(6)
    steps = b.diag(ids[0]).step
    assert steps >= 16 and all(b.diag(i).step == steps for i in ids)
    for i, sp in enumerate(specs):
        ref_state, ref = solo(sp, precision, steps, meshes[i])
        for x, y in zip(b.state(ids[i]), ref_state):
            np.testing.assert_array_equal(x, y)
        d = b.diag(ids[i])
        assert (d.dt, d.time, d.ke, d.enstrophy) == (ref.dt, ref.time, ref.ke, ref.enstrophy)
    b.close()


def test_served_requests_equal_solo_runs_and_oracle():
    from nextgen_amd.serve import MeshPool, Server, requests_poisson
    reqs = requests_poisson(40.0, 8, sizes=[28, 32, 36, 40, 48, 56, 64], seed=11, steps_lo=10, steps_hi=40,
                            cells_lo=2e4, cells_hi=3e5)
    pool = MeshPool()
    for r in reqs:
        pool.get(r.spec)
    print("serving: cells " + " ".join(str(r.spec.n ** 3) for r in reqs), flush=True)
    res, info = Server(pool, CG_PREC_MIXED, max_chunk=8).run(reqs)
    assert len(res) == len(reqs) and min(r.cells for r in res) >= 2e4
    by_id = {r.rid: r for r in reqs}
    for r in res:
        sp = by_id[r.rid].spec
        _, d = solo(sp, CG_PREC_MIXED, sp.steps, pool.get(sp))
        assert d.step == sp.steps
        assert (r.ke, r.enstrophy, r.time) == (d.ke, d.enstrophy, d.time)
    # one served request against the FP64 oracle (mixed tolerance)
    r0 = min(res, key=lambda r: r.cells * r.steps)
    sp = by_id[r0.rid].spec
    od = oracle_history(sp, pool.get(sp), sp.steps)[-1]
    assert abs(r0.ke - od["ke"]) <= 1e-5 * od["ke"]
    assert abs(r0.enstrophy - od["enstrophy"]) <= 1e-5 * od["enstrophy"]
    print(f"served request {r0.rid} ({r0.cells} cells, {r0.steps} steps) matches the oracle", flush=True)
