"""Batches (cg_batch_*): many solves in one fused launch per kernel.

Each instance of a batch must give bit-identical results to the same solve run
alone through cg_create/cg_step, in FP64 and in mixed precision, whatever
else is in the batch and whether instances join or leave midway; the HIP graph
replay and the LDS / global-gather variants of the extended-face pass must
agree bit for bit with eager launches.
"""
import numpy as np
import pytest

from nextgen_amd.batch import Batch
from nextgen_amd.ic import tgv_params
from nextgen_amd.ideal_gas import CG_PREC_FP64, CG_PREC_MIXED, IdealGasSolver, Mesh, parse_params

pytestmark = pytest.mark.gpu


def k4_params(Mt=0.3, Re=800.0):
    return (f"GAMMA 1.4\nP_REF {1 / 1.4!r}\nRHO_REF 1.0\nT_REF 1.0\nMU_REF {Mt / Re!r}\n"
            f"MU_POWER_LAW 0.75\nPR_LAM 0.7\nSGS_MODEL VREMAN\nTIMESTEP CFL 0.8\n")


# (lattice side, jitter seed, initial field, Re, Ma)
CASES = [(8, 3, "tgv", 400.0, 0.1), (10, 4, "k4", 1600.0, 0.2), (12, 5, "tgv", 3200.0, 0.05), (9, 6, "k4", 900.0, 0.3)]


def case(i, precision):
    n, seed, ic, Re, Ma = CASES[i]
    m = Mesh.jittered(n, 0.3, seed=seed)
    prec = "MIXED" if precision == CG_PREC_MIXED else "FP64"
    if ic == "tgv":
        text = tgv_params(Re=Re, Ma=Ma, sgs="VREMAN")
        state = m.ic_tgv(Ma, 1.4)
    else:
        text = k4_params(Mt=Ma, Re=Re)
        state = m.ic_k4(2.0, Ma, 1.4, seed=seed)
    return m, parse_params(text + f"CHECK_INTERVAL 3\nPRECISION {prec}\n"), state


def add_case(b, i, precision):
    m, p, state = case(i, precision)
    return b.add(m, p, *state)


def alone(i, precision, nsteps):
    m, p, state = case(i, precision)
    g = IdealGasSolver(m, p)
    g.set_state(*state)
    g.init_complete()
    g.advance_solution(nsteps)
    out = g.get_state(), g.report()
    g.close()
    return out


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", [CG_PREC_FP64, CG_PREC_MIXED])
def test_batch_instances_match_solo_runs(precision):
    b = Batch(precision)
    ids = []
    for i in range(len(CASES)):
        m, p, state = case(i, precision)
        ids.append(b.add(m, p, *state))
    assert b.counts()["instances"] == len(CASES)
    b.step(7)
    for i, iid in enumerate(ids):
        ref_state, ref_diag = alone(i, precision, 7)
        assert_same(b.state(iid), ref_state)
        d = b.diag(iid)
        assert d.step == ref_diag.step == 7
        assert d.dt == ref_diag.dt and d.time == ref_diag.time and d.ke == ref_diag.ke
    b.close()


def test_batch_join_and_leave():
    prec = CG_PREC_MIXED
    b = Batch(prec)
    a = add_case(b, 0, prec)
    c = add_case(b, 1, prec)
    b.step(4)
    b.remove(a)
    d = add_case(b, 2, prec)
    b.step(5)
    assert_same(b.state(c), alone(1, prec, 9)[0])
    assert_same(b.state(d), alone(2, prec, 5)[0])
    b.close()


@pytest.mark.parametrize("precision", [CG_PREC_FP64, CG_PREC_MIXED])
def test_removed_solves_leave_the_step_graphs_in_place(precision):
    # solves removed between graph replays: their block-map and table entries
    # are emptied in place (the captured graphs stay) until half of the blocks
    # are empty, then the maps are rebuilt; the others stay bit-identical to
    # their solo runs throughout, and a solve can join afterwards
    b = Batch(precision)
    ids = [add_case(b, i, precision) for i in range(4)]
    b.step(4)
    b.remove(ids[0])
    b.step(4)
    for i in (1, 2, 3):
        assert_same(b.state(ids[i]), alone(i, precision, 8)[0])
    b.remove(ids[2])  # 12^3 of 8^3 + 10^3 + 12^3 + 9^3 cells: past half, a rebuild
    b.step(2)
    b.remove(ids[1])
    b.step(3)
    assert_same(b.state(ids[3]), alone(3, precision, 13)[0])
    late = add_case(b, 0, precision)
    b.step(4)
    assert_same(b.state(ids[3]), alone(3, precision, 17)[0])
    assert_same(b.state(late), alone(0, precision, 4)[0])
    assert b.diag(ids[3]).step == 17
    b.close()


@pytest.mark.parametrize("option,values", [("graph", (1, 0)), ("lds", (1, 0)), ("tile", (4, 2, 8))])
def test_batch_variants_bitwise(option, values):
    prec = CG_PREC_FP64
    runs = []
    for value in values:
        b = Batch(prec)
        b.set_option(option, value)
        ids = [add_case(b, i, prec) for i in range(3)]
        snaps = []
        # replays interleaved with host reads (each chunk of 4 steps = 2 replays)
        for _ in range(3):
            b.step(4)
            snaps.append([b.state(i) for i in ids])
        runs.append(snaps)
        b.close()
    for other in runs[1:]:
        for s1, s0 in zip(runs[0], other):
            for x, y in zip(s1, s0):
                assert_same(x, y)


def test_tuner_keeps_results_exact():
    prec = CG_PREC_MIXED
    b = Batch(prec)
    ids = [add_case(b, i, prec) for i in range(3)]
    report = b.tune(2)  # every variant advances the instances (graph capture + timed steps)
    assert "chosen" in report
    b.step(3)
    for i, iid in enumerate(ids):
        n = b.diag(iid).step
        assert n > 3
        assert_same(b.state(iid), alone(i, prec, n)[0])
    b.close()


def test_instance_prepared_on_worker_thread_then_attached():
    import threading

    prec = CG_PREC_MIXED
    from nextgen_amd.batch import Instance

    b = Batch(prec)
    a = add_case(b, 0, prec)
    box = {}

    def prepare():
        m, p, state = case(1, prec)
        box["inst"] = Instance(m, p, *state)

    t = threading.Thread(target=prepare)
    t.start()
    b.step(3)  # the batch keeps stepping while the worker builds
    t.join()
    c = b.attach(box["inst"])
    b.step(4)
    assert_same(b.state(a), alone(0, prec, 7)[0])
    assert_same(b.state(c), alone(1, prec, 4)[0])
    b.close()


def test_served_results_equal_solo_runs():
    from nextgen_amd.ensemble import Spec
    from nextgen_amd.serve import MeshPool, Request, Server

    specs = [Spec(n=8, mesh_seed=3, ic="tgv", Re=500.0, Ma=0.1, ic_seed=1, steps=12),
             Spec(n=10, mesh_seed=4, ic="k4", Re=900.0, Ma=0.2, ic_seed=2, steps=7),
             Spec(n=9, mesh_seed=5, ic="tgv", Re=1500.0, Ma=0.05, ic_seed=3, steps=20)]
    reqs = [Request(i, 0.01 * i, sp) for i, sp in enumerate(specs)]
    pool = MeshPool()
    for r in reqs:
        pool.get(r.spec)
    res, info = Server(pool, CG_PREC_MIXED, max_chunk=4).run(reqs)
    assert len(res) == 3
    for r in res:
        sp = specs[r.rid]
        m = pool.get(sp)
        g = IdealGasSolver(m, sp.deck(CG_PREC_MIXED, check_interval=100))
        g.set_state(*sp.state(m))
        g.init_complete()
        g.advance_solution(sp.steps)
        d = g.report()
        g.close()
        assert d.step == sp.steps and r.ke == d.ke and r.time == d.time and r.enstrophy == d.enstrophy


def test_solves_sharing_a_layout_match_solo_runs():
    from nextgen_amd.batch import Instance, Layout

    prec = CG_PREC_MIXED
    m, p, state = case(1, prec)
    lay = Layout(m, prec)
    b = Batch(prec)
    a = b.attach(Instance(m, p, *state, layout=lay))
    # the same mesh with another deck and field
    n, seed, ic, Re, Ma = CASES[1]
    p2 = parse_params(tgv_params(Re=2000.0, Ma=0.15, sgs="VREMAN") + "CHECK_INTERVAL 3\nPRECISION MIXED\n")
    s2 = m.ic_tgv(0.15, 1.4)
    inst2 = Instance(m, p2, *s2, layout=lay)
    own, shared = inst2.bytes()
    assert shared == lay.bytes() and own < shared
    c = b.attach(inst2)
    lay.close()  # the solves keep the layout alive
    b.step(6)
    assert_same(b.state(a), alone(1, prec, 6)[0])
    g = IdealGasSolver(m, p2)
    g.set_state(*s2)
    g.init_complete()
    g.advance_solution(6)
    assert_same(b.state(c), g.get_state())
    g.close()
    cnt = b.counts()
    assert cnt["device_bytes"] == 2 * own + shared  # the layout is counted once
    b.close()


@pytest.mark.parametrize("precision", [CG_PREC_FP64, CG_PREC_MIXED])
def test_batch_with_mixed_sgs_models_matches_solo_runs(precision):
    # NONE, VREMAN, SIGMA and DSM_LOCAL solves in one batch: the Sigma and DSM
    # passes run for the batch and leave the other instances alone
    prec = "MIXED" if precision == CG_PREC_MIXED else "FP64"
    models = ["SIGMA", "VREMAN", "NONE", "DSM_LOCAL"]

    def make(i):
        n, seed, ic, Re, Ma = CASES[i]
        m = Mesh.jittered(n, 0.3, seed=seed)
        text = tgv_params(Re=Re, Ma=Ma, sgs=models[i])
        return m, parse_params(text + f"CHECK_INTERVAL 3\nPRECISION {prec}\n"), m.ic_tgv(Ma, 1.4)

    b = Batch(precision)
    ids = []
    for i in range(len(models)):
        m, p, st = make(i)
        ids.append(b.add(m, p, *st))
    b.step(5)
    for i, iid in enumerate(ids):
        m, p, st = make(i)
        g = IdealGasSolver(m, p)
        g.set_state(*st)
        g.init_complete()
        g.advance_solution(5)
        assert_same(b.state(iid), g.get_state())
        assert b.diag(iid).ke == g.report().ke
        g.close()
    b.close()
