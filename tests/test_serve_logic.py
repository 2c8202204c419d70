"""Scheduling logic of the serving loop (CPU, with a stand-in batch): every
request completes with its own horizon, admission is first come first served,
the device memory held by admitted solves stays within the budget, time to
solution is completion minus arrival, and the sequential reference applies the
same queueing arithmetic."""
import time
from types import SimpleNamespace

from nextgen_amd.ensemble import Spec
from nextgen_amd.serve import Request, Server, requests_poisson, summarize

BPC = 100.0  # stand-in device bytes per cell


class FakeInstance:
    def __init__(self, req):
        self.req = req

    def bytes(self):
        return int(BPC * self.req.spec.n ** 3), 0


class FakeBatch:
    def __init__(self, precision):
        self.members = {}
        self.next = 0
        self.log = []

    def attach(self, inst):
        i = self.next
        self.next += 1
        self.members[i] = [inst.req, 0]
        self.log.append(("attach", inst.req.rid))
        return i

    def step(self, k):
        time.sleep(0.0002 * k)
        for m in self.members.values():
            m[1] += k

    def diag(self, i):
        req, steps = self.members[i]
        return SimpleNamespace(ke=float(req.rid), enstrophy=float(steps), time=float(steps), step=steps)

    def remove(self, i):
        self.members.pop(i)

    def close(self):
        pass


def run(reqs, budget):
    batches = []

    def factory(p):
        batches.append(FakeBatch(p))
        return batches[-1]

    srv = Server(None, hbm_budget=budget, max_chunk=8, bytes_per_cell=BPC, batch_factory=factory,
                 prepare=FakeInstance)
    res, info = srv.run(reqs)
    return res, info, srv, batches[0]


def test_every_request_completes_with_its_horizon():
    reqs = [Request(i, 0.002 * i, Spec(n=10 + i, mesh_seed=1, ic="tgv", Re=1e3, Ma=0.1, ic_seed=1, steps=5 + 7 * i))
            for i in range(6)]
    res, info, srv, b = run(reqs, budget=1e12)
    assert sorted(r.rid for r in res) == list(range(6))
    for r in res:
        assert r.enstrophy == reqs[r.rid].spec.steps  # stepped exactly its horizon
        assert r.done >= r.start >= r.arrival - 1e-9
        assert abs(r.tts - (r.done - r.arrival)) < 1e-12
    # admitted in arrival order
    assert [rid for _, rid in b.log] == list(range(6))


def test_hbm_budget_is_respected():
    # every request holds 100 * 20^3 = 800 kB; a budget of 2.5 MB admits at most 3 at a time
    reqs = [Request(i, 0.0, Spec(n=20, mesh_seed=1, ic="tgv", Re=1e3, Ma=0.1, ic_seed=1, steps=40))
            for i in range(7)]
    res, info, srv, b = run(reqs, budget=2.5e6)
    assert len(res) == 7
    assert srv.peak_bytes <= 2.5e6
    # the batch never held more than three members at once
    starts = sorted((r.start, r.done) for r in res)
    for t, _ in starts:
        assert sum(1 for s, d in starts if s <= t < d) <= 3


def test_poisson_requests_and_summary():
    reqs = requests_poisson(2.0, 50, [27, 50, 126], seed=3)
    gaps = [b.arrival - a.arrival for a, b in zip(reqs, reqs[1:])]
    assert all(g > 0 for g in gaps)
    assert 0.2 < sum(gaps) / len(gaps) < 1.0  # mean 1/rate = 0.5
    assert {r.spec.n for r in reqs} <= {27, 50, 126}
    assert all(200 <= r.spec.steps <= 2000 for r in reqs)
    from nextgen_amd.serve import Result
    res = [Result(i, float(i), float(i), float(i) + 2.0, 1, 1, 0.0, 0.0, 0.0) for i in range(10)]
    s = summarize(res)
    assert s["tts_p50_s"] == 2.0 and s["solves_per_s"] == 10 / 11.0


def test_server_counts_net_steps_after_a_rollback():
    # a solve that a NaN rollback takes back 3 steps is stepped on to its
    # horizon in net steps, as the reference run loop does
    class RollbackBatch(FakeBatch):
        rolled = False

        def step(self, k):
            super().step(k)
            if not self.rolled:
                for m in self.members.values():
                    if m[1] >= 6:
                        m[1] -= 3
                        RollbackBatch.rolled = True
                        break

    reqs = [Request(0, 0.0, Spec(n=10, mesh_seed=1, ic="tgv", Re=1e3, Ma=0.1, ic_seed=1, steps=20))]
    res, info = Server(None, max_chunk=4, batch_factory=RollbackBatch, prepare=FakeInstance).run(reqs)
    assert RollbackBatch.rolled
    assert res[0].enstrophy == 20  # the device's own step count at retirement
    assert info["batch_steps"] == 23
