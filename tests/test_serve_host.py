"""The served path's scheduling, on the host: SimServer driving a stand-in
executor (no GPU).  Admission against the HBM budget, rejection of requests
that can never fit (serve() must end), cancellation of queued and running
requests, and isolation of a failed network: the others' results are the
same as in a run without the failure."""
import hashlib

import pytest

from brainpy_amd.dyn.serve import (BYTES_PER_NEURON, CANCELLED, DONE, FAILED, REJECTED, Request, SimServer,
                                   _Pool)


class FakeExecutor:
    """Host stand-in for DeviceExecutor: a network's result is a hash chain
    of its own seed and local step count, so it depends on nothing else in
    the batch (like the device step, which never reads another network)."""
    max_slots = 64

    def __init__(self, initial, max_capacity):
        self.capacity, self.max_capacity = initial, max_capacity
        self.state, self.bad = {}, set()
        self.peak_neurons = 0
        self.grows = 0

    def ensure(self, need):
        if need > self.max_capacity:
            return False
        while self.capacity < need:
            self.capacity = min(self.max_capacity, max(2 * self.capacity, 1))
            self.grows += 1
        return True

    def init(self, req):
        self.state[req.rid] = [0, hashlib.sha256(str(req.seed).encode()).hexdigest()]

    def changed(self):
        pass

    def step(self, active, step_idx):
        self.peak_neurons = max(self.peak_neurons, sum(r.n for r in active))
        for r in active:
            st = self.state[r.rid]
            assert step_idx - r.t0 == st[0]          # every network advances one local step per step
            st[0] += 1
            st[1] = hashlib.sha256((st[1] + str(st[0])).encode()).hexdigest()

    def read(self, reqs):
        return [(self.state[r.rid][0] * r.n, int(self.state[r.rid][1][:8], 16), None) for r in reqs]

    def unhealthy(self, active):
        return [r for r in active if r.rid in self.bad]

    def sync(self):
        pass


def _server(budget_neurons=10_000, initial=1024, **kw):
    budget = budget_neurons * BYTES_PER_NEURON
    return SimServer(hbm_budget=budget, executor=FakeExecutor(initial, budget // BYTES_PER_NEURON), **kw)


def _reqs(sizes, steps, seed0=100):
    return [Request(k, n, seed0 + k, 0.5, 6.0, 20.0, st, arrival=0.0) for k, (n, st) in enumerate(zip(sizes, steps))]


def test_pool_allocator():
    p = _Pool(100)
    a, b = p.alloc(30), p.alloc(50)
    assert (a, b) == (0, 30) and p.alloc(30) is None
    p.release(a, 30)
    assert p.alloc(25) == 0
    p.grow(200)
    assert p.alloc(110) == 80


def test_budget_binds_and_pools_grow():
    srv = _server(budget_neurons=10_000, initial=1000)
    reqs = _reqs([4000, 4000, 4000, 3000, 2000], [30, 20, 10, 25, 5])
    done = {r.rid: r for r in srv.serve(reqs)}
    assert all(r.status == DONE for r in done.values()) and len(done) == 5
    assert srv.ex.peak_neurons <= 10_000          # never more resident state than the budget
    assert srv.ex.grows > 0 and srv.capacity <= 10_000
    for r in reqs:
        assert r.steps_done == r.steps and r.events == r.steps * r.n


def test_never_fitting_requests_are_rejected_and_serve_returns():
    srv = _server(budget_neurons=5000)
    reqs = _reqs([6000, 3000, 10**9, 1], [10, 10, 10, 10])
    done = {r.rid: r for r in srv.serve(reqs, time_limit=30)}
    assert done[0].status == REJECTED and "budget" in done[0].reason
    assert done[2].status == REJECTED
    assert done[3].status == REJECTED         # one neuron: no E/I split
    assert done[1].status == DONE and done[1].steps_done == 10


def test_time_limit_ends_serve():
    srv = _server(budget_neurons=5000)
    reqs = _reqs([3000], [10**9])
    t = srv.serve(reqs, time_limit=0.2)
    assert t == [] and srv.active and srv.active[0].status == "active"


def test_cancel_queued_and_running():
    srv = _server(budget_neurons=5000)
    reqs = _reqs([3000, 3000, 1000], [50, 50, 50])   # 1 waits behind 0 (budget)
    done = {r.rid: r for r in srv.serve(reqs, cancel_at={0: 10, 1: 5})}
    assert done[0].status == CANCELLED and done[0].steps_done == 10 and done[0].events == 10 * 3000
    assert done[1].status == CANCELLED and done[1].steps_done == 0 and done[1].admitted is None
    assert done[2].status == DONE and done[2].steps_done == 50
    assert not srv.cancel(2) and not srv.cancel(99)


@pytest.mark.parametrize("bad", [0, 2])
def test_failed_network_is_isolated(bad):
    sizes, steps = [2000, 1500, 1000, 500], [200, 150, 300, 100]
    ref = {r.rid: r for r in _server(health_every=16).serve(_reqs(sizes, steps))}
    srv = _server(health_every=16)
    reqs = _reqs(sizes, steps)
    srv.ex.bad.add(bad)                      # its state goes non-finite
    done = {r.rid: r for r in srv.serve(reqs)}
    assert done[bad].status == FAILED and "non-finite" in done[bad].reason
    assert done[bad].steps_done == 16
    for k in range(4):
        if k != bad:
            assert done[k].status == DONE
            assert (done[k].events, done[k].spikes) == (ref[k].events, ref[k].spikes)


def test_fifo_mode_runs_one_at_a_time():
    srv = _server(budget_neurons=100_000, max_active=1)
    reqs = _reqs([1000, 2000, 3000], [5, 5, 5])
    done = srv.serve(reqs)
    assert [r.rid for r in done] == [0, 1, 2]
    assert srv.ex.peak_neurons == 3000 and srv.step_idx == 15
