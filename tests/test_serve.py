"""Scheduler of the imaging service on the CPU: a fake executor stands in for
the GPU, so admission, memory accounting, cancellation and fault isolation are
tested without a device."""
import pytest

from rick_amd import serve


class FakeExecutor:
    """Jobs take `dur[key]` seconds of fake-clock time and `mem[key]` bytes."""

    def __init__(self, clock, dur, mem, fail_keys=()):
        self.clock, self.dur, self.mem, self.fail = clock, dur, mem, set(fail_keys)
        self.started = []

    def footprint(self, key):
        return self.mem[key]

    def start(self, req):
        if req.key in self.fail:
            raise ValueError("bad job")
        self.started.append(req.rid)
        return (req.key, self.clock.t + self.dur[req.key])

    def done(self, handle):
        return self.clock.t >= handle[1]

    def finish(self, handle, keep_result=False):
        return ("image", handle[0]) if keep_result else None


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001          # time advances as the scheduler polls
        return self.t


@pytest.fixture(autouse=True)
def no_sleep(monkeypatch):
    monkeypatch.setattr(serve.time, "sleep", lambda s: None)


def run(reqs, dur, mem, budget, slots, fail=()):
    clock = FakeClock()
    ex = FakeExecutor(clock, dur, mem, fail)
    srv = serve.Server(ex, budget, slots, clock=clock)
    srv.run(reqs, keep_results=True)
    return srv, ex


def test_all_complete_and_results_kept():
    reqs = [serve.Request(i, i % 3, 0.01 * i) for i in range(20)]
    srv, ex = run(reqs, {0: 0.05, 1: 0.02, 2: 0.01}, {0: 10, 1: 20, 2: 30}, 100, 4)
    assert all(r.status == "done" for r in reqs)
    assert all(r.result == ("image", r.key) for r in reqs)
    assert all(r.finish >= r.start >= r.arrival for r in reqs)
    assert srv.used == 0 and not srv.running


def test_oversized_request_rejected_others_unaffected():
    reqs = [serve.Request(0, 0, 0.0), serve.Request(1, 1, 0.0), serve.Request(2, 0, 0.0)]
    run(reqs, {0: 0.01, 1: 0.01}, {0: 10, 1: 1000}, 100, 2)
    assert [r.status for r in reqs] == ["done", "rejected", "done"]
    assert "budget" in reqs[1].error


def test_memory_budget_limits_concurrency_and_keeps_arrival_order():
    reqs = [serve.Request(i, 0, 0.0) for i in range(6)]
    srv, ex = run(reqs, {0: 0.05}, {0: 40}, 100, 8)
    assert ex.started == list(range(6))               # FIFO
    # at most two 40-byte jobs fit in 100 bytes at once
    starts = sorted(r.start for r in reqs)
    finishes = sorted(r.finish for r in reqs)
    for i, s in enumerate(starts):
        running = sum(1 for r in reqs if r.start <= s < r.finish)
        assert running <= 2


def test_slot_limit():
    reqs = [serve.Request(i, 0, 0.0) for i in range(9)]
    run(reqs, {0: 0.02}, {0: 1}, 10**9, 3)
    for r in reqs:
        assert sum(1 for o in reqs if o.start <= r.start < o.finish) <= 3


def test_timeout_cancels_queued_request_only():
    reqs = [serve.Request(0, 0, 0.0), serve.Request(1, 0, 0.0, timeout=0.01),
            serve.Request(2, 0, 0.0)]
    run(reqs, {0: 0.1}, {0: 1}, 10**9, 1)
    assert [r.status for r in reqs] == ["done", "cancelled", "done"]


def test_failed_launch_isolated():
    reqs = [serve.Request(i, i % 2, 0.0) for i in range(6)]
    run(reqs, {0: 0.01, 1: 0.01}, {0: 1, 1: 1}, 10**9, 2, fail=(1,))
    assert [r.status for r in reqs] == ["done", "failed"] * 3
    stats = serve.latency_stats(reqs)
    assert stats["completed"] == 3 and stats["failed"] == 3


def test_footprint_formula_grows_with_grid():
    from rick_amd import FP32, FP64, synth
    small = synth.CONFIGS["small"]
    lofar = synth.CONFIGS["lofar"]
    assert serve.grid_footprint(lofar, FP64) > 8.5e9            # the 8.59 GB FP64 grid
    assert serve.grid_footprint(lofar, FP32) < serve.grid_footprint(lofar, FP64)
    assert serve.grid_footprint(small, FP32) < serve.grid_footprint(lofar, FP32)


def test_poisson_requests_rate():
    reqs = serve.poisson_requests(3, [0.5, 0.3, 0.2], rate=50.0, duration=20.0, seed=1)
    assert 850 < len(reqs) < 1150
    arr = [r.arrival for r in reqs]
    assert arr == sorted(arr) and arr[-1] <= 20.0
