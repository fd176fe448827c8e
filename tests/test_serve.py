"""Scheduler of the imaging service on the CPU: a fake executor stands in for
the GPU, so admission, memory accounting, cancellation and fault isolation are
tested without a device."""
import pytest

from rick_amd import serve


class FakeExecutor:
    """Jobs take `dur[key]` seconds of fake-clock time and `mem[key]` bytes;
    finished jobs' buffers go to an idle pool keyed by job, as on the GPU."""

    def __init__(self, clock, dur, mem, fail_keys=()):
        self.clock, self.dur, self.mem, self.fail = clock, dur, mem, set(fail_keys)
        self.started = []
        self.pool = serve.PlanPool()
        self.running_bytes = 0
        self.peak = 0                     # most device bytes held at once
        self.allocations = 0

    def footprint(self, key):
        return self.mem[key]

    def pooled_bytes(self):
        return self.pool.nbytes

    def reusable_bytes(self, key):
        return self.pool.reusable(key)

    def evict(self, nbytes, keep_key=None):
        return self.pool.evict(nbytes, keep_key)

    def start(self, req):
        if req.key in self.fail:
            raise ValueError("bad job")
        if self.pool.take(req.key) is None:
            self.allocations += 1
        self.running_bytes += self.mem[req.key]
        self.peak = max(self.peak, self.running_bytes + self.pool.nbytes)
        self.started.append(req.rid)
        return (req.key, self.clock.t + self.dur[req.key])

    def done(self, handle):
        return self.clock.t >= handle[1]

    def finish(self, handle, keep_result=False):
        key = handle[0]
        self.running_bytes -= self.mem[key]
        self.pool.put(key, self.mem[key], ("buffers", key))
        return ("image", key) if keep_result else None


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


def test_running_plus_pooled_bytes_stay_within_budget():
    """Six job kinds of 30-60 bytes against a 100-byte budget: finished jobs'
    buffers stay pooled, so admitting new kinds must free idle entries."""
    mem = {k: 30 + 6 * k for k in range(6)}
    dur = {k: 0.01 for k in range(6)}
    reqs = [serve.Request(i, (i * 7) % 6, 0.003 * i) for i in range(60)]
    srv, ex = run(reqs, dur, mem, 100, 3)
    assert all(r.status == "done" for r in reqs)
    assert ex.peak <= 100
    assert ex.pool.nbytes <= 100


def test_pool_reuses_same_job_without_allocating():
    reqs = [serve.Request(i, 0, 0.02 * i) for i in range(10)]
    srv, ex = run(reqs, {0: 0.005}, {0: 40}, 100, 2)
    assert all(r.status == "done" for r in reqs)
    assert ex.allocations == 1


def test_plan_pool_evicts_oldest_and_spares_the_reused_entry():
    p = serve.PlanPool()
    p.put("a", 10, 1)
    p.put("b", 20, 2)
    p.put("a", 10, 3)
    assert p.nbytes == 40 and p.reusable("a") == 10 and p.reusable("c") == 0
    assert p.evict(25, keep_sig="a") == 30          # oldest "a", then "b"
    assert p.take("a") == 3 and p.nbytes == 0


def test_gpu_executor_returns_slot_when_acquire_fails():
    """A failure while taking a job's plan and buffers (e.g. device out of
    memory) must hand the stream slot back: later jobs still start."""
    ex = serve.GpuExecutor.__new__(serve.GpuExecutor)
    ex.torch = None
    ex.free_streams = [0, 1]
    ex.pool = serve.PlanPool()

    def boom(key):
        raise MemoryError("hipMalloc failed")

    ex._acquire = boom
    ex._sig = lambda key: key
    ex.footprint = lambda key: 1
    for _ in range(3):
        with pytest.raises(MemoryError):
            ex.start(serve.Request(0, 0, 0.0))
    assert sorted(ex.free_streams) == [0, 1]


def test_progress_lines_during_a_run():
    reqs = [serve.Request(i, 0, 0.01 * i) for i in range(50)]
    clock = FakeClock()
    ex = FakeExecutor(clock, {0: 0.02}, {0: 1}, ())
    srv = serve.Server(ex, 10**9, 4, clock=clock)
    lines = []
    srv.run(reqs, progress=0.1, log=lines.append)
    assert all(r.status == "done" for r in reqs)
    # about one line per 0.1 s of the ~0.5 s run, each with the counts
    assert 3 <= len(lines) <= 8
    assert all(line.startswith("[serve] t=") and "finished" in line for line in lines)
    clock2 = FakeClock()
    quiet = serve.Server(FakeExecutor(clock2, {0: 0.02}, {0: 1}, ()), 10**9, 4, clock=clock2)
    silent = []
    quiet.run([serve.Request(i, 0, 0.01 * i) for i in range(10)], log=silent.append)
    assert silent == []
