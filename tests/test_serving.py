"""The serving loop's scheduling on the CPU, with a fake device backend: admission by footprint,
first come first served, chunks that never overshoot a request, continuous batching, latency
bookkeeping.  (The GPU version of the loop is exercised in test_gpu_serving.py.)"""
import pytest

from conftest import kb

serving = kb.serving


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 0.0)


class FakeBackend:
    """Each step of a chunk costs step_cost seconds of fake time; footprints from the request."""

    def __init__(self, clock, step_cost=1e-3):
        self.clock = clock
        self.step_cost = step_cost
        self.live = set()
        self.calls = []

    def prepare(self, r):
        r.points = 100
        return r

    def footprint(self, prep, r):
        return r.cfg["bytes"]

    def create(self, prep, r):
        self.live.add(r.rid)
        return r.rid, 1e-3

    def step(self, handles, dts, nsteps):
        assert set(handles) <= self.live
        self.calls.append((tuple(handles), nsteps))
        self.clock.t += self.step_cost * nsteps

    def finish(self, h, prep):
        self.live.remove(h)
        return ("state", h)


def make(reqs, budget, chunk=10):
    clock = FakeClock()
    be = FakeBackend(clock)
    srv = serving.Server(be, budget, chunk_steps=chunk, clock=clock, sleep=clock.sleep)
    return srv, be, clock


def R(rid, nsteps, arrival, nbytes=10):
    return serving.Request(rid=rid, cfg={"bytes": nbytes}, nsteps=nsteps, arrival=arrival)


def test_chunks_never_overshoot_and_all_finish():
    reqs = [R(0, 25, 0.0), R(1, 7, 0.0), R(2, 40, 0.005)]
    srv, be, clock = make(reqs, budget=100, chunk=10)
    st = srv.run(reqs)
    done = {}
    for handles, n in be.calls:
        for h in handles:
            done[h] = done.get(h, 0) + n
    assert done == {0: 25, 1: 7, 2: 40}
    assert all(r.result == ("state", r.rid) for r in reqs)
    assert st.summary()["requests"] == 3 and not be.live


def test_requests_advance_in_pairs_from_their_own_start():
    """Every request sees the step sequence of a solo run: pairs from its first step, then at most one
    single step at its end (a two-step pass may round differently from two one-step passes)."""
    reqs = [R(0, 25, 0.0), R(1, 7, 0.0), R(2, 40, 0.005), R(3, 13, 0.011)]
    srv, be, clock = make(reqs, budget=100, chunk=7)
    srv.run(reqs)
    for r in reqs:
        seq = [n for h, n in be.calls if r.rid in h]
        assert sum(seq) == r.nsteps
        assert all(n % 2 == 0 for n in seq[:-1]), (r.rid, seq)
        assert seq[-1] % 2 == 0 or seq[-1] == 1, (r.rid, seq)
        # a single step is the request's own last one, taken with no request that still continues
        for h, n in be.calls:
            if n == 1 and r.rid in h:
                assert all(seq2[-1] == 1 for seq2 in ([m for hh, m in be.calls if q in hh] for q in h))


def test_admission_by_memory_is_fifo():
    # budget fits two requests; the third waits for the first to finish
    reqs = [R(0, 10, 0.0, 40), R(1, 30, 0.0, 40), R(2, 10, 0.0, 40)]
    srv, be, clock = make(reqs, budget=80, chunk=10)
    st = srv.run(reqs)
    assert st.max_live == 2 and st.peak_bytes == 80
    assert reqs[2].admitted >= reqs[0].done  # admitted only after request 0 freed its memory
    assert set(be.calls[0][0]) == {0, 1}


def test_continuous_batching_joins_between_chunks():
    # request 1 arrives while request 0 is being stepped and joins at the next chunk boundary
    reqs = [R(0, 50, 0.0), R(1, 10, 0.012)]
    srv, be, clock = make(reqs, budget=100, chunk=10)
    srv.run(reqs)
    assert be.calls[0][0] == (0,)
    assert any(set(h) == {0, 1} for h, _ in be.calls)
    assert reqs[1].done < reqs[0].done


def test_latency_is_arrival_to_completion():
    reqs = [R(0, 20, 0.5)]
    srv, be, clock = make(reqs, budget=100, chunk=10)
    st = srv.run(reqs)
    assert reqs[0].latency == pytest.approx(20 * 1e-3)
    assert st.summary()["p99_ms"] == pytest.approx(20.0)


def test_request_larger_than_budget_is_refused():
    reqs = [R(0, 5, 0.0, 1000)]
    srv, be, clock = make(reqs, budget=100)
    with pytest.raises(ValueError):
        srv.run(reqs)


def test_serving_mix_is_reproducible():
    a = kb.configs.serving_requests(20, 50.0)
    b = kb.configs.serving_requests(20, 50.0)
    assert [(k, t) for k, _, _, t in a] == [(k, t) for k, _, _, t in b]
    assert {k for k, *_ in kb.configs.serving_requests(200, 50.0)} == {"1d", "2d", "3v"}


def test_autotuner_picks_the_fastest_and_caches(tmp_path):
    """The tuner with a fake context: schedules first, then velocities per wavefront at the best
    schedule, then graph replay; the fastest candidate wins, is left set and is cached."""
    times = {(60, 4, 0, False): 0.2, (60, 2, 0, False): 0.15, (30, 4, 0, False): 0.3,
             (60, 2, 32, False): 0.14, (60, 2, 16, False): 0.16, (60, 2, 32, True): 0.145}

    class Ctx:
        sched, vw, graphs = (60, 4), 0, False

        def set_schedule(self, a, b):
            self.sched = (a, b)

        def set_sweep_vw(self, v):
            self.vw = v

        def set_graphs(self, g):
            self.graphs = g

        def state(self):
            return (*self.sched, self.vw, self.graphs)

        def last_steps_ms(self):
            return times[self.state()] * 6

    class Dev:
        def __init__(self, nv):
            self.ks = kb.SolverSet(kb.configs.sod_1d1f1v(nu=nv))
            self.precision = "fp32"

    calls = []
    orig = kb.solver.batch_step
    kb.solver.batch_step = lambda devs, dts, n, ctx=None: calls.append(n)
    try:
        ctx = Ctx()
        devs = [Dev(32), Dev(64)]
        cands = [(60, 4), (60, 2), (30, 4)]
        best, res, taken = kb.autotune.tune(ctx, devs, [1e-3, 1e-3], candidates=cands, cache=tmp_path / "t.json")
        assert best == (60, 2, 32, False) and ctx.state() == best and taken == sum(calls)
        assert set(res) == set(times)
        ctx.set_schedule(30, 4)
        best2, _, taken2 = kb.autotune.tune(ctx, devs, [1e-3, 1e-3], candidates=cands, cache=tmp_path / "t.json")
        assert best2 == best and taken2 == 0 and ctx.state() == best
    finally:
        kb.solver.batch_step = orig


def test_overlap_prepares_arrivals_while_a_chunk_runs():
    """overlap=True: requests that arrive while a chunk runs are prepared during it (host work
    beside the device call), never admitted into the running chunk, and every request still gets
    exactly its steps."""
    import threading
    import time

    class SlowBackend(FakeBackend):
        def __init__(self):
            super().__init__(clock=None)
            self.prepared_during_step = 0
            self.stepping = threading.Event()

        def prepare(self, r):
            if self.stepping.is_set():
                self.prepared_during_step += 1
            return super().prepare(r)

        def step(self, handles, dts, nsteps):
            assert set(handles) <= self.live
            self.calls.append((tuple(handles), nsteps))
            self.stepping.set()
            time.sleep(0.02)
            self.stepping.clear()

    be = SlowBackend()
    reqs = [R(0, 30, 0.0), R(1, 20, 0.005), R(2, 10, 0.012), R(3, 10, 0.03)]
    srv = serving.Server(be, 1000, chunk_steps=10, overlap=True)
    st = srv.run(reqs)
    done = {}
    for handles, n in be.calls:
        for h in handles:
            done[h] = done.get(h, 0) + n
    assert done == {0: 30, 1: 20, 2: 10, 3: 10}
    assert be.prepared_during_step >= 1
    assert st.summary()["requests"] == 4 and not be.live


def test_transfers_overlap_the_chunks_and_never_touch_a_running_batch():
    """overlap with a transfer backend: uploads and downloads run while a chunk steps, never on a
    request inside the running chunk; memory is released only between chunks; every request still
    gets exactly its steps, in pairs from its own start, and its result."""
    import threading
    import time

    class Device(FakeBackend):
        def __init__(self):
            super().__init__(clock=None)
            self.running: set = set()
            self.lock = threading.Lock()
            self.overlapped = 0
            self.violations = []

        def step(self, handles, dts, nsteps):
            assert set(handles) <= self.live
            with self.lock:
                self.running = set(handles)
            self.calls.append((tuple(handles), nsteps))
            time.sleep(0.01)
            with self.lock:
                self.running = set()

        def check(self, what, h):
            with self.lock:
                if self.running:
                    self.overlapped += 1
                if h in self.running:
                    self.violations.append((what, h))

        def create(self, prep, r):
            h, dt = super().create(prep, r)
            self.check("create", h)
            return h, dt

        def download(self, h, prep):
            self.check("download", h)
            return ("state", h)

        def release(self, h):
            with self.lock:
                if self.running:
                    self.violations.append(("release during a chunk", h))
            self.live.remove(h)

    dev = Device()
    reqs = [R(0, 30, 0.0), R(1, 20, 0.004), R(2, 10, 0.012), R(3, 13, 0.02), R(4, 8, 0.021, 60)]
    srv = serving.Server(dev, 100, chunk_steps=6, overlap=True, transfer=dev)
    st = srv.run(reqs)
    assert not dev.violations and dev.overlapped > 0 and not dev.live
    assert all(r.result == ("state", r.rid) for r in reqs) and st.summary()["requests"] == 5
    for r in reqs:
        seq = [n for h, n in dev.calls if r.rid in h]
        assert sum(seq) == r.nsteps
        assert all(n % 2 == 0 for n in seq[:-1]) and (seq[-1] % 2 == 0 or seq[-1] == 1), (r.rid, seq)
    assert st.peak_bytes <= 100
