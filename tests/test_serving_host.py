"""Serving loop logic on the CPU with a fake engine and a fake clock."""
import numpy as np

from vissatsatellitestereo_amd.serving import ServeLoop, poisson_arrivals, percentile


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeEngine:
    """Each request needs `need` steps; at most `slots` run at once (admission)."""

    def __init__(self, clock, need=3, slots=2, step_s=0.01):
        self.clock, self.need, self.slots, self.step_s = clock, need, slots, step_s
        self.queue, self.active, self.finished, self.next = [], {}, {}, 1
        self.max_active = 0

    def submit(self, ref, srcs, depth_range):
        t = self.next
        self.next += 1
        self.queue.append((t, ref))
        return t

    def serve_step(self):
        while self.queue and len(self.active) < self.slots:
            t, ref = self.queue.pop(0)
            self.active[t] = [ref, 0]
        self.max_active = max(self.max_active, len(self.active))
        self.clock.t += self.step_s
        for t in list(self.active):
            self.active[t][1] += 1
            if self.active[t][1] >= self.need:
                self.finished[t] = {"depth": np.full((2, 2), self.active[t][0], np.float32)}
                del self.active[t]
        return len(self.active), len(self.queue)

    def poll(self, ticket, ref, want):
        return self.finished.pop(ticket, None)


def test_poisson_arrivals_seeded_and_rate():
    a = poisson_arrivals(2000, 10.0, seed=6)
    b = poisson_arrivals(2000, 10.0, seed=6)
    assert a == b and all(x < y for x, y in zip(a, a[1:]))
    rate = len(a) / a[-1]
    assert 9.0 < rate < 11.0


def test_drain_respects_slots_and_returns_each_result():
    clock = FakeClock()
    eng = FakeEngine(clock, need=3, slots=2)
    loop = ServeLoop(eng, clock=clock)
    views = [dict(ref=r, srcs=[0], depth_range=(0, 1)) for r in range(1, 6)]
    ts = [loop.submit(v) for v in views]
    loop.drain()
    assert eng.max_active == 2
    for t, v in zip(ts, views):
        assert np.all(loop.result(t)["depth"] == v["ref"])
    # 5 requests, 2 slots, 3 steps each -> finish at steps 3,3,6,6,9
    lat = sorted(loop.done[t][2] - loop.done[t][1] for t in ts)
    assert np.allclose(lat, [0.03, 0.03, 0.06, 0.06, 0.09])


def test_run_schedule_latency_from_arrival():
    clock = FakeClock()
    eng = FakeEngine(clock, need=2, slots=4, step_s=0.05)
    loop = ServeLoop(eng, clock=clock)
    views = [dict(ref=r, srcs=[0], depth_range=(0, 1)) for r in range(1, 4)]
    out = loop.run_schedule([0.0, 0.0, 0.0], views)
    assert np.allclose(out["latency_s"], [0.1, 0.1, 0.1])
    assert out["p50_ms"] == percentile([100.0] * 3, 50)


class WaitingEngine:
    """Requests need `need` steps; a request of ref 1 cannot start before a
    request of ref 2 has been submitted (as a geometric pass waits for a
    source's photometric pass).  A step with nothing to run takes no time."""

    def __init__(self, clock, need=2, step_s=0.01):
        self.clock, self.need, self.step_s = clock, need, step_s
        self.queue, self.active, self.finished, self.next = [], {}, {}, 1
        self.submitted_refs = set()
        self.idle_steps = 0

    def submit(self, ref, srcs, depth_range):
        t = self.next
        self.next += 1
        self.queue.append((t, ref))
        self.submitted_refs.add(ref)
        return t

    def serve_step(self):
        for q in list(self.queue):
            if q[1] != 1 or 2 in self.submitted_refs:
                self.queue.remove(q)
                self.active[q[0]] = [q[1], 0]
        if not self.active:
            self.idle_steps += 1
            return 0, len(self.queue)
        self.clock.t += self.step_s
        for t in list(self.active):
            self.active[t][1] += 1
            if self.active[t][1] >= self.need:
                self.finished[t] = {"depth": np.full((2, 2), self.active[t][0], np.float32)}
                del self.active[t]
        return len(self.active), len(self.queue)

    def poll(self, ticket, ref, want):
        return self.finished.pop(ticket, None)


def test_run_schedule_waits_for_an_arrival_without_spinning(monkeypatch):
    """A queued request that needs a later arrival: the loop sleeps until
    that arrival instead of calling serve_step over and over."""
    from vissatsatellitestereo_amd import serving
    clock = FakeClock()
    monkeypatch.setattr(serving.time, "sleep", lambda s: setattr(clock, "t", clock.t + max(s, 1e-4)))
    eng = WaitingEngine(clock)
    loop = ServeLoop(eng, clock=clock)
    views = [dict(ref=1, srcs=[2], depth_range=(0, 1)), dict(ref=2, srcs=[1], depth_range=(0, 1))]
    out = loop.run_schedule([0.0, 0.5], views, max_steps=100)
    # request 1 starts when request 2 arrives at 0.5 s; both finish two steps later
    assert np.allclose(out["latency_s"], [0.52, 0.02])
    assert eng.idle_steps <= 2
    assert out["steps"] <= 4
