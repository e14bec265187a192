"""Serving loop: per-view depth requests arriving over time, continuously
batched on one device.

The engine (``MvsContext`` or any object with ``submit(view) -> ticket``,
``serve_step() -> (n_active, n_queued)`` and ``poll(ticket, ref) -> result |
None``) admits a request when its HBM footprint fits the context's budget,
adds it to the running batch at the next half-step boundary and keeps its
image, hypotheses and view-selection state resident until it finishes.  One
``serve_step`` is one batched half-step launch over every admitted request,
each at its own iteration.  Requests that arrive while others run join as
slots free up; a request's result does not depend on its co-runners (per-view
counter-based random streams, tiles independent within a half-step).

With ``geometric=True`` a request is a view's full solve: its photometric
pass is submitted on arrival (no filter, as the reference's photometric call,
colmap_mvs_commands.py:39-55), and its geometric-consistency pass with the
filter (colmap_mvs_commands.py:58-82) as soon as the photometric passes of the
view and of all its sources have been submitted; the engine starts it once
they have finished.  The request completes with the geometric result.

``run_schedule`` replays an arrival schedule against wall-clock time and
reports p50/p99 latency (arrival -> result available) and views/s.
"""
import math
import time

import numpy as np


def poisson_arrivals(n, rate_per_s, seed=6, t0=0.0):
    """Arrival times of n requests with exponential inter-arrival gaps."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, n)
    return list(t0 + np.cumsum(gaps))


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=np.float64), q)) if values else math.nan


class ServeLoop:
    def __init__(self, engine, clock=time.perf_counter, want=("depth",), on_complete=None,
                 geometric=False):
        self.engine = engine
        self.clock = clock
        self.want = want
        self.on_complete = on_complete   # f(ticket, view, result) when a request finishes
        self.geometric = geometric
        self.pending = {}      # request ticket -> (view, arrival time)
        self.done = {}         # request ticket -> (result, arrival, completion)
        self.photo_submitted = set()     # image ids whose photometric pass was submitted
        self.awaiting_geo = {}           # request ticket -> view (geometric pass not yet submitted)
        self.engine_ticket = {}          # request ticket -> engine ticket it completes with
        self.inner = {}                  # engine ticket of a photometric pass -> ref (two-pass mode)

    def submit(self, view, arrival=None):
        """Submit a request; returns its ticket (the engine ticket of its
        first pass)."""
        arr = self.clock() if arrival is None else arrival
        if not self.geometric:
            t = self.engine.submit(view["ref"], view["srcs"], view["depth_range"])
            self.pending[t] = (view, arr)
            self.engine_ticket[t] = t
            return t
        t = self.engine.submit(view["ref"], view["srcs"], view["depth_range"], pass_=0, filter=False)
        self.inner[t] = view["ref"]
        self.photo_submitted.add(int(view["ref"]))
        self.pending[t] = (view, arr)
        self.awaiting_geo[t] = view
        self._release_geometric()
        return t

    def _release_geometric(self):
        for t, view in list(self.awaiting_geo.items()):
            if all(int(s) in self.photo_submitted for s in view["srcs"]):
                g = self.engine.submit(view["ref"], view["srcs"], view["depth_range"], pass_=1,
                                       filter=True)
                self.engine_ticket[t] = g
                del self.awaiting_geo[t]

    def step(self):
        """One batched half-step; collects finished requests."""
        n_active, n_queued = self.engine.serve_step()
        now = self.clock()
        for e, ref in list(self.inner.items()):          # photometric passes: release
            if self.engine.poll(e, ref, ()) is not None:
                del self.inner[e]
        for t in list(self.pending):
            e = self.engine_ticket.get(t)
            if e is None or e in self.inner:
                continue
            view, arr = self.pending[t]
            res = self.engine.poll(e, view["ref"], self.want)
            if res is not None:
                self.done[t] = (res, arr, now)
                del self.pending[t]
                if self.on_complete is not None:
                    self.on_complete(t, view, res)
        return n_active, n_queued

    def _check_progress(self, n_active, n_queued):
        # the engine has nothing left, yet requests wait: their geometric
        # pass needs photometric results of views that were never submitted
        if n_active == 0 and n_queued == 0 and self.pending:
            raise RuntimeError(f"{len(self.pending)} requests wait for the photometric pass of views "
                               "that were not submitted")

    def drain(self, max_steps=10**7):
        k = 0
        while self.pending and k < max_steps:
            self._check_progress(*self.step())
            k += 1
        if self.pending:
            raise RuntimeError(f"{len(self.pending)} requests still pending after {max_steps} steps")

    def run_schedule(self, arrivals, views, idle_sleep=0.0005, max_steps=10**7):
        """arrivals[i] (seconds from start) is when views[i] is submitted."""
        order = np.argsort(arrivals)
        start = self.clock()
        nxt = 0
        tickets = [None] * len(views)
        steps = 0
        idle = 0       # consecutive steps that ended with nothing on the device
        while nxt < len(order) or self.pending:
            now = self.clock() - start
            while nxt < len(order) and arrivals[order[nxt]] <= now:
                i = int(order[nxt])
                tickets[i] = self.submit(views[i], arrival=start + arrivals[i])
                nxt += 1
                idle = 0
            if self.pending and idle < 2:
                na, nq = self.step()
                if nxt >= len(order):
                    self._check_progress(na, nq)
                steps += 1
                if steps > max_steps:
                    raise RuntimeError("serving did not drain")
                # two idle steps in a row: what waits needs a request not yet
                # arrived (the first may have retired what the queue waited on)
                idle = idle + 1 if na == 0 else 0
            elif nxt < len(order):
                time.sleep(min(idle_sleep, max(0.0, arrivals[order[nxt]] - now)))
            else:
                idle = 0
        wall = self.clock() - start
        lat = [self.done[t][2] - self.done[t][1] for t in tickets]
        return dict(tickets=tickets, latency_s=lat, wall_s=wall, steps=steps,
                    p50_ms=1e3 * percentile(lat, 50), p99_ms=1e3 * percentile(lat, 99),
                    views_per_s=len(views) / wall if wall > 0 else math.nan)

    def result(self, ticket):
        return self.done[ticket][0]
