"""Batched and served scenes: many independent coupled-step contexts on one
GPU, each with its own HIP stream, advanced concurrently by host worker
threads.

A scene's step is a chain of small, latency-bound launches (the Krylov
loops) that leaves most of the GPU idle; independent scenes on their own
streams fill it, and each scene still advances at its own pace (its solves
converge in their own number of iterations).  esf_step is one C call per
step that releases the GIL (ctypes), so a Python thread per active scene is
enough to keep its stream fed.

  BatchRunner : a fixed set of scenes, K timed steps each, run sequentially
                (one scene at a time) or concurrently (a worker pool)
  Server      : requests (scene, K steps) arriving over time, admitted while
                a worker slot and the HBM budget allow, each one created,
                stepped K times and released by its own worker; reports per-step
                and completion latencies and aggregate throughput

Scenes follow the batched / served workloads of the benchmark: cubic
unit-cube grids of N^3 cells, 1-3 corotational spheres with radius in
[0.08, 0.2] and E log-uniform in [1e3, 1e5] (nu = 0.3), a fluid pool filling
20-50 % of the domain, dt = 1/240, mixed precision at 1e-8.
"""
import threading
import time
from collections import deque

import numpy as np

from . import _abi
from .cfg import random_spheres, scene_cfg
from .integrator import CoupledIntegrator3D, step_batch


def random_scene(rng, n, eps=1e-8):
    """cfg text of one scene of the batched / served workloads on n^3 cells"""
    k = int(rng.integers(1, 4))
    spheres = random_spheres(rng, k, (0.08, 0.2), (3.0, 5.0))
    fill = float(rng.uniform(0.2, 0.5))
    return scene_cfg(n, spheres, fill=fill, eps=eps, max_iterations=20000)


def step_scene(g, steps, lat=None):
    """`steps` coupled steps of one context; per-step wall latency (the step
    call returns when the step's work has completed on the GPU)"""
    for _ in range(steps):
        t0 = time.perf_counter()
        r = g.stepImplicit(-1.0)
        if lat is not None:
            lat.append(time.perf_counter() - t0)
        if r["status"] < 0:
            raise RuntimeError(f"step failed: {r}")


class BatchRunner:
    """A fixed set of scenes (contexts created up front), stepped either one
    scene at a time or concurrently by `workers` threads."""

    def __init__(self, texts, ns, precision=_abi.ESF_MIXED, warmup=2):
        self.ns = list(ns)
        self.scenes = [CoupledIntegrator3D.from_cfg_text(t, precision) for t in texts]
        for g in self.scenes:
            step_scene(g, warmup)

    def cells(self):
        return [n ** 3 for n in self.ns]

    def run_sequential(self, steps):
        lat = []
        t0 = time.perf_counter()
        for g in self.scenes:
            step_scene(g, steps, lat)
        return time.perf_counter() - t0, lat

    def run_batched(self, steps):
        """`steps` substeps of every scene, each substep one esf_step_batch
        call over all the scenes; the latency of a batch step is that of
        each of its scenes' steps"""
        lat = []
        t0 = time.perf_counter()
        for _ in range(steps):
            t1 = time.perf_counter()
            for r in step_batch(self.scenes):
                if r["status"] < 0:
                    raise RuntimeError(f"step failed: {r}")
            lat.append(time.perf_counter() - t1)
        return time.perf_counter() - t0, lat

    def run_concurrent(self, steps, workers):
        order = deque(range(len(self.scenes)))
        lock = threading.Lock()
        lats = [[] for _ in range(workers)]
        errors = []

        def work(w):
            try:
                while True:
                    with lock:
                        if not order:
                            return
                        i = order.popleft()
                    step_scene(self.scenes[i], steps, lats[w])
            except Exception as e:  # reported by the caller
                errors.append(e)

        threads = [threading.Thread(target=work, args=(w,)) for w in range(workers)]
        t0 = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return wall, [x for l in lats for x in l]

    def close(self):
        for g in self.scenes:
            g.close()
        self.scenes = []


class Server:
    """Continuous batching of scene requests.  A request (arrival time, N,
    K steps, cfg text) is admitted when fewer than `max_active` scenes are
    running and its estimated device memory fits the remaining HBM budget;
    its worker creates the context, takes K steps, records the per-step
    latencies, releases the context and its memory.  Requests that arrive
    while others run are admitted as slots and memory free up, and their
    steps overlap on the GPU with everyone else's."""

    def __init__(self, budget_bytes, max_active, precision=_abi.ESF_MIXED,
                 bytes_per_node=1200.0, on_done=None):
        self.budget = budget_bytes
        self.max_active = max_active
        self.precision = precision
        # on_done(request, context): called by the request's worker after its
        # K steps, before the context is released (e.g. to read results)
        self.on_done = on_done
        # admission estimate per grid node, raised to what contexts really hold
        self.bytes_per_node = bytes_per_node
        self.lock = threading.Lock()

    def estimate(self, n):
        return int(self.bytes_per_node * (n + 1) ** 3)

    def run(self, requests):
        """requests: list of dicts with keys t (arrival, s), n, k, text.
        Returns per-request records and the makespan."""
        reqs = sorted(requests, key=lambda r: r["t"])
        pending = deque()
        nxt = 0
        used = 0
        active = {}
        done = []
        peak_active = 0
        peak_bytes = 0
        t_start = time.perf_counter()

        def worker(req, est):
            rec = {"id": req["id"], "n": req["n"], "k": req["k"], "arrival": req["t"],
                   "lat": [], "error": None}
            rec["start"] = time.perf_counter() - t_start
            g = None
            try:
                g = CoupledIntegrator3D.from_cfg_text(req["text"], self.precision)
                rec["bytes"] = g.device_bytes()
                with self.lock:
                    self.bytes_per_node = max(self.bytes_per_node,
                                              rec["bytes"] / float((req["n"] + 1) ** 3))
                step_scene(g, req["k"], rec["lat"])
                if self.on_done is not None:
                    self.on_done(req, g)
            except Exception as e:  # kept in the record, reported by the caller
                rec["error"] = repr(e)
            finally:
                if g is not None:
                    g.close()
            rec["end"] = time.perf_counter() - t_start
            rec["est"] = est
            return rec

        results = {}

        def run_one(req, est):
            results[req["id"]] = worker(req, est)

        while nxt < len(reqs) or pending or active:
            now = time.perf_counter() - t_start
            while nxt < len(reqs) and reqs[nxt]["t"] <= now:
                pending.append(reqs[nxt])
                nxt += 1
            # reap finished workers
            for rid in [rid for rid, (th, est) in active.items() if not th.is_alive()]:
                th, est = active.pop(rid)
                th.join()
                used -= est
                done.append(results[rid])
            # admit in arrival order while a slot and the budget allow
            while pending and len(active) < self.max_active:
                est = self.estimate(pending[0]["n"])
                if used + est > self.budget and active:
                    break
                req = pending.popleft()
                used += est
                th = threading.Thread(target=run_one, args=(req, est))
                active[req["id"]] = (th, est)
                th.start()
            peak_active = max(peak_active, len(active))
            peak_bytes = max(peak_bytes, used)
            time.sleep(0.0005)
        makespan = time.perf_counter() - t_start
        return done, makespan, {"peak_active": peak_active, "peak_budget_bytes": peak_bytes}


def poisson_requests(rng, rate, seconds, n_range=(32, 160), k_range=(10, 100), eps=1e-8):
    """Poisson arrivals at `rate` per second over `seconds`: cubic scenes
    with N uniform in n_range, K uniform in k_range steps"""
    t = 0.0
    out = []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > seconds:
            break
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        out.append({"id": len(out), "t": t, "n": n, "k": int(rng.integers(k_range[0], k_range[1] + 1)),
                    "text": random_scene(rng, n, eps)})
    return out


def summarize(records, makespan):
    lat = np.array([x for r in records for x in r["lat"]])
    comp = np.array([r["end"] - r["arrival"] for r in records])
    wait = np.array([r["start"] - r["arrival"] for r in records])
    steps = int(sum(r["k"] for r in records if r["error"] is None))
    cells_steps = float(sum(r["k"] * r["n"] ** 3 for r in records if r["error"] is None))
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3) if len(a) else None
    return {"requests": len(records), "errors": sum(r["error"] is not None for r in records),
            "steps": steps, "makespan_s": round(makespan, 3),
            "steps_per_s": round(steps / makespan, 2),
            "cells_steps_per_s": round(cells_steps / makespan, 1),
            "step_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
            "completion_latency_ms": {"p50": pct(comp, 50), "p99": pct(comp, 99)},
            "queue_wait_ms": {"p50": pct(wait, 50), "p99": pct(wait, 99)}}
