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


class BatchedServer:
    """Continuous batching of scene requests, each step one batched launch
    per kernel over the scenes in flight: every in-flight scene is advanced
    by one substep per esf_step_batch call.  Scenes are grouped by size into
    `lanes` (N ranges of equal width over `n_range`); each lane's thread
    steps its scenes as one batch, so a step of small scenes does not wait
    for the solves of large ones, and the lanes' batches run side by side on
    the GPU.  Requests are admitted in arrival order while fewer than
    `max_active` scenes are in flight (or being built) and their estimated
    device memory fits the HBM budget; `builders` threads create the
    admitted scenes' contexts (grids, level set, masses, Krylov state
    resident in HBM) beside the stepping, and a scene joins its lane's batch
    at the lane's next step once built.  A scene's step latency is its batch
    step's; it leaves after its K steps and its memory returns to the
    budget."""

    def __init__(self, budget_bytes, max_active, precision=_abi.ESF_MIXED,
                 bytes_per_node=1200.0, builders=16, lanes=8, n_range=(32, 160), on_done=None):
        self.budget = budget_bytes
        self.max_active = max_active
        self.precision = precision
        self.bytes_per_node = bytes_per_node
        self.builders = builders
        self.lanes = lanes
        self.n_range = n_range
        self.on_done = on_done

    def estimate(self, n):
        return int(self.bytes_per_node * (n + 1) ** 3)

    def lane_of(self, n):
        lo, hi = self.n_range
        w = (hi - lo + 1) / float(self.lanes)
        return min(self.lanes - 1, max(0, int((n - lo) / w)))

    def run(self, requests):
        """requests: list of dicts with keys id, t (arrival, s), n, k, text.
        Returns per-request records, the makespan and admission counters."""
        reqs = sorted(requests, key=lambda r: r["t"])
        t_start = time.perf_counter()
        cv = threading.Condition()
        st = {"used": 0, "inflight": 0, "peak_active": 0, "peak_bytes": 0, "stop": False}
        todo = deque()                         # admitted, waiting for a builder
        ready = [deque() for _ in range(self.lanes)]   # built, joining a lane
        done = []
        batch_sizes = []

        def builder():
            while True:
                with cv:
                    while not todo and not st["stop"]:
                        cv.wait()
                    if not todo:
                        return
                    req, est = todo.popleft()
                rec = {"id": req["id"], "n": req["n"], "k": req["k"], "arrival": req["t"],
                       "lat": [], "error": None, "est": est,
                       "start": time.perf_counter() - t_start}
                g = None
                try:
                    g = CoupledIntegrator3D.from_cfg_text(req["text"], self.precision)
                    rec["bytes"] = g.device_bytes()
                except Exception as e:  # kept in the record, reported by the caller
                    rec["error"] = repr(e)
                with cv:
                    if rec["error"] is None:
                        self.bytes_per_node = max(self.bytes_per_node,
                                                  rec["bytes"] / float((req["n"] + 1) ** 3))
                        ready[self.lane_of(req["n"])].append([rec, g, est, 0])
                    else:
                        rec["end"] = time.perf_counter() - t_start
                        st["used"] -= est
                        st["inflight"] -= 1
                        done.append(rec)
                    cv.notify_all()

        def lane(li):
            active = []
            while True:
                with cv:
                    while not active and not ready[li] and not st["stop"]:
                        cv.wait(0.01)
                    while ready[li]:
                        active.append(ready[li].popleft())
                    if not active and st["stop"]:
                        return
                if not active:
                    continue
                t0 = time.perf_counter()
                try:
                    reps = step_batch([a[1] for a in active])
                    errs = [None if r["status"] >= 0 else f"step failed: {r}" for r in reps]
                except _abi.EsfError as e:
                    errs = [repr(e)] * len(active)
                dt = time.perf_counter() - t0
                keep, leaving = [], []
                for a, err in zip(active, errs):
                    rec, g = a[0], a[1]
                    rec["lat"].append(dt)
                    a[3] += 1
                    if err is not None:
                        rec["error"] = err
                    if a[3] >= rec["k"] or err is not None:
                        if self.on_done is not None and err is None:
                            self.on_done({"id": rec["id"], "n": rec["n"], "k": rec["k"]}, g)
                        g.close()
                        rec["end"] = time.perf_counter() - t_start
                        leaving.append(a)
                    else:
                        keep.append(a)
                with cv:
                    batch_sizes.append(len(active))
                    for a in leaving:
                        st["used"] -= a[2]
                        st["inflight"] -= 1
                        done.append(a[0])
                    cv.notify_all()
                active = keep

        threads = [threading.Thread(target=builder, daemon=True) for _ in range(self.builders)]
        threads += [threading.Thread(target=lane, args=(i,), daemon=True)
                    for i in range(self.lanes)]
        for th in threads:
            th.start()
        nxt = 0
        pending = deque()
        try:
            while True:
                now = time.perf_counter() - t_start
                while nxt < len(reqs) and reqs[nxt]["t"] <= now:
                    pending.append(reqs[nxt])
                    nxt += 1
                with cv:
                    if len(done) == len(reqs):
                        break
                    # admit in arrival order while a slot and the budget allow
                    while pending and st["inflight"] < self.max_active:
                        est = self.estimate(pending[0]["n"])
                        if st["used"] + est > self.budget and st["inflight"]:
                            break
                        req = pending.popleft()
                        st["used"] += est
                        st["inflight"] += 1
                        todo.append((req, est))
                        cv.notify_all()
                    st["peak_active"] = max(st["peak_active"], st["inflight"])
                    st["peak_bytes"] = max(st["peak_bytes"], st["used"])
                    cv.wait(0.0005)
        finally:
            with cv:
                st["stop"] = True
                cv.notify_all()
            for th in threads:
                th.join()
        makespan = time.perf_counter() - t_start
        mean_batch = float(np.mean(batch_sizes)) if batch_sizes else 0.0
        return done, makespan, {"peak_active": st["peak_active"],
                                "peak_budget_bytes": st["peak_bytes"],
                                "lanes": self.lanes, "batch_steps": len(batch_sizes),
                                "mean_scenes_per_batch_step": round(mean_batch, 2)}


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
