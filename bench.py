#!/usr/bin/env python3
"""Coupled steps/s of the MI355X drop-in on the unit-cube rubber-ball scene.

Workload (the 128^3 configuration): 129^3 nodes, dh = 1/128, one
corotational sphere (r = 0.15 at (0.5, 0.65, 0.5), E = 1e4, nu = 0.3,
solid:fluid density 1000:1000), smoke carpet (fluid pool) over the bottom
30%, gravity 9.8, dt = 1/240, every Krylov solve to 1e-8 relative residual.

A step is one esf_step: CFL, corotational stress, body forces, the three
Schur-complement solves, solid finalize, mass gathering, advection of
velocity and intensity, boundary conditions -- with all fields resident in
HBM.  The scene runs twice: FP64 (the reference's CR, "plain") and mixed
(FP32 CR inside FP64 iterative refinement).  `value` is the mixed form's
steps/s; the FP64 rate, the speedup and the per-field difference between
the two forms are reported beside it.

roofline: the dominant Krylov kernel of the step.  After the timed steps,
up to 10 more steps of the same run are taken with every launch of the
timed kernel classes (pressure SpMV modes, CR direction, solid cell and node
passes) bracketed by HIP events on the context's stream; the class with the
most device time is the dominant kernel and `achieved` = its algorithmic
bytes per launch / its average launch time.  (The events add small gaps
between launches, so they stay out of the throughput measurement.)
The pressure SpMV alone (algorithmic bytes w * (2 N_cells + N_interior_nodes))
is also timed over back-to-back launches at the workload size and at 256^3,
where the working set exceeds the last-level cache.

cpu_baseline: the CPU oracle (C++/OpenMP restatement of the reference step)
on the same scene, teacher-forced from the GPU's state after warm-up, timed
for a bounded number of steps on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "coupled steps/s @128^3 mixed vs FP64 at 1e-8 resid; p50/p99 served step latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(esf, text, precision, warmup, steps, profiled_steps):
    """warm-up, then `steps` timed steps back to back (wall clock between two
    synchronisations; per-step latency is each call's time: a step returns
    once what the host reads of it has landed, its last launches finishing
    behind the next step's first), then `profiled_steps` more steps with HIP
    events around every launch of the timed kernel classes (the events add
    gaps between launches, so they are kept out of the throughput
    measurement)"""
    g = esf.CoupledIntegrator3D.from_cfg_text(text, precision)
    g.u0 = g.download(esf._abi.S_U, 0)
    # about 0.3 s of back-to-back pressure applies (results unused) first, so
    # the warm-up and timed steps run at the clocks sustained load settles to
    ms = g.bench_pressure_apply(fp32=False, reps=10)
    g.bench_pressure_apply(fp32=False, reps=max(10, min(20000, int(300.0 / max(ms, 1e-3)))))
    for _ in range(warmup):
        g.stepImplicit(-1.0)
    g.synchronize()
    lat = []
    iters = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        r = g.stepImplicit(-1.0)
        lat.append(time.perf_counter() - s0)
        iters.append(r)
    g.synchronize()
    wall = time.perf_counter() - t0
    g.kernel_timing(reset=True)
    g.set_kernel_timing(True)
    for _ in range(profiled_steps):
        g.stepImplicit(-1.0)
    g.synchronize()
    g.set_kernel_timing(False)
    return g, wall, np.array(lat), iters, g.kernel_timing(reset=True)


def kernel_bytes(cls, n_cells, n_int, n_sc=0, n_sn=0):
    """algorithmic HBM bytes per launch of each timed Krylov kernel class:
    every cell/node stream the kernel must read or write once.
    Solid classes (n_sc valid solid cells, n_sn solid DOF nodes):
      cell pass: per cell R, L, F, 1/J (28 f64), 8 corner DOF indices, 24 f64
                 out; per node r, D^-1 s, D^-1 Ap (3 x 3 f64)
      node pass: per node 8 gather slots, x, r, p, s, Ap, D^-1 s, D^-1 Ap, D,
                 M~ read and x, r, p, s, Ap, D^-1 s, D^-1 Ap written
                 (16 x 3 f64); per cell 24 f64 gathered"""
    if cls == "solid_cell_cr":
        return n_sc * (28 * 8 + 8 * 4 + 24 * 8) + n_sn * 3 * 24
    if cls == "solid_node_cr":
        return n_sn * (8 * 4 + 16 * 24) + n_sc * 24 * 8
    # the same passes carrying the true-residual check: x read at the
    # corners and its 24 rows per cell written; the node pass reads those
    # rows and b and writes the kept x
    if cls == "solid_cell_crx":
        return n_sc * (28 * 8 + 8 * 4 + 2 * 24 * 8) + n_sn * 4 * 24
    if cls == "solid_node_crx":
        return n_sn * (8 * 4 + 18 * 24) + 2 * n_sc * 24 * 8
    w = 4 if "32" in cls else 8
    c, i = n_cells, n_int
    table = {"apply": w * (2 * c + i), "setup_r": w * (5 * c + i), "setup_s": w * (4 * c + i),
             "update": w * (11 * c + i), "check": w * (2 * c + i),
             "ir_resid": 8 * c + 4 * c + 8 * c + 8 * i + 8 * c + 4 * c, "dir": w * 7 * c}
    for k, v in table.items():
        if cls.endswith(k) or (k == "dir" and cls.startswith("cr_dir")):
            return v
    raise KeyError(cls)


def dominant_roofline(kt, n_cells, n_int, n_sc, n_sn):
    """the Krylov kernel class with the most device time in the timed steps"""
    cls = max(kt, key=lambda k: kt[k][0])
    total, count = kt[cls]
    if count == 0:
        return None
    ms = total / count
    b = kernel_bytes(cls, n_cells, n_int, n_sc, n_sn)
    gbs = b / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": load_traffic(cls), "kernel_class": cls,
            "algorithmic_bytes": b, "avg_launch_ms": round(ms, 5), "launches": count,
            "share_of_timed_kernel_ms": round(total / sum(v[0] for v in kt.values()), 3)}


def spmv_roofline(g, n_cells, n_int, fp32, reps):
    w = 4 if fp32 else 8
    # about 0.2 s of back-to-back applies first, so the timed ones run at
    # the clocks sustained load settles to
    ms = g.bench_pressure_apply(fp32=fp32, reps=10)
    g.bench_pressure_apply(fp32=fp32, reps=max(10, min(20000, int(200.0 / max(ms, 1e-3)))))
    ms = g.bench_pressure_apply(fp32=fp32, reps=reps)
    bytes_ = w * (2 * n_cells + n_int)
    gbs = bytes_ / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "kSpmv<%s,M_APPLY>" % ("float" if fp32 else "double"),
            "algorithmic_bytes": bytes_, "avg_launch_ms": round(ms, 5)}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(text, label, g_fp64, esf, alpha, warmup, timed, seconds):
    """The CPU oracle on the same scene: its state set from the GPU FP64
    context after the timed steps (grid fields, warm start, the solid's
    fields, mass and boxes), `warmup` untimed steps, then up to `timed`
    steps timed one by one (fewer if `seconds` run out); the median step."""
    import ctypes
    from oracle import oracle as O
    o = O.OracleScene.from_cfg_text(text)
    A = esf._abi
    o.set_field(O.VELOCITY, g_fp64.download(A.F_VELOCITY))
    o.set_field(O.INTENSITY, g_fp64.download(A.F_INTENSITY))
    o.set_field(O.LAMBDA, g_fp64.download(A.F_LAMBDA))
    b = (ctypes.c_int * 12)(*g_fp64.boxes(0))
    o.L.orc_set_boxes(o.h, 0, b)
    for f_o, f_g in ((O.S_U, A.S_U), (O.S_X, A.S_X), (O.S_VELOCITY, A.S_VELOCITY),
                     (O.S_VEL_WORKSPACE, A.S_VEL_WORKSPACE), (O.S_MASS, A.S_MASS),
                     (O.S_QFRAC, A.S_QFRAC), (O.S_VALID, A.S_VALID), (O.S_PHI, A.S_PHI)):
        o.set_field(f_o, g_fp64.download(f_g, 0), 0)
    o.L.orc_gather(o.h)
    for _ in range(warmup):
        o.step(alpha)
    times = []
    t0 = time.perf_counter()
    while len(times) < timed:
        s0 = time.perf_counter()
        o.step(alpha)
        times.append(time.perf_counter() - s0)
        if time.perf_counter() - t0 > seconds:
            break
    cores = O.lib().orc_num_threads()
    o.close()
    med = float(np.median(times))
    return {"value": round(1.0 / med, 4), "unit": "coupled steps/s", "cores": cores,
            "kind": "port",
            "sample": f"median of {len(times)} oracle steps (after {warmup} warm-up steps) of the "
                      f"same {label} scene from the GPU FP64 context's state; step times "
                      f"{[round(t, 3) for t in times]} s",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND",
                                                   "OMP_PLACES")}}


def sweeps(esf, cfg, n, reps):
    """The fast-sweeping entry points on the n^3 ball's mass (the headline
    scene's sphere, r = 0.15 at (0.5, 0.65, 0.5)): esf_compute_sdf and
    esf_extend_vector_field of a random field, wall time per call with the
    host copies, best of `reps`, beside the oracle on one host thread
    (checker only), and whether the two agree bitwise"""
    N = n + 1
    g = esf.CoupledIntegrator3D.from_cfg_text(cfg(n))
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    m = (np.sqrt((x / n - 0.5) ** 2 + (y / n - 0.65) ** 2 + (z / n - 0.5) ** 2) < 0.15).astype(
        np.float64)
    f = np.random.default_rng(1).normal(size=(N, N, N, 3))
    t_sdf, t_ext = [], []
    for _ in range(reps):
        t = time.perf_counter()
        phi, box, tight = g.computeSDF(m)
        t_sdf.append(time.perf_counter() - t)
        t = time.perf_counter()
        out = g.extendVectorField(box, m, phi, f)
        t_ext.append(time.perf_counter() - t)
    g.close()
    from oracle import oracle as O
    t = time.perf_counter()
    phi_o, box_o, _ = O.compute_sdf(m, 5, 20)
    c_sdf = time.perf_counter() - t
    t = time.perf_counter()
    out_o = O.extend_vector_field(m, phi_o, box_o, f)
    c_ext = time.perf_counter() - t
    print(json.dumps({
        "workload": f"fast sweeping on the {n}^3 ball's mass, box {[int(v) for v in box]}",
        "gpu_ms": {"compute_sdf": round(1e3 * min(t_sdf), 2),
                   "extend_vector_field": round(1e3 * min(t_ext), 2)},
        "cpu_oracle_ms_1_thread": {"compute_sdf": round(1e3 * c_sdf, 2),
                                   "extend_vector_field": round(1e3 * c_ext, 2)},
        "host_copies_included": True,
        "bitwise": {"phi": bool(np.array_equal(phi.view(np.int64), phi_o.view(np.int64))),
                    "field": bool(np.array_equal(out.view(np.int64), out_o.view(np.int64)))},
    }), flush=True)


def spmv_only(esf, cfg, n, reps):
    g = esf.CoupledIntegrator3D.from_cfg_text(cfg(n), esf.ESF_MIXED)
    g.forces()
    r64 = spmv_roofline(g, n ** 3, (n - 1) ** 3, False, reps)
    r32 = spmv_roofline(g, n ** 3, (n - 1) ** 3, True, reps)
    nbytes = r64["algorithmic_bytes"] // 2
    g.bench_copy(nbytes, reps)
    ms = g.bench_copy(nbytes, reps)
    copy = {"bytes_read_plus_written": 2 * nbytes, "avg_launch_ms": round(ms, 5),
            "achieved": round(2 * nbytes / (ms * 1e-3) / 1e9, 1), "unit": "GB/s"}
    print(json.dumps({"workload": f"pressure SpMV, {n}^3 cells", "fp64": r64, "fp32": r32,
                      "copy": copy}), flush=True)
    g.close()


def divergence_field(u, n):
    """b = J u on an n^3-cell grid for an AoS node field u ((n+1)^3 x 3,
    zero on the boundary nodes): b_c = sum over the cell's 8 corners of
    +-u_x +-u_y +-u_z, sign + on the corner's high side
    (COUPLED_INTEGRATOR_3D.cpp:239-249)"""
    U = u.reshape(n + 1, n + 1, n + 1, 3)  # z, y, x
    b = np.zeros((n, n, n))
    for oz in (0, 1):
        for oy in (0, 1):
            for ox in (0, 1):
                C = U[oz:oz + n, oy:oy + n, ox:ox + n]
                b += (1 if ox else -1) * C[..., 0]
                b += (1 if oy else -1) * C[..., 1]
                b += (1 if oz else -1) * C[..., 2]
    return b.reshape(-1)


def isolated_system(esf, n, seed, eps, reps, max_it=100000):
    """The isolated coupled linear system at n^3 cells: k = 8 spheres from
    the seed (centres uniform in [0.15, 0.85]^3, non-overlapping, radii
    uniform in [0.05, 0.12], E log-uniform in [1e3, 1e6]), the domain 60 %
    fluid-filled; the Schur pressure system S x = b with S = J M^-1 J^T on
    those masses and b a random divergence field (b = J u, u ~ N(0,1) at
    interior nodes, so b lies in the range of the singular S), solved from
    x = 0 to `eps` in FP64 and in the mixed form."""
    from eulersf_amd.cfg import random_spheres, scene_cfg
    rng = np.random.default_rng(seed)
    spheres = random_spheres(rng, 8, (0.05, 0.12), (3.0, 6.0), center_range=(0.15, 0.85))
    text = scene_cfg(n, spheres, fill=0.6, eps=eps)
    u = np.zeros(((n + 1) ** 3, 3))
    inner = np.zeros((n + 1, n + 1, n + 1), dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    u[inner.reshape(-1)] = rng.standard_normal((int(inner.sum()), 3))
    b = divergence_field(u, n)
    del u
    out = {"workload": f"isolated pressure system, {n}^3 cells, 8 spheres (seed {seed}), "
                       f"60% fluid, b = J u with u ~ N(0,1), x0 = 0, eps {eps:g}",
           "cells": n ** 3, "seed": seed}
    sols = {}
    for name, prec in (("fp64", esf.ESF_PLAIN_FP64), ("mixed", esf.ESF_MIXED)):
        g = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
        g.forces()
        g.pressureDiag()
        times = []
        for r in range(reps):
            g.upload(esf._abi.F_PRESSURE_RHS, b)
            g.upload(esf._abi.F_LAMBDA, np.zeros(n ** 3))
            it, ref, rr, ms, st = g.pressureSolveResident(eps, max_it)
            times.append(ms)
        sols[name] = g.download(esf._abi.F_LAMBDA)
        spmv = spmv_roofline(g, n ** 3, (n - 1) ** 3, prec == esf.ESF_MIXED, 20)
        out[name] = {"iterations": it, "refinements": ref, "rel_residual": rr, "status": st,
                     "time_to_solution_ms": round(float(np.median(times)), 3),
                     "time_to_solution_ms_all": [round(t, 3) for t in times],
                     "spmv": spmv}
        g.close()
    # the pressure is fixed up to null(S); compare residuals of the two
    # solutions through S itself on the FP64 context
    g = esf.CoupledIntegrator3D.from_cfg_text(text, esf.ESF_PLAIN_FP64)
    g.forces()
    d = sols["fp64"] - sols["mixed"]
    Sd = g.getMatVecMult(d)
    Sx = g.getMatVecMult(sols["fp64"])
    out["mixed_vs_fp64"] = {"rel_diff_S_x": float(np.linalg.norm(Sd) / np.linalg.norm(Sx))}
    g.close()
    out["solid"] = isolated_solid_system(esf, text, len(spheres), rng, eps, reps, max_it)
    print(json.dumps(out), flush=True)


def isolated_solid_system(esf, text, k, rng, eps, reps, max_it=100000):
    """the solid half of the isolated system on the same scene: each valid
    cell's deformation gradient a random rotation times diag(N(1, 0.05)),
    the elastic stiffness assembled from those, A = dt^2 K + M~ (beta = 0)
    solved for A x = b, b ~ N(0, 1) per solid dof, from x = 0 to `eps` in
    FP64 by Jacobi-preconditioned CR"""
    from eulersf_amd.cfg import random_deformations
    g = esf.CoupledIntegrator3D.from_cfg_text(text, esf.ESF_PLAIN_FP64)
    cells = 0
    for b in range(k):
        nc = g.solidNumCells(b)
        g.setDeformation(random_deformations(rng, nc), b)
        cells += nc
    g.setDt(1.0 / 240)
    g.forces()
    ns = g.num_solid_nodes()
    rhs = rng.standard_normal(3 * ns)
    times = []
    for r in range(reps):
        g.upload(esf._abi.F_SOLID_RHS, rhs)
        it, rr, ms, st = g.solidSolveResident(eps, max_it)
        times.append(ms)
    g.close()
    return {"solid_nodes": ns, "valid_cells": cells, "iterations": it, "rel_residual": rr,
            "status": st, "time_to_solution_ms": round(float(np.median(times)), 3),
            "time_to_solution_ms_all": [round(t, 3) for t in times]}


def batched(esf, count, steps, workers, seed):
    """The batched uneven scenes: `count` scenes with N drawn from
    {32, 48, 64, 80, 96}, built like the served scenes; K timed steps of
    every scene through esf_step_batch (each kernel of a step launched once
    over all scenes), then one scene at a time, then concurrently on
    `workers` host threads (one HIP stream per scene)."""
    from eulersf_amd.serve import BatchRunner, random_scene
    rng = np.random.default_rng(seed)
    ns = [int(rng.choice([32, 48, 64, 80, 96])) for _ in range(count)]
    texts = [random_scene(rng, n) for n in ns]
    log(f"[bench] batched: {count} scenes, {steps} steps each, {workers} workers")
    br = BatchRunner(texts, ns, esf.ESF_MIXED)
    cells = sum(br.cells())
    wb, lb = br.run_batched(steps)
    log(f"[bench] batched: one launch per kernel {wb:.3f} s")
    ws, ls = br.run_sequential(steps)
    log(f"[bench] batched: one scene at a time {ws:.3f} s")
    wc, lc = br.run_concurrent(steps, workers)
    log(f"[bench] batched: {workers} threads {wc:.3f} s")
    mem = sum(g.device_bytes() for g in br.scenes)
    br.close()
    pct = lambda a, q: round(float(np.percentile(a, q)) * 1e3, 3)
    out = {"workload": f"{count} scenes, N in {{32,48,64,80,96}} (seed {seed}), 1-3 spheres, "
                       f"mixed precision at 1e-8, {steps} steps per scene after 2 warm-up steps",
           "cells_total": cells, "device_bytes_total": mem,
           "batched": {"wall_s": round(wb, 3), "cells_steps_per_s": round(cells * steps / wb, 1),
                       "step_latency_ms": {"p50": pct(lb, 50), "p99": pct(lb, 99)},
                       "note": "esf_step_batch: every scene's step in one call, each kernel "
                               "launched once over the scenes that reach it"},
           "sequential": {"wall_s": round(ws, 3), "cells_steps_per_s": round(cells * steps / ws, 1),
                          "step_latency_ms": {"p50": pct(ls, 50), "p99": pct(ls, 99)}},
           "concurrent": {"workers": workers, "wall_s": round(wc, 3),
                          "cells_steps_per_s": round(cells * steps / wc, 1),
                          "step_latency_ms": {"p50": pct(lc, 50), "p99": pct(lc, 99)}},
           "batched_over_sequential": round(ws / wb, 3),
           "batched_over_concurrent": round(wc / wb, 3)}
    return out


def served(esf, seconds, rate, workers, seed, one_at_a_time=True, batch_slots=64):
    """The serving stream: Poisson arrivals at `rate`/s for `seconds`, cubic
    scenes with N uniform in [32, 160] asking for K uniform in [10, 100]
    steps, admitted against a 200 GB HBM budget.  Served three ways: the
    batched loop (every in-flight scene advanced by one esf_step_batch per
    step, up to `batch_slots` in flight), `workers` host threads stepping
    their scenes concurrently, and one scene at a time."""
    from eulersf_amd.serve import BatchedServer, Server, poisson_requests, summarize
    rng = np.random.default_rng(seed)
    reqs = poisson_requests(rng, rate, seconds)
    log(f"[bench] served: {len(reqs)} requests over {seconds} s")
    out = {"workload": f"Poisson arrivals at {rate}/s for {seconds} s (seed {seed}): N ~ U[32,160], "
                       "K ~ U[10,100] steps, 1-3 spheres, mixed precision at 1e-8, "
                       "200 GB HBM budget", "requests": len(reqs)}
    out["hw_queues"] = os.environ.get("GPU_MAX_HW_QUEUES", "runtime default (4)")
    recs, span, info = BatchedServer(200e9, batch_slots, esf.ESF_MIXED).run(reqs)
    out["batched"] = {"slots": batch_slots, **summarize(recs, span), **info}
    log(f"[bench] served, batched: {out['batched']['steps_per_s']} steps/s")
    recs, span, info = Server(200e9, workers, esf.ESF_MIXED).run(reqs)
    out["concurrent"] = {"slots": workers, **summarize(recs, span), **info}
    log(f"[bench] served, {workers} threads: {out['concurrent']['steps_per_s']} steps/s")
    if one_at_a_time:
        log("[bench] served, one scene at a time")
        recs1, span1, info1 = Server(200e9, 1, esf.ESF_MIXED).run(reqs)
        out["one_at_a_time"] = {**summarize(recs1, span1), **info1}
        out["throughput_ratio"] = round(out["batched"]["steps_per_s"] /
                                        out["one_at_a_time"]["steps_per_s"], 3)
        out["throughput_ratio_threads"] = round(out["concurrent"]["steps_per_s"] /
                                                out["one_at_a_time"]["steps_per_s"], 3)
    return out


def drift(esf, n, frames, velocity, flip, eps, every):
    """The free-running drift of the GPU step from the oracle's over
    `frames` frames of the n^3 ball (APPLICATION::step, frame dt = the cfg
    dt, CFL alpha 1.8), both with the full finalizeSolution (FLIP particles,
    velocity extension) when `flip`: per frame the relative difference of
    the velocity (against max(|v|, g dt)), the displacement, the intensity
    and the solid mass"""
    from oracle import oracle as O
    extra = {"flip particles": 1, "number of particles per material cell": 8} if flip else None
    text = esf.unit_cube_ball_cfg(n, eps=eps, max_iterations=20000, velocity=velocity, extra=extra)
    g = esf.CoupledIntegrator3D.from_cfg_text(text)
    o = O.OracleScene.from_cfg_text(text)
    A = esf._abi
    frame_dt = 1.0 / 240.0
    rows = []
    t_gpu = t_cpu = 0.0
    sub_g = sub_o = 0
    for f in range(frames):
        t0 = time.perf_counter()
        sub_g += g.frame(frame_dt, 1.8)
        t_gpu += time.perf_counter() - t0
        t0 = time.perf_counter()
        sim = 0.0
        while sim < frame_dt:
            o.L.orc_set_max_dt(o.h, frame_dt - sim)
            rc, st = o.step(1.8)
            sim += st.dt
            sub_o += 1
        t_cpu += time.perf_counter() - t0
        if f % every == every - 1 or f == frames - 1:
            vg, vo = g.download(A.F_VELOCITY), o.field(O.VELOCITY)
            floor = 9.8 * frame_dt * np.sqrt(vo.size)
            row = {"frame": f + 1,
                   "velocity": float(np.linalg.norm(vg - vo) / max(np.linalg.norm(vo), floor)),
                   "displacement": float(np.linalg.norm(g.download(A.S_U, 0) - o.field(O.S_U, 0)) /
                                         np.linalg.norm(o.field(O.S_U, 0))),
                   "intensity": float(np.linalg.norm(g.download(A.F_INTENSITY) -
                                                     o.field(O.INTENSITY)) /
                                      np.linalg.norm(o.field(O.INTENSITY))),
                   "solid_mass": float(np.linalg.norm(g.download(A.S_MASS, 0) -
                                                      o.field(O.S_MASS, 0)) /
                                       np.linalg.norm(o.field(O.S_MASS, 0)))}
            rows.append(row)
            log(f"[bench] drift frame {f + 1}: " + ", ".join(f"{k} {v:.2e}" for k, v in row.items()
                                                            if k != "frame"))
    g.close()
    o.close()
    print(json.dumps({"workload": f"{n}^3 ball, initial velocity {velocity}, {frames} frames of "
                                  f"dt 1/240, solves to {eps:g}, "
                                  f"{'FLIP particles and velocity extension' if flip else 'no FLIP'}",
                      "substeps": {"gpu": sub_g, "oracle": sub_o},
                      "seconds": {"gpu": round(t_gpu, 2), "oracle": round(t_cpu, 2)},
                      "max": {k: max(r[k] for r in rows) for k in rows[0] if k != "frame"},
                      "frames": rows}), flush=True)


def traffic_run(esf, text):
    """The command the HBM byte counters are collected over (rocprofv3 --pmc
    FETCH_SIZE, then WRITE_SIZE): a few steps of each form, so every Krylov
    kernel class runs, then a streaming copy of known size (256 MiB read +
    256 MiB written per launch) that calibrates the counters."""
    for prec in (esf.ESF_PLAIN_FP64, esf.ESF_MIXED):
        g = esf.CoupledIntegrator3D.from_cfg_text(text, prec)
        for _ in range(3):
            g.stepImplicit(-1.0)
        g.synchronize()
        g.close()
    g = esf.CoupledIntegrator3D.from_cfg_text(text, esf.ESF_PLAIN_FP64)
    nbytes = 256 << 20
    ms = g.bench_copy(nbytes, 3)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "copy_avg_ms": round(ms, 5)}), flush=True)
    g.close()


def load_traffic(cls):
    """HBM bytes per launch of kernel class `cls` from the committed PMC
    summary (profiles/traffic_unit_ball_128.md, produced by two rocprofv3
    --pmc runs over `bench.py --traffic-run`), or None"""
    path = os.path.join(ROOT, "profiles", "traffic_unit_ball_128.md")
    if not os.path.exists(path):
        return None
    text = open(path).read()
    try:
        block = text.split("```json", 1)[1].split("```", 1)[0]
        table = json.loads(block)["bytes_per_launch"]
    except (IndexError, ValueError, KeyError):
        return None
    v = table.get(cls)
    return None if v is None else int(v)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--velocity", type=str, default="0,0,0",
                    help="the ball's initial velocity (vx,vy,vz); the workload is at rest")
    ap.add_argument("--spmv-reps", type=int, default=100)
    ap.add_argument("--roofline-n", type=int, default=256,
                    help="grid of the HBM-bound SpMV measurement (0 to skip)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="bound on the CPU baseline's timed steps")
    ap.add_argument("--cpu-warmup", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-run", action="store_true",
                    help="a few steps of both forms and a calibrating copy, for the "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--system", type=int, default=0, metavar="N",
                    help="the isolated coupled linear system on N^3 cells (8 seeded spheres): "
                         "FP64 and mixed solves to --eps from x = 0; prints one JSON line")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--system-reps", type=int, default=3)
    ap.add_argument("--system-iters", type=int, default=100000,
                    help="iteration cap of the isolated system's solves (short profiling runs)")
    ap.add_argument("--batch", type=int, default=0, metavar="S",
                    help="the batched uneven scenes: S scenes stepped one at a time and "
                         "concurrently; prints one JSON line")
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--workers", type=int, default=16,
                    help="concurrent scenes (host threads / HIP streams) of --batch and --serve")
    ap.add_argument("--serve", type=float, default=0.0, metavar="SECONDS",
                    help="the serving stream over SECONDS of Poisson arrivals, with the "
                         "concurrent server and one scene at a time; prints one JSON line")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--batched-scenes", type=int, default=64,
                    help="scenes of the batched-workload block of the default run (0 to skip)")
    ap.add_argument("--served-seconds", type=float, default=3.0,
                    help="arrival window of the short served-stream block of the default "
                         "run (0 to skip)")
    ap.add_argument("--drift", type=int, default=0, metavar="FRAMES",
                    help="free-running GPU vs oracle over FRAMES frames of the --n ball "
                         "(full finalizeSolution with FLIP unless --no-flip); prints one JSON line")
    ap.add_argument("--no-flip", action="store_true")
    ap.add_argument("--drift-every", type=int, default=10)
    ap.add_argument("--sweep", type=int, default=0, metavar="N",
                    help="only time the fast-sweeping entry points on the N^3 ball; prints one "
                         "JSON line")
    ap.add_argument("--spmv-only", type=int, default=0, metavar="N",
                    help="only time the pressure SpMV (FP64, FP32) on the N^3 scene and a "
                         "streaming copy of the same byte count; prints one JSON line")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the coupled step has no multi-GPU path in this build", file=sys.stderr)
        sys.exit(2)

    # the CPU baseline's OpenMP threads, bound one per core (read when the
    # oracle library loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import eulersf_amd as esf
    esf.load()
    from eulersf_amd.cfg import unit_cube_ball_cfg
    if args.spmv_only:
        spmv_only(esf, unit_cube_ball_cfg, args.spmv_only, args.spmv_reps)
        return
    if args.sweep:
        sweeps(esf, unit_cube_ball_cfg, args.sweep, 3)
        return
    if args.drift:
        vel = tuple(float(v) for v in args.velocity.split(","))
        drift(esf, args.n, args.drift, vel, not args.no_flip, args.eps, args.drift_every)
        return
    if args.system:
        isolated_system(esf, args.system, args.seed, args.eps, args.system_reps,
                        args.system_iters)
        return
    if args.batch:
        print(json.dumps(batched(esf, args.batch, args.batch_steps, args.workers, args.seed)),
              flush=True)
        return
    if args.serve:
        print(json.dumps(served(esf, args.serve, args.serve_rate, args.workers, args.seed)),
              flush=True)
        return
    n = args.n
    vel = tuple(float(v) for v in args.velocity.split(","))
    text = unit_cube_ball_cfg(n, eps=args.eps, max_iterations=20000, velocity=vel)
    n_cells, n_int = n ** 3, (n - 1) ** 3
    if args.traffic_run:
        traffic_run(esf, text)
        return

    log(f"[bench] FP64 form: {args.warmup} warm-up + {args.steps} timed steps at {n}^3")
    prof = max(1, min(args.steps, 10))
    g64, wall64, lat64, it64, kt64 = run_form(esf, text, esf.ESF_PLAIN_FP64, args.warmup,
                                              args.steps, prof)
    log(f"[bench] FP64: {args.steps / wall64:.2f} steps/s")
    log(f"[bench] mixed form: {args.warmup} warm-up + {args.steps} timed steps")
    gmx, wallmx, latmx, itmx, ktmx = run_form(esf, text, esf.ESF_MIXED, args.warmup, args.steps,
                                              prof)
    log(f"[bench] mixed: {args.steps / wallmx:.2f} steps/s")

    A = esf._abi
    g64_u0 = g64.u0
    errs = {}
    # max |mixed - fp64| over the field, relative to the field's scale; the
    # velocity scale is at least g*dt (the velocity one step of gravity adds
    # before the projection removes it), the displacement scale at least dh
    dt = 1.0 / 240.0
    floors = {"velocity": 9.8 * dt, "intensity": 0.0, "displacement": 1.0 / n, "pressure": 0.0}
    for name, fid, solid in (("velocity", A.F_VELOCITY, -1), ("intensity", A.F_INTENSITY, -1),
                             ("displacement", A.S_U, 0), ("pressure", A.F_LAMBDA, -1)):
        a = g64.download(fid, solid)
        b = gmx.download(fid, solid)
        if name == "displacement":  # compare the motion, not the placement
            u0 = g64_u0
            a, b = a - u0, b - u0
        scale = max(float(np.max(np.abs(a))), floors[name]) or 1.0
        errs[name] = {"max_abs": float(np.max(np.abs(a - b))), "scale": scale,
                      "rel": float(np.max(np.abs(a - b)) / scale)}

    roof64 = spmv_roofline(g64, n_cells, n_int, False, args.spmv_reps)
    roof32 = spmv_roofline(gmx, n_cells, n_int, True, args.spmv_reps)
    roof_big = None
    if args.roofline_n:
        nb = args.roofline_n
        log(f"[bench] SpMV roofline at {nb}^3")
        big = esf.CoupledIntegrator3D.from_cfg_text(unit_cube_ball_cfg(nb, eps=args.eps),
                                                    esf.ESF_MIXED)
        big.forces()  # masses -> 1/m
        roof_big = {"fp64": spmv_roofline(big, nb ** 3, (nb - 1) ** 3, False, args.spmv_reps),
                    "fp32": spmv_roofline(big, nb ** 3, (nb - 1) ** 3, True, args.spmv_reps)}
        big.close()

    serve_block = None
    if args.served_seconds > 0:
        serve_block = served(esf, args.served_seconds, args.serve_rate, args.workers, args.seed)

    batch_block = None
    if args.batched_scenes > 0:
        batch_block = batched(esf, args.batched_scenes, args.batch_steps, args.workers, args.seed)

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU oracle baseline ({args.cpu_warmup} warm-up + median of "
            f"{args.cpu_steps} steps)")
        cpu = cpu_baseline(text, f"{n}^3", g64, esf, -1.0, args.cpu_warmup, args.cpu_steps,
                           args.cpu_seconds)
        log(f"[bench] CPU: {cpu['value']} steps/s on {cpu['cores']} threads")

    def summary(lat, iters):
        pr = [r["iterations"][1] for r in iters]
        return {"p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
                "pressure_iterations_mean": float(np.mean(pr)),
                "solid_iterations_mean": float(np.mean([r["iterations"][0] + r["iterations"][2]
                                                        for r in iters])),
                "max_rel_residual": float(max(max(r["rel_residual"]) for r in iters))}

    value = args.steps / wallmx
    n_sc, n_sn = itmx[-1]["n_solid_cells"], itmx[-1]["n_solid_nodes"]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "coupled steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wallmx / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 Krylov in f64 refinement (f64 elsewhere)",
        "data": "synthetic scene (analytic sphere), no external data",
        "config": {"workload": f"unit-cube rubber ball, {n}^3 cells ({n + 1}^3 nodes), "
                               "sphere r=0.15 E=1e4 nu=0.3, rho 1000:1000, dt=1/240, "
                               f"initial velocity {vel}, solves to {args.eps:g}",
                   "grid_cells": n_cells, "solid_nodes": n_sn, "solid_cells": n_sc,
                   "parallelism": "single GPU"},
        "fp64": {"steps_per_s": round(args.steps / wall64, 3),
                 "ms_per_step": round(wall64 / args.steps * 1e3, 3), **summary(lat64, it64)},
        "mixed": {"steps_per_s": round(value, 3), **summary(latmx, itmx),
                  "refinements_mean": float(np.mean([r["refinements"] for r in itmx]))},
        "mixed_over_fp64": round(value / (args.steps / wall64), 3),
        "mixed_vs_fp64_max_rel_err": errs,
        "step_latency_ms": {"p50": round(float(np.percentile(latmx, 50)) * 1e3, 3),
                            "p99": round(float(np.percentile(latmx, 99)) * 1e3, 3)},
        "roofline": dominant_roofline(ktmx, n_cells, n_int, n_sc, n_sn),
        "roofline_fp64_form": dominant_roofline(kt64, n_cells, n_int, n_sc, n_sn),
        "kernel_ms_per_step": {"mixed": {k: round(v[0] / prof, 4) for k, v in ktmx.items() if v[1]},
                               "fp64": {k: round(v[0] / prof, 4) for k, v in kt64.items() if v[1]},
                               "profiled_steps": prof},
        "kernel_us_per_launch": {"mixed": {k: round(v[0] / v[1] * 1e3, 2)
                                           for k, v in ktmx.items() if v[1]},
                                 "fp64": {k: round(v[0] / v[1] * 1e3, 2)
                                          for k, v in kt64.items() if v[1]}},
        "spmv_apply_at_workload": {"fp64": roof64, "fp32": roof32},
        "roofline_hbm_sized": roof_big,
        "served_stream": serve_block,
        "batched_scenes": batch_block,
        "cpu_baseline": cpu,
    }
    g64.close()
    gmx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
