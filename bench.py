"""Benchmark: particle-steps/s of the adaptive 3D SPH step on one MI355X.

Workload (default): the seeded 3D-2 dam break -- 262,144 large particles on a jittered 64^3
lattice in a 64^3-cell box (reference-native units), split and merge on, capacity 1,000,000,
marching cubes on the 257^3 density grid every 10th step. A "step" is one SPHSystem::
animation() step (hash, sort, cell ranges, density, force, integrate, [density grid,
normals, marching cubes], energy, split, merge) on device-resident state.

value = sum over the K timed steps of the live particle count / wall time of those steps.
roofline: the kernel with the largest time per launch over the later warm-up steps (the
first steps carry one-off first-launch costs), timed with HIP events on the library's
stream during the timed region; achieved = algorithmic bytes per launch / mean launch time
against the 8 TB/s HBM peak. roofline.traffic is the HBM bytes per launch of that kernel
measured by rocprofv3 PMC runs (FETCH_SIZE, WRITE_SIZE), scaled by what the same counters
read for a plain device copy of known size (--copy-calibration-mib), read from
profiles/<kernel>_<workload>_traffic.json. roofline.valu_frac is the fraction of the VALU
issue peak (one wave64 instruction per SIMD per 4 cycles) that the kernel's VALU
instructions (SQ_INSTS_VALU per live particle, profiles/<kernel>_<workload>_valu.json)
take of its measured launch time -- the neighbour kernels' secondary bound.
full_run: the workload's whole configured step count (2000 steps for the 3D-2 dam break)
from the seeded start on a fresh context, beside the K-step window.
cpu_baseline: the fp32 CPU oracle (a restatement of the reference kernels, OpenMP over
particles / cells / voxels) on a bounded sample of the same workload, started from the GPU
state at the beginning of the timed region.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import sca2012_amd as sa  # noqa: E402
from sca2012_amd import scenes  # noqa: E402
from sca2012_amd._abi import PHASES  # noqa: E402

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per live particle (each array read once and written once per
# phase; neighbour re-reads counted as cache hits) for the 3D step, by profiling phase.
ALG_BYTES = {
    "hash": 20,       # pos 12 + slot 4 -> key 4 (+ slot map 4 written, counted in sort)
    "sort": 16,       # keys + values, read and written once
    "cells": 4,       # sorted keys read once (cell table writes counted per cell below)
    "reorder": 0,     # layout change of this implementation, not in the reference algorithm
    "density": 28,    # pos 12, level 4, index 4 -> dens 4, press 4
    "force": 56,      # pos, ev, dens, press, level, index -> acc 12, surface 4
    "integrate": 92,  # pos, vel, ev, acc, dens, level -> pos, vel, ev
    "energy": 36,     # pos, vel, level, index -> energy
    "split": 12,      # energy, surface, level -> flag
    "merge": 16,      # per-particle status scan
    "compact": 0,
}

PHASE_ID = {name: k for k, name in enumerate(PHASES)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n))))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, params, state, first_step, budget_s, cores):
    """Time the fp32 CPU oracle on consecutive steps of the same workload from `state`."""
    from oracle import orc
    orc.set_threads(cores)
    w = scenes.WORKLOADS[workload]
    o = orc.Oracle(w["variant"], params, w["capacity"])
    o.upload(state)
    # one untimed warm-up step so first-touch page faults are not timed
    o.step(scenes.step_flags(workload, first_step))
    steps, psteps, t_total = 0, 0, 0.0
    k = first_step + 1
    mc_seen = False
    mc_steps = 0
    while True:
        f = scenes.step_flags(workload, k)
        n = o.num_particles()
        t0 = time.perf_counter()
        rc = o.step(f)
        dt = time.perf_counter() - t0
        if rc != 0:
            break
        t_total += dt
        psteps += n
        steps += 1
        mc_seen |= bool(f & sa.STEP_MC)
        mc_steps += 1 if f & sa.STEP_MC else 0
        k += 1
        cover_mc = (not w["mc_every"]) or mc_seen
        if (t_total >= budget_s and cover_mc) or steps >= 50:
            break
    # the same steps on one thread, for the scaling of the CPU figure
    orc.set_threads(1)
    o1 = orc.Oracle(w["variant"], params, w["capacity"])
    o1.upload(state)
    o1.step(scenes.step_flags(workload, first_step))
    steps1, psteps1, t1 = 0, 0, 0.0
    while steps1 < steps and t1 < budget_s / 2:
        n = o1.num_particles()
        t0 = time.perf_counter()
        rc = o1.step(scenes.step_flags(workload, first_step + 1 + steps1))
        dt = time.perf_counter() - t0
        if rc != 0:
            break
        t1 += dt
        psteps1 += n
        steps1 += 1
    orc.set_threads(cores)
    return {
        "value": psteps / t_total if t_total > 0 else None,
        "unit": "particle-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": "%d consecutive steps (%d of them marching-cubes steps) of %s from the GPU "
                  "state at the start of the timed region (%d live particles), fp32 CPU oracle "
                  "with %d OpenMP threads on %s, %.1f s" % (
                      steps, mc_steps, workload, len(state), cores, cpu_model(), t_total),
        "value_1_thread": psteps1 / t1 if t1 > 0 else None,
        "sample_1_thread": "the first %d of those steps on one thread, %.1f s" % (steps1, t1),
    }


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_traffic.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t["fetch_bytes_per_launch"] + t["write_bytes_per_launch"]


# MI355X VALU issue peak: 256 CUs x 4 SIMDs, one wave64 VALU instruction per SIMD every 4
# cycles (16 lanes per SIMD) at 2.4 GHz = 614.4e9 wave-instructions/s, i.e. 78.6 TFLOP/s
# of non-packed fp32 FMA.
VALU_PEAK_WAVE_INST_S = 256 * 4 * 2.4e9 / 4


def measured_valu(kernel, workload):
    """VALU instructions per launch and per live particle of `kernel` on `workload` from the
    committed PMC summary (SQ_INSTS_VALU), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_valu.json" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def valu_fraction(kernel, workload, mean_live, mean_launch_ms):
    """Fraction of the VALU issue peak: the profiled VALU instructions per live particle x the
    live particles of the timed steps, over the launch time measured live."""
    v = measured_valu(kernel, workload)
    if not v or not mean_launch_ms:
        return None, None
    insts = v["valu_insts_per_particle"] * mean_live
    return insts / (mean_launch_ms * 1e-3) / VALU_PEAK_WAVE_INST_S, insts


def tune_in_warmup(obj, warmup, lists=None):
    """Place the pair-list autotuner's first round (six tuning steps) at the end of the
    warm-up; its later rounds (twelve steps, every 128 steps) fall in the timed steps, as in
    any long run (fewer than 7 warm-up steps: the library's own schedule). `lists` = "F,E"
    pins the lengths instead."""
    if lists:
        f, e = (int(x) for x in lists.split(","))
        obj.set_lists(f, e)
    elif warmup >= 7:
        obj.autotune(0xffffffff, after=warmup - 6, every=128)


def copy_calibration(mib):
    """A plain device-to-device copy of known size, for scaling the PMC byte counters."""
    rc = sa.load_library().sph_copy_calibration(mib << 20, 5)
    if rc != 0:
        raise sa.SphError(rc, "sph_copy_calibration")
    log("copy calibration: 5 copies of %d MiB (k_copy_calibration)" % mib)


def batch_stream_run(args, w, tanks, caps, flags):
    """The batch with stream steps (kernels launched one by one, one host synchronisation
    per step): the dominant phase, its mean launch time over the timed steps, the wall time
    of those steps and a per-phase breakdown from further steps."""
    b = sa.SPHBatch(w["variant"], [p for p, _ in tanks], caps)
    b.set_graph(False)
    tune_in_warmup(b, args.warmup, args.lists)
    for s_, (_, a) in enumerate(tanks):
        b.upload(s_, a)
    # the first warm-up step carries one-off first-launch costs: profile the later ones
    b.step(flags, 1)
    b.prof_enable(True)
    b.step(flags, max(args.warmup - 1, 1))
    prof = b.prof_read()
    cand = {p: v[0] for p, v in prof.items() if p in ALG_BYTES and ALG_BYTES[p] > 0}
    dominant = max(cand, key=cand.get) if any(cand.values()) else "force"
    b.prof_reset()
    b.prof_phases([dominant])  # timed region: events around the dominant phase only
    t0 = time.perf_counter()
    for k in range(args.steps):
        b.step(flags)
    wall = time.perf_counter() - t0
    prof = b.prof_read()
    dom_ms, dom_launches = prof[dominant]
    phase_steps = max(1, min(args.steps, 10))
    b.prof_reset()
    b.prof_enable(True)
    b.step(flags, phase_steps)
    phase_ms = {p: round(v[0] / phase_steps, 4) for p, v in b.prof_read().items() if v[1]}
    tuned = b.tuned_lists()
    b.close()
    return dict(dominant=dominant, mean_launch_ms=dom_ms / dom_launches, wall=wall,
                phase_ms=phase_ms, phase_steps=phase_steps, lists=tuned[:2])


def run_batched(args, w):
    """Batched tanks: scene-steps/s of graph steps over all tanks (each step one captured
    hipGraph of persistent launches, the host launching step after step), beside the same
    batch with stream steps (kernels launched one by one, one synchronisation per step) and
    the same tanks stepped one after another through their own single-scene contexts."""
    variant, flags = w["variant"], w["flags"]
    sizes, tanks = w["build"]()
    caps = [int(w["capacity_factor"] * n) for n in sizes]
    ns = len(tanks)
    log("workload %s: %d tanks, %d particles, %d cells" % (
        args.workload, ns, int(sizes.sum()), sum(int(p["total_cell"][0]) for p, _ in tanks)))
    b = sa.SPHBatch(variant, [p for p, _ in tanks], caps)
    b.set_graph(True)
    if args.lists:
        f_, e_ = (int(x) for x in args.lists.split(","))
        b.set_lists(f_, e_)
    if args.groups_per_cu >= 0:
        b.set_launch(args.groups_per_cu)
    if not args.lists and args.groups_per_cu < 0 and args.warmup >= 24:
        # one round of each tuned dimension (grid, force list, energy list) in the warm-up
        b.autotune(0xffffffff, after=args.warmup - 20, every=128)
    for s_, (_, a) in enumerate(tanks):
        b.upload(s_, a)
    b.step(flags, args.warmup)
    gpc, n_cu, grid_ms = b.tuned_launch()
    tuned = b.tuned_lists()
    log("graph steps: %d workgroups per CU on %d CUs (round ms %s), lists force %d energy %d" % (
        gpc, n_cu, grid_ms, tuned[0], tuned[1]))
    n_first = int(b.num_particles().sum())
    state0 = [b.download(s_) for s_ in range(ns)] if not args.no_cpu else None
    t0 = time.perf_counter()
    b.step(flags, args.steps)  # all timed steps in one call: one synchronisation at the end
    wall = time.perf_counter() - t0
    value = ns * args.steps / wall
    n_last = int(b.num_particles().sum())
    b.close()
    mean_n = 0.5 * (n_first + n_last)
    log("graph steps: %.4g scene-steps/s (%.3f ms per step)" % (value, wall * 1e3 / args.steps))

    stream = batch_stream_run(args, w, tanks, caps, flags)
    stream_value = ns * args.steps / stream["wall"]
    dominant = stream["dominant"]
    phase_ms = stream["phase_ms"]
    phase_steps = stream["phase_steps"]
    achieved = ALG_BYTES[dominant] * mean_n / (stream["mean_launch_ms"] * 1e-3) / 1e9
    log("stream steps: %.4g scene-steps/s, phases over %d further steps (ms/step): %s" % (
        stream_value, phase_steps, phase_ms))

    # the same tanks, one context each, stepped one after another (same warm-up)
    ctxs = []
    for (p, a), c in zip(tanks, caps):
        g = sa.SPHContext(variant, p, c, max_triangles=1)
        tune_in_warmup(g, args.warmup, args.lists)
        g.upload(a)
        ctxs.append(g)
    for g in ctxs:
        for _ in range(args.warmup):
            g.step(flags)
    seq_steps = max(1, min(args.steps, 20))
    t0 = time.perf_counter()
    for k in range(seq_steps):
        for g in ctxs:
            g.step(flags)
    seq_wall = time.perf_counter() - t0
    seq_value = ns * seq_steps / seq_wall
    log("one by one: %.4g scene-steps/s (%d steps)" % (seq_value, seq_steps))
    for g in ctxs:
        g.close()

    cpu = None
    if not args.no_cpu:
        from oracle import orc
        cores = host_cores()
        orc.set_threads(cores)
        t_cpu, done = 0.0, 0
        for s_, (p, _) in enumerate(tanks):
            o = orc.Oracle(variant, p, caps[s_])
            o.upload(state0[s_])
            t = time.perf_counter()
            o.step(flags)
            t_cpu += time.perf_counter() - t
            done += 1
            if t_cpu > args.cpu_seconds:
                break
        cpu = {"value": done / t_cpu, "unit": "scene-steps/s", "cores": cores, "kind": "port",
               "sample": "one step of each of the first %d tanks from the GPU state after %d "
                         "warm-up steps, fp32 CPU oracle with %d OpenMP threads on %s, %.1f s" % (
                             done, args.warmup, cores, cpu_model(), t_cpu)}
        log("CPU baseline: %.4g scene-steps/s" % cpu["value"])

    out = {
        "metric": "scene-steps/s @256 batched adaptive 3D SPH tanks",
        "value": value,
        "unit": "scene-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "variant": "SPH-3D-2",
            "tanks": ns,
            "initial_particles": int(sizes.sum()),
            "tank_particles_min_max": [int(sizes.min()), int(sizes.max())],
            "capacity_per_tank": "4 x initial",
            "split_merge": True,
            "live_particles_first_timed_step": n_first,
            "live_particles_last_timed_step": n_last,
            "live_particles_first_last_mean": mean_n,
            "particle_steps_per_s_approx": mean_n * args.steps / wall,
            "parallelism": "single-gpu",
        },
        "one_by_one": {"value": seq_value, "unit": "scene-steps/s", "steps": seq_steps,
                       "speedup_batched": value / seq_value},
        "stream_steps": {"value": stream_value, "unit": "scene-steps/s",
                         "ms_per_step": stream["wall"] * 1e3 / args.steps,
                         "note": "the same batch with the kernels launched one by one, one "
                                 "host synchronisation per step",
                         "graph_speedup": value / stream_value},
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant, args.workload),
            "alg_bytes_per_particle": ALG_BYTES[dominant],
            "mean_launch_ms": stream["mean_launch_ms"],
            "note": "the dominant kernel timed with events in stream steps (a graph step has "
                    "no events inside)",
        },
        "cpu_baseline": cpu,
        "phase_ms_per_step": phase_ms,
        "phase_steps": phase_steps,
        "autotune": {"force_list": tuned[0], "energy_list": tuned[1], "best_ms": tuned[2],
                     "groups_per_cu": gpc, "num_cu": n_cu, "grid_best_ms": grid_ms},
    }
    print(json.dumps(out), flush=True)


def run_rollout(args, w):
    """Differentiable rollout: time of the backward pass (adjoint incl. checkpoint replay)
    over the time of the forward pass, 200 steps of 32^3 particles."""
    r = w["rollout"]
    nf, ns, steps = r["field_nodes"], r["sample_n"], w["steps"]
    params, parts = w["build"]()
    theta = scenes.field_theta(nf, scenes.SEED_BASE + 50)
    theta_star = scenes.field_theta(nf, scenes.SEED_BASE + 51)
    log("workload %s: %d particles, %d^3 x 3 field, %d^3 sample, %d steps" % (
        args.workload, len(parts), nf, ns, steps))
    out_runs = {}
    for ck in (r["checkpoint_every"], 0):
        adj = sa.SPHAdjoint(params, len(parts), nf, ns, checkpoint_every=ck, flags=w["flags"])
        adj.upload(parts)
        target = np.zeros(ns ** 3, dtype=np.float32)
        adj.forward(theta_star, steps, target)
        adj.set_target(target)
        grad = np.zeros_like(theta)
        for _ in range(max(1, args.warmup // 10)):  # warm-up pair
            adj.forward(theta, steps)
            adj.backward(grad)
        t0 = time.perf_counter()
        loss = adj.forward(theta, steps)
        t_f = time.perf_counter() - t0
        t0 = time.perf_counter()
        adj.backward(grad)
        t_b = time.perf_counter() - t0
        ev_f, ev_b = adj.timing()
        topo = adj.topology(steps)
        out_runs[ck] = dict(loss=loss, fwd_ms=t_f * 1e3, bwd_ms=t_b * 1e3, fwd_ms_events=ev_f,
                            bwd_ms_events=ev_b, ratio=t_b / t_f, grad_norm=float(np.linalg.norm(grad)),
                            live_particles_end=adj.num_particles(),
                            splits=int(topo[:, 1].sum()), merge_groups=int(topo[:, 2].sum()))
        log("checkpoint_every %d: forward %.1f ms, backward %.1f ms, ratio %.2f, loss %.6g" % (
            ck, t_f * 1e3, t_b * 1e3, t_b / t_f, loss))
        adj.close()
    ck = r["checkpoint_every"]
    cpu = None
    if not args.no_cpu:
        from oracle import orc
        cores = host_cores()
        orc.set_threads(cores)
        o = orc.Oracle(sa.V3D2, params, 2 * len(parts))
        o.upload(parts)
        o.set_field(theta, nf)
        done, t_cpu = 0, 0.0
        while done < steps and t_cpu < args.cpu_seconds:
            t0 = time.perf_counter()
            o.step(w["flags"])
            t_cpu += time.perf_counter() - t0
            done += 1
        cpu = {"value": len(parts) * done / t_cpu, "unit": "particle-steps/s (forward)",
               "cores": cores, "kind": "port",
               "sample": "%d forward steps of the rollout model (fp32 CPU oracle with the field, "
                         "no adjoint on the CPU), %d OpenMP threads on %s, %.1f s" % (
                             done, cores, cpu_model(), t_cpu)}
    res = out_runs[ck]
    out = {
        "metric": "adjoint/forward time @32k-particle 200-step rollout",
        "value": res["ratio"],
        "unit": "x",
        "n_gpus": 1,
        "steps": steps,
        "warmup": max(1, args.warmup // 10),
        "ms_per_step": res["fwd_ms"] / steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "variant": "SPH-3D-2", "particles": int(len(parts)),
                   "field_nodes": nf, "sample_n": ns, "checkpoint_every": ck,
                   "split_merge": bool(w["flags"]), "time_step": float(params["time_step"][0]),
                   "parallelism": "single-gpu"},
        "forward_ms": res["fwd_ms"],
        "backward_ms": res["bwd_ms"],
        "forward_particle_steps_per_s": len(parts) * steps / (res["fwd_ms"] * 1e-3),
        "no_checkpoint": out_runs[0],
        "checkpointed": res,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="adaptive_3d2_262k", choices=sorted(scenes.WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lists", default=None,
                    help="pin the force,energy pair-list lengths (e.g. 32,16) instead of autotuning")
    ap.add_argument("--groups-per-cu", type=int, default=-1,
                    help="batched workload: pin the graph steps' workgroups per CU (0: launches "
                         "sized by the buffers' bounds; 4, 8, 16: persistent) instead of autotuning")
    ap.add_argument("--no-full-run", dest="full_run", action="store_false",
                    help="skip the run over all of the workload's configured steps")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB 5 times on the device "
                         "(the plain copy kernel that scales the PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would "
              "step its own slab of cells; see the sharding design). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    w = scenes.WORKLOADS[args.workload]
    if w.get("batched"):
        return run_batched(args, w)
    if w.get("rollout"):
        return run_rollout(args, w)
    variant = w["variant"]
    params, parts = w["build"]()
    log("workload %s: %d particles, %d cells" % (args.workload, len(parts), int(params["total_cell"][0])))
    lib = sa.load_library()
    log(lib.sph_version().decode())
    ctx = sa.SPHContext(variant, params, w["capacity"])
    tune_in_warmup(ctx, args.warmup, args.lists)
    ctx.upload(parts)

    # warm-up; its later steps (from the second on, at most the last half) run with every
    # phase timed, to find the kernel with the largest steady-state time per step. The first
    # steps are left out: a kernel's first launch carries one-off costs (code-object loading,
    # cold caches) that say nothing about the timed steps.
    prof_from = min(max(1, args.warmup // 2), max(args.warmup - 1, 0))
    for k in range(args.warmup):
        if k == prof_from:
            ctx.synchronize()
            ctx.prof_reset()
            ctx.prof_enable(True)
        ctx.step(scenes.step_flags(args.workload, k))
    ctx.synchronize()
    prof = ctx.prof_read()
    ctx.prof_enable(False)
    ctx.prof_reset()
    cand = {p: v[0] / v[1] for p, v in prof.items()
            if p in ALG_BYTES and ALG_BYTES[p] > 0 and v[1]}
    dominant = max(cand, key=cand.get) if cand else "force"
    log("warm-up steps %d-%d, ms per launch:" % (prof_from, args.warmup - 1),
        {p: round(v[0] / v[1], 4) for p, v in prof.items() if v[1]})
    tuned = ctx.tuned_lists()
    log("autotuned pair lists: force %d, energy %d" % tuned[:2])

    state0 = ctx.download() if not args.no_cpu else None
    first_timed = args.warmup

    # timed region: only the dominant phase bracketed by HIP events on the library's stream
    # (an event pair per phase would cost the stream a few us each)
    ctx.prof_phases([dominant])
    ctx.synchronize()
    n_live = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        n_before = ctx.num_particles()
        ctx.step(scenes.step_flags(args.workload, first_timed + k))
        n_live.append(n_before)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    prof = ctx.prof_read()
    psteps = float(np.sum(n_live))
    value = psteps / wall
    ms_per_step = wall * 1e3 / args.steps

    dom_ms, dom_launches = prof[dominant]
    mean_n = psteps / args.steps
    alg_bytes = ALG_BYTES[dominant] * mean_n
    achieved = alg_bytes / (dom_ms / dom_launches * 1e-3) / 1e9 if dom_launches else None

    # phase breakdown: further steps after the timed region with every phase timed
    phase_steps = max(1, min(args.steps, 20))
    ctx.prof_reset()
    ctx.prof_enable(True)
    for k in range(phase_steps):
        ctx.step(scenes.step_flags(args.workload, first_timed + args.steps + k))
    ctx.synchronize()
    phase_ms = {p: round(v[0] / phase_steps, 4) for p, v in ctx.prof_read().items() if v[1]}
    log("phases over %d further steps (ms/step):" % phase_steps, phase_ms)

    traffic = measured_traffic(dominant, args.workload)
    mean_launch_ms = dom_ms / dom_launches if dom_launches else None
    valu_frac, valu_insts = valu_fraction(dominant, args.workload, mean_n, mean_launch_ms)

    # the whole configured run (all of its steps, from the seeded start) on a fresh context:
    # the rate over the full trajectory, beside the rate of the timed window
    full = None
    if args.full_run and w.get("steps"):
        full_steps = int(w["steps"])
        g2 = sa.SPHContext(variant, params, w["capacity"])
        g2.upload(parts)
        g2.synchronize()
        n_sum, n_min, n_max = 0, None, 0
        t0 = time.perf_counter()
        for k in range(full_steps):
            n_b = g2.num_particles()
            g2.step(scenes.step_flags(args.workload, k))
            n_sum += n_b
            n_min = n_b if n_min is None else min(n_min, n_b)
            n_max = max(n_max, n_b)
            if (k + 1) % 500 == 0:
                log("full run: step %d, %d live particles, %.2f s" % (
                    k + 1, g2.num_particles(), time.perf_counter() - t0))
        g2.synchronize()
        full_wall = time.perf_counter() - t0
        full = {"steps": full_steps, "value": n_sum / full_wall, "unit": "particle-steps/s",
                "wall_s": full_wall, "ms_per_step": full_wall * 1e3 / full_steps,
                "live_particles_min_max": [int(n_min), int(n_max)],
                "live_particles_mean": n_sum / full_steps,
                "live_particles_end": int(g2.num_particles()),
                "note": "all %d configured steps from the seeded start on a fresh context, "
                        "library-default list tuning, marching cubes every %d steps" % (
                            full_steps, w["mc_every"])}
        g2.close()
        log("full run: %.4g particle-steps/s over %d steps (%.2f s)" % (
            full["value"], full_steps, full_wall))

    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        log("CPU baseline on %d threads ..." % cores)
        cpu = cpu_baseline(args.workload, params, state0, first_timed, args.cpu_seconds, cores)
        log("CPU baseline: %.4g particle-steps/s" % cpu["value"])

    out = {
        "metric": "particle-steps/s @1M adaptive 3D SPH",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "variant": {0: "SPH-2D", 1: "SPH-3D-2", 2: "SPH-3D-4"}[variant],
            "initial_particles": int(len(parts)),
            "capacity": w["capacity"],
            "cells": int(params["total_cell"][0]),
            "density_grid_voxels": int(params["tot_dens"][0]) if variant else 0,
            "mc_every": w["mc_every"],
            "split_merge": bool(w["flags"] & (sa.STEP_SPLIT | sa.STEP_MERGE)),
            "live_particles_first_timed_step": int(n_live[0]),
            "live_particles_last_timed_step": int(n_live[-1]),
            "live_particles_mean": mean_n,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": traffic,
            "alg_bytes_per_particle": ALG_BYTES[dominant],
            "alg_bytes_per_launch": alg_bytes,
            "mean_launch_ms": mean_launch_ms,
            "valu_frac": valu_frac,
            "valu_insts_per_launch": valu_insts,
            "valu_peak_wave_inst_per_s": VALU_PEAK_WAVE_INST_S,
        },
        "full_run": full,
        "cpu_baseline": cpu,
        "phase_ms_per_step": phase_ms,
        "phase_steps": phase_steps,
        "autotune": {"force_list": tuned[0], "energy_list": tuned[1], "best_ms": tuned[2]},
    }
    print(json.dumps(out), flush=True)
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)


if __name__ == "__main__":
    main()
