"""Benchmark of the SPH step hot path on one MI355X.

Workload (default): the 1M-particle IISPH dam break (100^3 fluid particles at
0.01 m in a closed 4x2x2 m box of 400,002 boundary particles, dt 5e-4,
omega 0.5, stop at 0.1% average density error, <= 100 iterations). A step is
the reference's whole step (neighbour search, density, gravity, integrate
velocity, IISPH setup + solve, integrate velocity, integrate position) on
device-resident data. Prints ONE JSON line.

    python bench.py [--steps K] [--warmup W] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes per launch of each stage (compulsory traffic: each field
# read or written once per pass; neighbour re-reads served on chip), per fluid
# particle and per active boundary particle (boundary particle with a fluid
# neighbour), FP32 fields.
STAGE_BYTES = {
    "iisph_pressure_acceleration": (36, 16),  # x, p, rho, m -> a
    "iisph_pressure": (52, 16),               # x, a, m, rho, s, AA, p -> Ap, p
    "density": (20, 16),                      # x, m -> rho
    # the fused setup pass does the work of density (20), a = g (12),
    # integrate v (40), setup c (32) and setup s/AA (56), three boundary
    # passes, and IISPH iteration 1 without neighbours (p, p/rho^2 and the
    # (x, m p/rho^2) record take the place of the zeros the setup writes; Ap
    # and a = 0 are new: 20)
    "iisph_setup": (180, 48),
    # the fused integration: integrate v (40) + integrate x (36)
    "integrate_position": (76, 0),
    "nsearch_sort": (96, 0),
    "gravity_integrate_v": (40, 0),           # v, a, t_birth -> v
    "wkbb18_diagonal": (56, 16),              # x, rho, m -> sp
    # the diagonal and the CG set-up in one walk (fused steps): x, m, rho, v
    # in; sp, the boundary matrix B, the (x, m/rho) record, x0, r0, y0, p0 out
    "wkbb18_diagonal_setup": (156, 16),
    "gradient_correction": (56, 0),           # x, rho, m -> L (3x3)
    "pcg_spmv": (44, 16),                     # q = A p and p.q, per CG iteration
    "pcg_vector": (156, 0),                   # x, r, y updates, dots, new p
}
# the pt16 solves are scalar systems (one vector component at a time): the
# product reads x, m of the particle, p, the diagonal and writes q
STAGE_BYTES_PT16 = {"pcg_spmv": (28, 0)}


def profiled_traffic(stage):
    """HBM bytes per launch of a stage of the 1M dam break, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by a plain copy
    kernel); None when the summary has no row for it."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_dambreak_1m.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 7 and cells[0] == stage:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dambreak_1m",
                    choices=["dambreak_1m", "dambreak_8k", "viscous_250k", "viscous_gc_250k",
                             "pt16_250k",
                             "served_stream", "checkpoint_1m"])
    ap.add_argument("--duration", type=float, default=30.0,
                    help="served stream: length of the arrival trace in seconds")
    ap.add_argument("--rate", type=float, default=200.0,
                    help="served stream: Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-scenes", type=int, default=512)
    ap.add_argument("--serve-warm-steps", type=int, default=3)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", type=int, default=-1,
                    help="force the schedule mask of the main scene (-1: autotune)")
    ap.add_argument("--batch-schedule", type=int, default=-1,
                    help="force the schedule mask of batched steps (-1: autotune)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="target length of the CPU baseline sample")
    ap.add_argument("--batch-scenes", type=int, default=64,
                    help="scenes of the batched workload (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=100,
                    help="timed steps of the batched workloads (the 64-scene batch runs "
                         "100 steps, as its configuration; the served mix at most 30)")
    ap.add_argument("--batch-viscous-scenes", type=int, default=32,
                    help="scenes of the batched mix with WKBB18 scenes (0 skips it)")
    ap.add_argument("--no-viscous", action="store_true",
                    help="skip the viscous-column sub-run of the default workload")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the served-stream sub-run of the default workload")
    ap.add_argument("--serve-duration", type=float, default=20.0,
                    help="trace length (s) of the served-stream sub-run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (reference kernel "
                         "for scaling the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the step "
              "path of this library uses one GPU", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "served_stream":
        bench_serve(args)
        return
    if args.workload == "checkpoint_1m":
        bench_checkpoint(args)
        return
    from prtcl_amd import hip, scenes

    sc = {"dambreak_1m": scenes.config1, "dambreak_8k": scenes.config0,
          "viscous_250k": scenes.config2, "viscous_gc_250k": scenes.config2,
          "pt16_250k": scenes.config2}[args.workload]()
    if args.workload == "pt16_250k":
        # the same tank with pt16 (vorticity diffusion and velocity
        # reconstruction, the solver settings of prtcl-lua.bunny.lua:150-155)
        sc.implicit_viscosity = 3
        sc.name = "pt16_column_250k"
    if args.workload == "viscous_gc_250k":
        # the same tank with the gradient correction and wkbb18_gc
        sc.implicit_viscosity = 2
        sc.name = "viscous_gc_column_250k"
    sc.permute_every = 8
    ctx = hip.Context(0)
    out, g = bench_scene(ctx, sc, args.steps, args.warmup, schedules=args.schedule < 0,
                         traffic=args.workload == "dambreak_1m", schedule=args.schedule)
    out = {
        "metric": "IISPH particle-steps/s at 1M particles; batched-scene p50/p99 step latency",
        "value": out.pop("value"), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": out.pop("ms_per_step"),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic", **out,
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, g, args.cpu_seconds)
    g.destroy()

    if args.workload == "dambreak_1m":
        if args.batch_schedule >= 0:
            ctx.set_batch_schedule(args.batch_schedule)
        if args.batch_scenes > 0:
            out["batched"] = bench_batch(ctx, args.batch_scenes, args.batch_steps, args.warmup)
        if args.batch_viscous_scenes > 0:
            # the served stream's scene mix: 30 % with WKBB18, stepped batched
            vs = scenes.config4_scenes(args.batch_viscous_scenes)
            out["batched_viscous"] = bench_batch(
                ctx, len(vs), min(args.batch_steps, 30), min(args.warmup, 3), scs=vs,
                name=f"served_mix{len(vs)}")
        if not args.no_viscous:
            # config 2: the 250k viscous column (WKBB18 PCG every step)
            vsc = scenes.config2()
            vsc.permute_every = 8
            v, gv = bench_scene(ctx, vsc, min(args.steps, 20), min(args.warmup, 3))
            gv.destroy()
            out["viscous"] = v
        ctx.close()
        if not args.no_served:
            out["served"] = bench_serve(args, sub=True)
    print(json.dumps(out), flush=True)


def step_bytes(nf, nb_active, k, kv=0.0, viscous=False):
    """Algorithmic HBM bytes of one whole step: 352 + 88 (k - 1) per fluid
    particle and 16 (4 + 2 (k - 1)) per active boundary particle for IISPH
    (iteration 1 needs no neighbour pass: every pressure starts at 0, so it
    costs the 20 bytes of Ap and a it adds to the setup's writes); a WKBB18
    step adds 376 + 200 k_v per fluid particle and 16 (2 + k_v) per active
    boundary particle (diagonal, the set-up product and one product per PCG
    iteration)."""
    kp = max(k - 1.0, 0.0)
    b = (352 + 88 * kp) * nf + 16 * (4 + 2 * kp) * nb_active
    if viscous:
        b += (376 + 200 * kv) * nf + 16 * (2 + kv) * nb_active
    return b


def bench_scene(ctx, sc, steps, warmup, schedules=False, traffic=False, schedule=-1):
    """Whole steps of one device-resident scene: the timed region (K steps,
    no instrumentation), a pass with per-stage HIP events, the roofline of
    the dominant stage and of the whole step. schedule >= 0 forces the
    scene's schedule mask (bit-identical variants) instead of the tuner's
    choice. Returns (dict, scene)."""
    from prtcl_amd import hip
    g = hip.HipScene(ctx, sc)
    g.prepare()
    if schedule >= 0:
        g.set_schedule(schedule)
    for _ in range(warmup):
        g.step(1)
    ctx.synchronize()

    # active boundary particles (with at least one fluid neighbour)
    off, _ = g.neighbors(hip.BOUNDARY, hip.FLUID)
    nb_active = int(np.count_nonzero(np.diff(off)))

    # timed region: K whole steps, no instrumentation
    iters, wk_iters, lat = [], [], []
    ctx.synchronize()
    t_start = time.perf_counter()
    prev = t_start
    for _ in range(steps):
        g.step(1)
        info = g.step_info()
        iters.append(info["iisph_iterations"])
        wk_iters.append(info["wkbb18_iterations"])
        now = time.perf_counter()
        lat.append(now - prev)
        prev = now
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start

    # kernel timing pass: the same steps continued with hipEvents recorded on
    # the library's stream around every stage (the events add small gaps, so
    # this pass is kept out of the timed region above)
    ctx.profile(True)
    prof_steps = min(steps, 20)
    for _ in range(prof_steps):
        g.step(1)
    ctx.synchronize()
    prof = ctx.profile_read()
    ctx.profile(False)
    tuned = g.schedule()

    sched = {}
    if schedules:
        # per-launch stage times of both schedules of the fused passes (global
        # gathers vs LDS-staged slabs), each forced for a few further steps
        for name, mask in (("global", 8), ("lds", 7 | 8)):
            g.set_schedule(mask)
            g.step(1)
            ctx.profile(True)
            for _ in range(min(prof_steps, 10)):
                g.step(1)
            ctx.synchronize()
            pr = ctx.profile_read()
            ctx.profile(False)
            sched[name] = {k: round(v[0] / v[1], 5) for k, v in pr.items()
                           if v[1] > 0 and k in ("iisph_setup", "iisph_pressure_acceleration",
                                                 "iisph_pressure")}
        g.set_schedule(tuned)

    nf = sc.n_fluid
    viscous = sc.implicit_viscosity == 1
    # dominant kernel by total time (pcg_vector is not one kernel: the vector
    # kernels of every CG iteration plus the set-up and apply of the solve)
    stages = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max((k for k in stages if k != "pcg_vector"), key=lambda k: stages[k][0])
    dom_ms, dom_n = stages[dom]
    table = dict(STAGE_BYTES, **STAGE_BYTES_PT16) if sc.implicit_viscosity == 3 else STAGE_BYTES
    per_f, per_b = table.get(dom, (0, 0))
    bytes_launch = per_f * nf + per_b * nb_active
    avg_s = dom_ms / dom_n / 1e3
    achieved = bytes_launch / avg_s / 1e9
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": profiled_traffic(dom) if traffic else None,
        "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_s * 1e3, 5),
        "launches": dom_n,
    }
    k_mean = float(np.mean(iters))
    kv_mean = float(np.mean(wk_iters))
    step_frac = None
    if sc.implicit_viscosity in (0, 1):
        sb = step_bytes(nf, nb_active, k_mean, kv_mean, viscous)
        step_frac = round(sb / (elapsed / steps) / (HBM_PEAK_GBS * 1e9), 5)
    out = {
        "workload": sc.name,
        "value": round(nf * steps / elapsed, 1),
        "unit": "particle-steps/s",
        "steps": steps, "warmup": warmup,
        "ms_per_step": round(1e3 * elapsed / steps, 4),
        "config": {
            "workload": sc.name, "fluid_particles": nf, "boundary_particles": sc.n_boundary,
            "active_boundary_particles": nb_active, "spacing_m": sc.spacing,
            "time_step_s": sc.time_step, "iisph_tolerance": sc.iisph_tolerance,
            "iisph_max_iterations": sc.iisph_max_iterations, "reductions": "f64",
        },
        "iisph_iterations_mean": round(k_mean, 3),
        "wkbb18_iterations_mean": round(kv_mean, 3),
        "p50_step_ms": round(1e3 * float(np.percentile(lat, 50)), 4),
        "p99_step_ms": round(1e3 * float(np.percentile(lat, 99)), 4),
        "step_roofline_frac": step_frac,
        "roofline": roofline,
        "stages_ms_per_step": {k: round(v[0] / prof_steps, 4) for k, v in stages.items()},
    }
    if viscous:
        out["config"].update(dynamic_viscosity=sc.fluid_viscosity,
                             wkbb18_tolerance=sc.wkbb18_maximum_error,
                             wkbb18_max_iterations=sc.wkbb18_maximum_iterations)
    if schedules:
        out["schedule"] = {"tuned_mask": tuned, "ms_per_launch": sched}
    return out, g


def bench_checkpoint(args):
    """Checkpoint and resume of the 1M dam break (model archive of the
    reference's native binary format): archive bytes per second written by
    prtcl_hip_scene_save_archive to a file and to /dev/null (the device gather
    to host order and the copies alone), and read back by
    prtcl_hip_scene_load_archive into a new scene; and the fluid group's VTK
    text (prtcl_hip_scene_save_vtk)."""
    import tempfile
    from prtcl_amd import hip, scenes
    sc = scenes.config1()
    sc.permute_every = 8
    ctx = hip.Context(0)
    g = hip.HipScene(ctx, sc)
    g.prepare()
    for _ in range(max(1, args.warmup)):
        g.step(1)
    ctx.synchronize()
    reps = max(1, min(args.steps, 10))
    d = tempfile.mkdtemp(prefix="prtcl_ckpt_")
    path = os.path.join(d, "scene.bin")
    nbytes = g.save_archive(path)
    g.save_archive(os.devnull)

    def timed(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    t_file = timed(lambda: g.save_archive(path))
    t_null = timed(lambda: g.save_archive(os.devnull))

    def load():
        h = hip.HipScene.load_archive(ctx, path, settings=sc)
        ctx.synchronize()
        h.destroy()

    t_load = timed(load)
    # VTK text of the fluid group (formatted on the device)
    vpath = os.path.join(d, "f.vtk")
    vbytes = g.save_vtk("fluid", vpath)
    t_vtk = timed(lambda: g.save_vtk("fluid", vpath))
    t_vtk_null = timed(lambda: g.save_vtk("fluid", os.devnull))
    # sphere tracing of the fluid at the scripts' 800 x 600 sensor
    cam_args = ((2.5, -3.0, 1.6), (-1.9, 3.5, -1.2), (0, 0, 1), 1.0, 800, 600)
    img, par = g.trace(hip.camera(*cam_args), "fluid")
    t_trace = timed(lambda: g.trace(hip.camera(*cam_args), "fluid"))
    out = {
        "metric": "checkpoint write bytes/s of the 1M dam-break scene (model archive)",
        "value": round(nbytes / t_file, 1),
        "unit": "bytes/s",
        "n_gpus": 1, "steps": reps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_file, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "checkpoint_" + sc.name, "fluid_particles": sc.n_fluid,
                   "boundary_particles": sc.n_boundary, "archive_bytes": nbytes,
                   "file": path},
        "save_ms": round(1e3 * t_file, 3),
        "save_to_devnull_ms": round(1e3 * t_null, 3),
        "save_to_devnull_bytes_per_s": round(nbytes / t_null, 1),
        "load_ms": round(1e3 * t_load, 3),
        "load_bytes_per_s": round(nbytes / t_load, 1),
        "vtk_fluid_bytes": vbytes,
        "vtk_fluid_ms": round(1e3 * t_vtk, 3),
        "vtk_fluid_to_devnull_ms": round(1e3 * t_vtk_null, 3),
        "vtk_fluid_particles_per_s": round(sc.n_fluid / t_vtk, 1),
        "trace_800x600_ms": round(1e3 * t_trace, 3),
        "trace_surface_pixels": int(np.count_nonzero(img != 0.3)),
        "trace_pixels_per_s": round(800 * 600 / t_trace, 1),
    }
    if not args.no_cpu_baseline:
        # the restatement writing the same fields from host arrays
        from oracle import archive as A
        m = A.read_model(open(path, "rb").read())
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0) or n == 0:
            with open(path + ".cpu", "wb") as fh:
                fh.write(A.write_model(m))
            n += 1
        dt = (time.perf_counter() - t0) / n
        out["cpu_baseline"] = {"value": round(nbytes / dt, 1), "unit": "bytes/s", "cores": 1,
                               "kind": "port",
                               "sample": f"{n} writes of the same archive from host arrays "
                                         "by oracle/archive.py"}
        os.remove(path + ".cpu")
        # the restatement of SaveVTK on a bounded sample of the fluid group
        from oracle import vtk as V
        k = 20000
        f = {name: g.download("fluid", name)[:k] for name in
             ("position", "velocity", "acceleration", "mass", "density", "pressure",
              "time_of_birth")}
        t0 = time.perf_counter()
        V.write_vtk("f", "fluid", f)
        dt = time.perf_counter() - t0
        out["cpu_baseline_vtk"] = {"value": round(k / dt, 1), "unit": "particles/s",
                                   "cores": 1, "kind": "port",
                                   "sample": f"the first {k} fluid particles by oracle/vtk.py"}
        # the tracer restatement on a 16 x 12 sensor with the same field of view
        from oracle import pyoracle as P
        xs = g.download("fluid", "position").astype(np.float64)
        h = float(np.float32(sc.smoothing_scale))
        small = list(cam_args[:4]) + [16, 12]
        t0 = time.perf_counter()
        P.sphere_trace(xs, h, P.camera(*small))
        dt = time.perf_counter() - t0
        out["cpu_baseline_trace"] = {"value": round(16 * 12 / dt, 1), "unit": "pixels/s",
                                     "cores": 1, "kind": "port",
                                     "sample": "a 16 x 12 sensor (same camera and scene) "
                                               "by oracle/trace_ref.cpp"}
    os.remove(vpath)
    os.remove(path)
    os.rmdir(d)
    print(json.dumps(out), flush=True)


def bench_serve(args, sub=False):
    """Served stream (config 4): Poisson arrivals of "advance scene s by
    k = 10 steps" over resident scenes with N_f log-uniform in [5k, 500k], 30%
    with implicit viscosity; scenes admitted against the HBM budget, each
    warmed by a few steps before the trace. Reports request latency
    percentiles, the aggregate particle-steps/s of the run and its HBM
    roofline fraction (algorithmic bytes of the fluid-particle terms of every
    step served, with each request's iteration counts taken as those of its
    last step; boundary terms left out, so a lower bound), over the wall time
    and over the time the request loop spent stepping. With sub=True the
    result is returned as a sub-object of the default workload's line."""
    from prtcl_amd import hip, scenes, serving
    duration = args.serve_duration if sub else args.duration
    t0 = time.perf_counter()
    scs = scenes.config4_scenes(args.serve_scenes)
    gen_s = time.perf_counter() - t0
    ctx = hip.Context(0, budget=int(args.budget_gb * 1e9))
    ss = serving.SceneSet(ctx, lambda i: scs[i], warm_steps=args.serve_warm_steps)
    # admission and set-up of every scene before the trace (the scenes are
    # resident and running when the stream starts: each has taken
    # serve_warm_steps steps, past the start-up transient of its first solves)
    t0 = time.perf_counter()
    for i in range(len(scs)):
        ss.get(i)
    ctx.synchronize()
    admit_s = time.perf_counter() - t0
    arr = scenes.poisson_arrivals(4, args.rate, duration, len(scs))
    out = serving.replay(ctx, ss, arr, k=10)
    lat = out["latencies"]
    visc = np.array([scs[i].implicit_viscosity for i in out["latency_scene"]], bool)
    tk = out["tick_ms"]
    nf = sum(s.n_fluid for s in scs)
    nvisc = sum(1 for s in scs if s.implicit_viscosity)
    # algorithmic bytes of the steps served
    its = out["last_step_iterations"]
    req_bytes = 0.0
    for (k, kv), idx in zip(its, out["latency_scene"]):
        n = scs[idx].n_fluid
        req_bytes += 10 * step_bytes(n, 0, float(k), float(kv), bool(scs[idx].implicit_viscosity))
    busy_s = float(tk.sum()) / 1e3
    res = {
        "metric": "served stream: request latency p50/p99 and particle-steps/s",
        "value": round(out["particle_steps"] / out["wall"], 1),
        "unit": "particle-steps/s", "n_gpus": 1, "steps": out["ticks"], "warmup": 0,
        "ms_per_step": round(1e3 * out["wall"] / max(out["ticks"], 1), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "served_stream", "scenes": len(scs), "viscous_scenes": nvisc,
                   "fluid_particles_total": nf, "rate_per_s": args.rate, "k_steps": 10,
                   "trace_s": duration, "hbm_budget_gb": args.budget_gb},
        "requests": out["requests"],
        "p50_request_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_request_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "mean_request_ms": round(1e3 * float(np.mean(lat)), 3),
        "p50_request_ms_inviscid": round(1e3 * float(np.percentile(lat[~visc], 50)), 3),
        "p50_request_ms_viscous": round(1e3 * float(np.percentile(lat[visc], 50)), 3)
        if visc.any() else None,
        "tick_ms_p50": round(float(np.percentile(tk, 50)), 3),
        "tick_ms_p99": round(float(np.percentile(tk, 99)), 3),
        "tick_ms_max": round(float(tk.max()), 3),
        "scenes_per_tick_mean": round(float(out["active"].mean()), 2),
        "roofline_frac_wall": round(req_bytes / out["wall"] / (HBM_PEAK_GBS * 1e9), 5),
        "roofline_frac_busy": round(req_bytes / max(busy_s, 1e-9) / (HBM_PEAK_GBS * 1e9), 5),
        "busy_s": round(busy_s, 3),
        "wall_s": round(out["wall"], 3), "scene_generation_s": round(gen_s, 2),
        "admission_s": round(admit_s, 2), "warm_steps_per_scene": args.serve_warm_steps,
        "device_gb": round(ctx.bytes_in_use() / 1e9, 3),
    }
    ctx.close()
    if sub:
        for key in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype",
                    "data", "warmup"):
            res.pop(key)
        res["workload"] = "served_stream"
        return res
    print(json.dumps(res), flush=True)


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE per dispatch give the counters' scale (KiB per
    byte actually moved) for the traffic figure of the roofline."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes": nbytes, "reps": reps,
                      "read_plus_write_GBs": round(2 * nbytes * reps / el / 1e9, 1)}),
          flush=True)


def bench_batch(ctx, count, steps, warmup, scs=None, name=None):
    """The batched workload: `count` independent dam-break scenes (N_f
    log-uniform in [2k, 200k], uneven aspect ratios, random initial
    velocities, each with its own IISPH convergence at 0.1%), or the given
    scenes (`scs`, e.g. the served mix with WKBB18 scenes). Two identical
    copies are stepped from the same initial state: one with one
    prtcl_hip_step call per scene (the plain per-scene path), one with a
    single call over all scenes (one launch per stage for the whole batch).
    Both do the same work (their results are bit-identical)."""
    from prtcl_amd import hip, scenes
    scs = scs if scs is not None else scenes.config3(count=count)
    nf_total = sum(s.n_fluid for s in scs)
    nvisc = sum(1 for s in scs if s.implicit_viscosity)

    def make():
        hs = [hip.HipScene(ctx, s) for s in scs]
        for h in hs:
            h.prepare()
        return hs

    def timed(hs, together):
        for _ in range(warmup):
            if together:
                ctx.step(hs, 1)
            else:
                for h in hs:
                    h.step(1)
        ctx.synchronize()
        lat, its, wks = [], [], []
        t0 = prev = time.perf_counter()
        for _ in range(steps):
            if together:
                ctx.step(hs, 1)
            else:
                for h in hs:
                    h.step(1)
            now = time.perf_counter()
            lat.append(now - prev)
            prev = now
            info = [h.step_info() for h in hs]
            its.append(max(d["iisph_iterations"] for d in info))
            wks.append(max(d["wkbb18_iterations"] for d in info))
        ctx.synchronize()
        el = time.perf_counter() - t0
        return el, lat, its, wks

    per = make()
    el_p, lat_p, _, _ = timed(per, False)
    for h in per:
        h.destroy()
    bat = make()
    el_b, lat_b, its_b, wks_b = timed(bat, True)
    for h in bat:
        h.destroy()
    v_b = nf_total * steps / el_b
    v_p = nf_total * steps / el_p
    out = {
        "workload": name or f"batch{count}", "scenes": len(scs), "fluid_particles_total": nf_total,
        "steps": steps, "warmup": warmup,
        "value": round(v_b, 1), "unit": "particle-steps/s",
        "per_scene_value": round(v_p, 1), "speedup_vs_per_scene": round(v_b / v_p, 3),
        "p50_step_ms": round(1e3 * float(np.percentile(lat_b, 50)), 4),
        "p99_step_ms": round(1e3 * float(np.percentile(lat_b, 99)), 4),
        "per_scene_p50_step_ms": round(1e3 * float(np.percentile(lat_p, 50)), 4),
        "per_scene_p99_step_ms": round(1e3 * float(np.percentile(lat_p, 99)), 4),
        "max_iisph_iterations_mean": round(float(np.mean(its_b)), 2),
    }
    out["schedule_mask"] = ctx.batch_schedule()
    if nvisc:
        out["viscous_scenes"] = nvisc
        out["max_wkbb18_iterations_mean"] = round(float(np.mean(wks_b)), 2)
    return out


def cpu_baseline(sc, g, seconds):
    """The oracle (CPU/OpenMP restatement of the reference scheme) timed on
    this host for a bounded sample of the same workload: it starts from the
    GPU's state at the end of the timed region and runs whole steps until
    `seconds` of wall time have passed."""
    from oracle import pyoracle
    st = sc
    o = pyoracle.OracleScene(st, fp64_reductions=False)
    o.set("fluid", "position", g.download("fluid", "position"))
    o.set("fluid", "velocity", g.download("fluid", "velocity"))
    o.set("boundary", "position", g.download("boundary", "position"))
    o.set_params(permute_every=8)
    o.L.oracle_set_time(o.h, g.get_global("current_time"))
    o.prepare()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    el = time.perf_counter() - t0
    return {
        "value": round(sc.n_fluid * steps / el, 1), "unit": "particle-steps/s",
        "cores": pyoracle.lib().oracle_num_threads(), "kind": "port",
        "sample": f"{steps} whole steps of {sc.name} from the GPU state after the timed "
                  f"region ({el:.1f} s wall)",
    }


if __name__ == "__main__":
    main()
