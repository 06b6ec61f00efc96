#!/usr/bin/env python3
"""Throughput benchmark of the rigid-body step on one MI355X.

Workload (default): the dense stacking pile — 100,000 bodies (thirds spheres /
boxes / capsules) on a 1.5 m lattice over a 20 x 20 m footprint inside four wall
planes, friction 0.6, dt = 1/120 s, 20 solver sweeps.  A "step" is one full
physics step (broadphase, narrowphase, colouring, solve, integration) of the
whole scene with its state resident in HBM.

Prints one JSON line: body-steps/s, the roofline of the dominant kernel (the
persistent sweep kernel k_pgs_own: algorithmic bytes / summed event-timed
duration of its launches during the timed region) and the CPU oracle timed on a
bounded sample of the same workload on the host.  Two further sections: the
1M-body scene iterated to contact residual <= 1e-4 every step, the residual
checked after every sweep and after every fourth (with the oracle at the same
sweep count beside it), and the plain vs learned warm start at that
residual on one 1k-body scene and on 4096 batched worlds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "axiom-physics-engine_amd"))

import numpy as np  # noqa: E402

import axm  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED_BASE = 0x4158494F4D000000
WORKLOADS = {
    0: "mixed spheres/boxes, 1k bodies, ground plane",
    1: "dense stacking pile: 100k spheres/boxes/capsules, 20x20 m footprint, 4 walls + ground",
    2: "sparse 1M spheres/boxes in a 400 m cube above the ground",
    3: "4096 independent worlds, 16-4096 bodies each",
}
DEFAULT_ITERS = {0: 10, 1: 20, 2: 10, 3: 10}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(desc, bodies, iters, dt, budget_s):
    """Times the CPU oracle (one thread) on the given state for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc

    ow = orc.OracleWorld(orc.params_from_desc(desc))
    ow.set_bodies(bodies)
    t1 = ow.run(dt, iters, 1)
    steps = max(1, min(200, int(budget_s / max(t1, 1e-6))))
    t = ow.run(dt, iters, steps)
    total_steps = steps
    return dict(value=bodies.n * total_steps / t, unit="body-steps/s", cores=1, kind="port",
                sample="oracle (C++, 1 thread) continuing from the GPU state after the timed region: "
                       "%d bodies x %d steps (%.1f s)" % (bodies.n, total_steps, t))


def tolerance_comparison(tol, max_sweeps, worlds):
    """Body-steps/s and sweeps per step when every step iterates to residual <= tol (capped at
    max_sweeps): the plain previous-step warm start against the learned one (network trained here on
    the GPU on contacts logged from the plain solver), on one held-out 1k-body scene and on a batch of
    `worlds` independent worlds stepped together."""
    import warmstart

    t0 = time.time()
    net, losses, rows = warmstart.train_on_config0()
    out = {"residual_tol": tol, "max_sweeps": max_sweeps,
           "training": {"rows": rows, "epochs": len(losses), "loss_first": round(float(losses[0]), 5),
                        "loss_last": round(float(losses[-1]), 5), "seconds": round(time.time() - t0, 2),
                        "data": "16 seeded 1k-body scenes x 240 steps, plain solver at 60 sweeps"}}
    log("warm-start network trained on %d contacts in %.1fs" % (rows, time.time() - t0))
    b, d = axm.scene(0, SEED_BASE + 2, 1000)
    b3, d3 = axm.scene(3, SEED_BASE + 3000, 0, worlds)
    cases = (("single_1k_plain", b, d, axm.WARM_CACHE, None, 60, 120, 1),
             ("single_1k_learned", b, d, axm.WARM_MLP, net, 60, 120, 1),
             ("batched_plain", b3, d3, axm.WARM_CACHE, None, 60, 30, 5),
             ("batched_learned", b3, d3, axm.WARM_MLP, net, 60, 30, 5))
    for name, bb, dd, warm, nn, warmup, steps, every in cases:
        tm = {}
        it, reached = warmstart.iterations_to_tolerance(bb, dd, steps, tol, max_sweeps, warm, net=nn, warmup=warmup,
                                                        every=every, timing=tm)
        out[name] = {"bodies": bb.n, "steps": steps, "body_steps_per_s": round(bb.n * steps / tm["s"], 1),
                     "mean_sweeps": round(float(it.mean()), 2), "median_sweeps": float(np.median(it)),
                     "fraction_reached": round(float(reached.mean()), 3), "residual_check_every": every}
        log("%s: %.0f body-steps/s, mean sweeps %.1f, reached %.2f" % (name, out[name]["body_steps_per_s"],
                                                                       it.mean(), reached.mean()))
    net.close()
    return out


def scene_at_residual(cfg, tol, max_sweeps, steps, warmup, cpu_seconds, no_cpu, every=1):
    """Scene `cfg` with every step iterated until its contact residual is <= tol (checked inside
    the sweep launch after every `every` sweeps, capped at max_sweeps): body-steps/s, sweeps per
    step, and the oracle at the GPU's mean sweep count on the state the GPU reached."""
    dt = 1.0 / 120.0
    bodies, desc = axm.scene(cfg, SEED_BASE + 1000 * cfg, 0, 0)
    desc.residual_tol = tol
    desc.max_iterations = max_sweeps
    desc.residual_every = every
    w = axm.World(desc)
    w.set_bodies(bodies)
    for _ in range(warmup):
        w.step(dt, DEFAULT_ITERS[cfg], 0)
    w.sync()
    sweeps, reached = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = w.step(dt, max_sweeps, axm.STEP_RESIDUAL)
        sweeps.append(st.iterations)
        reached.append(st.residual <= tol)
    w.sync()
    el = time.perf_counter() - t0
    out = {"workload": WORKLOADS[cfg], "scene_config": cfg, "bodies": bodies.n, "residual_tol": tol,
           "check_every": every, "max_sweeps": max_sweeps, "warmup": warmup, "steps": steps,
           "body_steps_per_s": round(bodies.n * steps / el, 1), "ms_per_step": round(el * 1e3 / steps, 4),
           "mean_sweeps": round(float(np.mean(sweeps)), 2), "max_sweeps_used": int(np.max(sweeps)),
           "fraction_reached": round(float(np.mean(reached)), 3)}
    log("scene %d at residual %g (checked every %d): %.0f body-steps/s, mean sweeps %.1f" % (
        cfg, tol, every, out["body_steps_per_s"], np.mean(sweeps)))
    if not no_cpu:
        g = w.get_bodies()
        state = bodies.copy()
        state.pos, state.quat, state.lin_vel, state.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
        desc.residual_tol = 0.0
        out["cpu_baseline"] = cpu_baseline(desc, state, max(1, int(round(float(np.mean(sweeps))))), dt,
                                           cpu_seconds)
        out["gpu_over_cpu"] = round(out["body_steps_per_s"] / out["cpu_baseline"]["value"], 2)
    w.close()
    return out


def measured_traffic(cfg, n, iters, args):
    """HBM bytes per sweep-kernel launch from the committed counter runs (rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE, scaled by a plain copy of known size), when they cover this exact workload; else None."""
    path = os.path.join(ROOT, "profiles", "k_pgs_own_pile100k_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except OSError:
        return None
    w = t["workload"]
    if (w["scene_config"], w["bodies"], w["sweeps"], w["warmup"], w["steps"]) != (cfg, n, iters, args.warmup,
                                                                                   args.steps):
        return None
    return t["traffic_bytes_per_launch"]


def copy_calibration(mib, reps):
    """Plain device-to-device copies of a known size (read `mib` MiB, write `mib` MiB per
    dispatch), for scaling the HBM counters FETCH_SIZE / WRITE_SIZE of a counter run.
    Uses PyTorch only (the physics library is not loaded in this mode)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n * 4, "bytes_written": n * 4, "dispatches": reps}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=150)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--bodies", type=int, default=0)
    ap.add_argument("--worlds", type=int, default=0)
    ap.add_argument("--iters", type=int, default=0)
    ap.add_argument("--tol", type=float, default=0.0, help="iterate each step until the residual <= tol")
    ap.add_argument("--max-iters", type=int, default=200)
    ap.add_argument("--tol-every", type=int, default=1, help="with --tol: sweeps between residual checks")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--eager", action="store_true", help="launch every step's kernels instead of replaying a graph")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default sweep launch shape")
    ap.add_argument("--no-tolerance", action="store_true", help="skip the plain vs learned run at residual <= tol")
    ap.add_argument("--no-scene-1m", action="store_true", help="skip the 1M-body scene at residual <= tol")
    ap.add_argument("--tolerance", type=float, default=1e-4)
    ap.add_argument("--tolerance-worlds", type=int, default=4096)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run plain copies of MIB MiB (counter scaling), then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build is single-GPU (every box has one MI355X); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if axm.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    cfg = args.config
    iters = args.iters or DEFAULT_ITERS[cfg]
    dt = 1.0 / 120.0
    bodies, desc = axm.scene(cfg, SEED_BASE + 1000 * cfg, args.bodies, args.worlds)
    if os.environ.get("AXM_SOLVER_MODE"):
        desc.solver_mode = int(os.environ["AXM_SOLVER_MODE"])
    if args.tol > 0:
        desc.residual_tol = args.tol
        desc.max_iterations = args.max_iters
        desc.residual_every = max(1, args.tol_every)
    log("scene config %d: %d bodies" % (cfg, bodies.n))
    w = axm.World(desc)
    if os.environ.get("AXM_SWEEP"):
        g, lds, pol = (int(x, 0) for x in os.environ["AXM_SWEEP"].split(","))
        w.set_sweep_config(g, lds, pol)
    w.set_bodies(bodies)
    t0 = time.time()
    for k in range(args.warmup):
        w.step(dt, iters, 0)
        if (k + 1) % 50 == 0:
            log("warmup %d/%d (%.1fs)" % (k + 1, args.warmup, time.time() - t0))
    w.sync()
    tuned = None
    if not args.no_autotune and args.tol <= 0:
        # tuned on a copy of the warmed-up state, so the timed world's trajectory is untouched
        gs = w.get_bodies()
        tb = bodies.copy()
        tb.pos, tb.quat, tb.lin_vel, tb.ang_vel = gs.pos, gs.quat, gs.lin_vel, gs.ang_vel
        tw = axm.World(desc)
        tw.set_bodies(tb)
        tw.step(dt, iters, 0)
        g, lds, pol, ms = tw.autotune(dt, iters, 5)
        tw.close()
        w.set_sweep_config(g, lds, pol)
        tuned = {"groups": g, "lds_kib": lds, "policy": pol, "sweep_ms": round(ms, 4)}
        log("autotuned sweep: %d workgroups, %d KiB LDS, policy 0x%x (%.3f ms)" % (g, lds, pol, ms))
    w.sync()
    w.profile(True)
    w.sync()
    t0 = time.perf_counter()
    sweeps, reached = [], []
    if args.tol > 0:
        # residual-stopped steps: the sweep launch checks the residual itself; the host reads
        # each step's sweep count and residual once the step is done
        for _ in range(args.steps):
            st = w.step(dt, args.max_iters, axm.STEP_RESIDUAL)
            sweeps.append(st.iterations)
            reached.append(st.residual <= args.tol)
    else:
        st = w.run(dt, iters, args.steps, graph=not args.eager)
    w.sync()
    elapsed = time.perf_counter() - t0
    prof = w.profile(False)
    log("timed %d steps in %.3fs" % (args.steps, elapsed))
    # diagnostics on one extra (untimed) step
    diag = w.step(dt, iters, axm.STEP_RESIDUAL | axm.STEP_TIMINGS)
    value = bodies.n * args.steps / elapsed
    achieved = (prof.pgs_bytes / (prof.pgs_ms * 1e-3) / 1e9) if prof.pgs_ms > 0 else 0.0
    roof = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(cfg, bodies.n, iters, args),
        "kernel": "k_pgs_own",
        "launches": int(prof.pgs_launches),
        "avg_launch_us": round(prof.pgs_ms * 1e3 / max(prof.pgs_launches, 1), 3),
        "bytes_per_launch": round(prof.pgs_bytes / max(prof.pgs_launches, 1), 1),
    }
    out = {
        "metric": "body-steps/sec at contact residual <=1e-4: plain GPU vs batched+learned warm-start",
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded DeterministicRNG scene)",
        "config": {
            "workload": WORKLOADS[cfg], "scene_config": cfg, "bodies": bodies.n, "sweeps": iters,
            "residual_tol": args.tol, "dt": dt,
            "mode": "plain GPU, previous-step warm start" + (", residual-stopped" if args.tol > 0 else ""),
            "launch": "eager" if args.eager else "hipGraph replay",
            "sweep_launch": tuned if tuned else "default (one workgroup per CU)",
        },
        "residual_last_step": diag.residual,
        "last_step": {k: diag.as_dict()[k] for k in ("pairs", "manifolds", "contacts", "colours", "colour_rounds",
                                                    "iterations", "broadphase_ms", "narrowphase_ms", "solver_ms",
                                                    "integration_ms", "total_ms")},
        "roofline": roof,
    }
    w_state = w.get_bodies() if not args.no_cpu else None
    w.close()
    if args.tol > 0:
        out["residual_stopped"] = {"residual_tol": args.tol, "max_sweeps": args.max_iters,
                                   "check_every": max(1, args.tol_every),
                                   "mean_sweeps": round(float(np.mean(sweeps)), 2),
                                   "max_sweeps_used": int(np.max(sweeps)),
                                   "fraction_reached": round(float(np.mean(reached)), 3)}
    if not args.no_scene_1m and not (cfg == 2 and args.tol > 0):
        r1 = scene_at_residual(2, args.tolerance, 300, 20, 20, 10.0, args.no_cpu)
        # the same with the residual checked after every fourth sweep (each check is a grid-wide
        # meeting of the sweep's workgroups); the oracle's rate is the one measured above
        r4 = scene_at_residual(2, args.tolerance, 300, 20, 20, 10.0, True, every=4)
        keep = ("check_every", "body_steps_per_s", "ms_per_step", "mean_sweeps", "max_sweeps_used",
                "fraction_reached")
        r1["checked_every_4"] = {k: r4[k] for k in keep}
        if "cpu_baseline" in r1:
            r1["checked_every_4"]["gpu_over_cpu"] = round(r4["body_steps_per_s"] / r1["cpu_baseline"]["value"], 2)
        out["scene_1m_at_residual"] = r1
    if not args.no_tolerance:
        out["at_tolerance"] = tolerance_comparison(args.tolerance, args.max_iters * 3 // 2, args.tolerance_worlds)
    if not args.no_cpu:
        g = w_state
        state = bodies.copy()
        state.pos, state.quat, state.lin_vel, state.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
        # at a residual tolerance the oracle (the same arithmetic, so the same residuals) runs the
        # GPU's mean sweep count per step
        cpu_iters = max(1, int(round(float(np.mean(sweeps))))) if args.tol > 0 else iters
        out["cpu_baseline"] = cpu_baseline(desc, state, cpu_iters, dt, args.cpu_seconds)
        out["gpu_over_cpu"] = round(value / out["cpu_baseline"]["value"], 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
