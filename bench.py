#!/usr/bin/env python3
"""Benchmark of the QG timestep hot path on one MI355X.

Workload: 2048^2 x 10 layers (L = 4000 km, log-depth layers, seeded
exponential stratification, double-gyre wind, dt = 900 s), seeded k^-3
initial condition.  A step is one full AB3 step of the model (PV inversion by
sine transforms and vertical modes, Arakawa Jacobian, dissipation, forcing,
update).  The same initial state is stepped in two forms:

  * FP64 plain        -- every field and pass in FP64
  * FP32 compensated  -- FP32 fields and passes, Kahan-compensated AB3
                         increment (FP64 only for the Helmholtz denominators
                         and the energy/enstrophy sums)

`value` is the FP32 form's grid-point.layer.steps/s (NX*NY*nl*steps/s) with
all inputs resident in HBM, timed with HIP events on the model's stream over
exactly --steps steps after --warmup untimed steps.  The line also carries
the FP64 figure and the ratio, the energy/enstrophy error of the FP32 form
against the FP64 form after --accuracy-steps steps, a roofline entry for the
dominant kernel and a CPU baseline (the FP64 oracle, a C restatement of the
reference step, on the host cores, over a bounded sample).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic HBM bytes per storage point.layer for each pass, in units of the
# element size s (FP64: 8, FP32: 4).  P4 with Kahan adds a read and a write of
# the compensation array.
PASS_WORDS = {"P1": 2, "T1": 2, "P2": 2, "T2": 2, "P3": 2, "P4": 6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run_form(Q, p, modes, q0, prec, steps, warmup, prof_steps, tune=True):
    m = Q.Model(p, prec, modes)
    tuned = m.autotune() if tune else None
    m.set_q(q0)
    m.run(warmup)
    m.synchronize()
    ms = m.time_run(steps)
    # per-pass timing on a few more steps (events around every launch)
    m.profile(True)
    m.run(prof_steps)
    m.profile(False)
    passes = m.profile_read()
    m.close()
    return ms, passes, tuned


def accuracy(Q, p, modes, q0, nsteps):
    out = {}
    for name, prec in (("fp64", Q.FP64), ("fp32", Q.FP32)):
        m = Q.Model(p, prec, modes)
        m.set_q(q0)
        m.run(nsteps)
        m.invert_pv()
        El, Zl, E, Z = m.diagnostics()
        out[name] = (E, Z, m.q, m.psi)
        m.close()
    E64, Z64, q64, p64 = out["fp64"]
    E32, Z32, q32, p32 = out["fp32"]
    return {
        "steps": nsteps,
        "energy_rel_err": abs(E32 - E64) / abs(E64),
        "enstrophy_rel_err": abs(Z32 - Z64) / abs(Z64),
        "q_l2_rel_err": float(np.linalg.norm(q32 - q64) / np.linalg.norm(q64)),
        "psi_l2_rel_err": float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64)),
        "energy_fp64": E64,
        "enstrophy_fp64": Z64,
    }


def cpu_baseline(p, modes, q0, seconds):
    """The FP64 oracle (C restatement of the reference step) on all host cores."""
    import oracle as O
    threads = O.threads(0)
    o = O.Oracle(p)
    o.set_modes(*modes)
    o.set_q(q0)
    t0 = time.perf_counter()
    o.timestep()
    one = time.perf_counter() - t0
    nsteps = max(1, min(20, int(seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    o.run(nsteps)
    el = time.perf_counter() - t0
    o.close()
    work = p.NX * p.NY * p.nl
    return {"value": work * nsteps / el, "unit": "grid-pt*layer*steps/s", "cores": threads, "kind": "port",
            "sample": "%d steps of the same 2048^2 x 10 workload after 1 untimed step, FP64 oracle "
                      "(gcc -O3, OpenMP), %.1f s; CPU: %s" % (nsteps, el, cpu_model())}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nx", type=int, default=2048)
    ap.add_argument("--accuracy-steps", type=int, default=1000)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,fp32")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default tile shapes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the hot path is a single-GPU build (the ensemble and "
              "slab sharding designs are not launched across GPUs); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import qgw_amd as Q
    from qgw_amd import configs

    p = configs.c3()
    if args.nx != 2048:
        p = configs.scaled(p, args.nx)
    log("workload NX=%d nl=%d dt=%g nu4=%g" % (p.NX, p.nl, p.dt_fixed, p.nu_hyper))
    t0 = time.perf_counter()
    modes = Q.eigmode(p)
    psi0, q0 = Q.make_ic(p, modes)
    log("initial condition %.1f s" % (time.perf_counter() - t0))
    work = p.NX * p.NY * p.nl
    g = Q.grid(p)[0]

    res = {}
    tunes = {}
    for form in args.forms.split(","):
        prec = {"fp64": Q.FP64, "fp32": Q.FP32, "fp32plain": Q.FP32_PLAIN}[form]
        log("timing form", form)
        ms, passes, tuned = run_form(Q, p, modes, q0, prec, args.steps, args.warmup, args.profile_steps,
                                     not args.no_tune)
        res[form] = (ms, passes)
        tunes[form] = tuned
        log("  %s: %.3f ms/step" % (form, ms / args.steps), {k: round(v[0] / max(v[1], 1), 4)
                                                          for k, v in passes.items() if v[1]})

    head = "fp32" if "fp32" in res else list(res)[0]
    ms, passes = res[head]
    ms_step = ms / args.steps
    value = work * args.steps / (ms / 1e3)

    # roofline of the dominant kernel of the headline form
    s = 8 if head == "fp64" else 4
    words = dict(PASS_WORDS)
    if head == "fp32":
        words["P4"] = 8
    per_launch = {k: v[0] / v[1] for k, v in passes.items() if v[1] and k in words}
    dom = max(per_launch, key=lambda k: per_launch[k] * passes[k][1])
    alg_bytes = words[dom] * s * work
    achieved = alg_bytes / (per_launch[dom] * 1e-3) / 1e9
    pass_table = {k: {"ms_per_launch": round(per_launch[k], 5),
                      "GBs_algorithmic": round(words[k] * s * work / (per_launch[k] * 1e-3) / 1e9, 1)}
                  for k in per_launch}
    total_alg = sum(words.values()) * s * work
    step_gbs = total_alg / (ms_step * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        prefix = "k_%s<%s" % (dom.lower(), "double" if head == "fp64" else "float")
        if dom in ("T1", "T2"):
            prefix = "k_tproj<%s" % ("double" if head == "fp64" else "float")
        for kname, tv in json.load(open(tpath))["kernels"].items():
            if kname.startswith(prefix):
                traffic = tv["fetch_bytes"] + tv["write_bytes"]
    out = {
        "metric": "grid-pt*layer*steps/s at 2048^2x10, FP32 batched vs FP64 plain, <=1e-4 energy err",
        "value": value,
        "unit": "grid-pt*layer*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if head != "fp64" else "fp64",
        "data": "synthetic (seeded k^-3 initial condition, seeded stratification)",
        "config": {"workload": "2048^2 x 10-layer QG step (L=4000 km, dt=900 s, nu4=2e8, double-gyre wind)",
                   "nx": p.NX, "ny": p.NY, "nl": p.nl, "form": head + ("+kahan" if head == "fp32" else ""),
                   "interior_points": g.n * g.m,
                   "p4_tile": {f: (Q.P4_TILES[t[Q.TUNE_P4_TILE]] if t else "stream") for f, t in tunes.items()}},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg_bytes,
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE, profiles/traffic.json"},
        "step_algorithmic_GBs": step_gbs,
        "step_roofline_frac": step_gbs / HBM_PEAK_GBS,
        "passes": pass_table,
    }
    if "fp64" in res:
        ms64 = res["fp64"][0] / args.steps
        out["fp64_value"] = work / (ms64 * 1e-3)
        out["fp64_ms_per_step"] = ms64
        out["fp32_over_fp64"] = ms64 / ms_step
    if args.accuracy_steps > 0 and "fp64" in res and "fp32" in res:
        log("accuracy leg: %d steps per form" % args.accuracy_steps)
        out["accuracy"] = accuracy(Q, p, modes, q0, args.accuracy_steps)
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(p, modes, q0, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
