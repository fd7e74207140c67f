#!/usr/bin/env python3
"""Timesteps/s of the TURBINS hot path on one MI355X.

Workload: lock-exchange gravity current over an immersed Gaussian bump on a
512 x 256 x 128 MAC grid (16 x 1 x 1 domain, Re = 2000, Sc = 1, 2nd-order
ENO, dt = 1e-3), every Poisson solve to a 1e-8 relative residual, every
Helmholtz solve to 1e-12.  Two device paths are timed on the same state:
  plain    -- FP64 GMG-preconditioned CG from the previous solution (the
              reference's initial guess)
  headline -- learned mode: Galerkin guess on the last eight solutions and
              the trained network's direction (the network is dropped at
              run time if its saved iterations cost more than its
              inference), then FP32 CG with FP32 V-cycles inside FP64
              iterative refinement (FP64 GMG-PCG takes over if it stalls)
The JSON line's `value` is the headline path's timesteps/s.  Inputs are
resident in HBM before the timed region; each library step ends with a
device synchronisation, so the wall clock brackets exactly K steps.

Both paths start from one snapshot taken after 30 untimed FP64 steps and
run W untimed then K timed steps from it, so they time the same stretch of
the trajectory.  The seeded start is not solenoidal; its divergence halves
from step to step, and over the first ~25 steps the Poisson solves start
from large residuals that a production run (hundreds of steps) does not
see, so the snapshot is taken after them (at 512 x 256 x 128, measured
with earlier kernels: headline 39.4 / plain 30.2 timesteps/s from step 4,
45.2 / 33.5 from step 30, 45.8 / 34.1 from step 60).

`roofline` reports the instrumented kernel with the largest total time in
a few further headline-path steps run with HIP events on (outside the timed
region): algorithmic bytes per launch / average launch time
from HIP events on the library's stream.  `cpu_baseline` times the FP64
oracle (the reference's algorithm: unpreconditioned BiCGStab, GMG-PCG
Poisson) on the same workload from the same snapshot on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "timesteps/s at 512x256x128, 1e-8 Poisson residual: plain vs learned+mixed"
HBM_PEAK_GBS = 8000.0

# algorithmic bytes per point (each array read or written once per point at
# its storage width)
BYTES_PER_POINT = {
    # ENO as tb_step runs it, with the right-hand sides in the same pass:
    # momentum u, v, w, three flags, three pressure gradients and c in, three
    # right-hand sides out; concentration c, u, v, w and a flag in, one out
    "ENO_MOM": 83, "ENO_CONC": 41, "PROJECT": 115, "BICG_UPD": 56, "PCG_UPD": 48, "POIS_OP": 16,
    "CG32_OP": 8, "CG32_UPD": 24,
}
WEIGHTS = os.path.join(ROOT, "turbins_amd", "weights", "poisson_guess.f32")
# HBM bytes per instrumented launch measured with rocprofv3 FETCH_SIZE /
# WRITE_SIZE on this workload, scaled by a plain copy of known size
# (profiles/hbm_traffic_bump512x256.md); only valid for the kernels as built.
# HELM_GS: the sweep pair, and the first pair of a solve, which also reads
# the last solve's increment
MEASURED_TRAFFIC = {"HELM_GS": 463.8e6, "HELM_GS_FIRST": 552.2e6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


STATE = ["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ", "PHI"]


def snapshot(g):
    return {n: g.download(n) for n in STATE}


def restore(g, snap):
    # uploading PHI also empties the learned mode's solution history
    for n in STATE:
        g.upload(n, snap[n])


def run_path(g, c, snap, steps, warmup, mode, rtol):
    """W untimed and K timed steps from the snapshot: both paths run the same
    stretch of the trajectory (steps W+1 .. W+K after the snapshot)."""
    restore(g, snap)
    g.set_poisson_mode(mode, rtol)
    for _ in range(warmup):
        g.step(c["dt"])
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(g.step(c["dt"]))
    t1 = time.perf_counter()
    return (t1 - t0), stats


def profile_path(g, c, steps):
    """Per-kernel device time from HIP events on the library's stream over a
    few further steps (kept out of the timed region: with events on, the
    V-cycles launch kernel by kernel instead of replaying their graphs)."""
    from turbins_amd import abi
    g.profile(True)
    stats = [g.step(c["dt"]) for _ in range(steps)]
    prof = {}
    for name, kid in abi.K.items():
        ms, n = g.profile_read(kid)
        if n:
            prof[name] = (ms, n)
    g.profile(False)
    return prof, stats


def mixed_vs_fp64(g, c, mode):
    """One step from the same state with the headline Poisson path and with
    the plain FP64 one: relative L2 difference of the fields after the step
    (both solve to the same 1e-8 residual)."""
    from turbins_amd import abi
    names = ["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ", "PHI"]
    state = {n: g.download(n) for n in names}
    g.set_poisson_mode(mode, 1e-8)
    g.step(c["dt"])
    a = {n: g.download(n) for n in names}
    for n in names:
        g.upload(n, state[n])
    g.set_poisson_mode(abi.POIS_FP64_PCG, 1e-8)
    g.step(c["dt"])
    out = {}
    for n in names:
        b = g.download(n)
        den = float(np.linalg.norm(b))
        out[n] = float(np.linalg.norm(a[n] - b) / den) if den > 0 else 0.0
    return out


def roofline(prof, stats, c):
    NT = (c["nx"] + 1) * (c["ny"] + 1) * (c["nz"] + 1)
    ncell = c["nx"] * c["ny"] * c["nz"]
    helm_its = sum(h.iterations for st in stats for stage in st.helm for h in stage)
    # solves that swept (in the steady stepping measured here each starts
    # from the field plus its last increment)
    n_first = sum(1 for st in stats for stage in st.helm for h in stage if h.iterations > 0 and not h.fallback)
    best = None
    for name, (ms, n) in prof.items():
        if name in BYTES_PER_POINT:
            total = BYTES_PER_POINT[name] * NT * n
        elif name == "HELM_OP":
            # 17 B/point (x, status, y) per launch; the rhat.v launch of each
            # BiCGStab iteration also reads rhat (8 B/point)
            total = 17 * NT * n + 8 * NT * helm_its
        elif name == "HELM_GS":
            # one fused red+black sweep pair: x, b and the status byte read
            # once, the new x written once; the first pair of each solve also
            # reads the last solve's increment (FP32, 4 B/point)
            total = 25 * NT * n + 4 * NT * n_first
        elif name == "RBGS0":
            total = 24 * ncell * n      # FP64 fused pair: x, b read, x written
        elif name == "RBGS0_F32":
            total = 12 * ncell * n      # FP32 fused pair: x, b read, x written
        else:
            continue
        if best is None or ms > best[1]:
            best = (name, ms, n, total)
    if best is None:
        return None
    name, ms, n, total = best
    per_launch = total / n
    avg_s = ms / n * 1e-3
    achieved = per_launch / avg_s / 1e9
    traffic = MEASURED_TRAFFIC.get(name) if c["nx"] == 512 and c["ny"] == 256 and c["nz"] == 128 else None
    if name == "HELM_GS" and traffic is not None:
        traffic = ((n - n_first) * traffic + n_first * MEASURED_TRAFFIC["HELM_GS_FIRST"]) / n
    return {"kernel": name, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_s * 1e6, 2), "launches": n}


def _omp_threads(n):
    """Thread count of the oracle's OpenMP loops (libgomp, loaded with it)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(c, g, snap, threads, warmup, reps):
    """The same workload on the FP64 oracle (the reference's algorithm:
    unpreconditioned BiCGStab Helmholtz, GMG-PCG Poisson), from the bench's
    snapshot, with the device's geometry installed (the setup chain is not
    part of a step) and the matrices assembled before the clock starts:
    `warmup` untimed and `reps` timed SSP-RK2 steps on `threads` threads
    (median per step), then one RK stage on one thread."""
    # OpenMP reads the binding when it starts, i.e. with the oracle's first
    # parallel loop
    os.environ.setdefault("OMP_PROC_BIND", "close")
    import oracle
    from turbins_amd import workload as W
    _omp_threads(threads)
    cfg = W.config_for(c, pois_rtol=1e-8)
    o = oracle.Oracle(cfg)
    for q in range(4):
        o.set_geometry(q, g.flags(q), g.ib_table(q))
    for name in STATE:
        o.upload(name, snap[name])
    x = np.zeros(o.NT)
    t = time.perf_counter()
    for q in range(4):
        o.helmholtz_apply(q, c["dt"], x)
    o.poisson_apply(x)
    o.gmg_levels()
    log(f"cpu baseline: assembly {time.perf_counter() - t:.1f}s")
    for _ in range(warmup):
        o.step(c["dt"])
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.step(c["dt"])
        times.append(time.perf_counter() - t0)
    log("cpu baseline: " + ", ".join(f"{v:.2f}" for v in times) + f" s per step on {threads} threads")
    # one stage of ora_step's sequence on one thread
    _omp_threads(1)
    dt = c["dt"]
    t0 = time.perf_counter()
    o.store_old()
    for q in range(4):
        o.eno_conv(q)
    for q in range(4):
        o.set_rhs(q, dt)
    for q in range(4):
        o.helmholtz_solve(q, dt, rtol=cfg.helm_rtol, maxit=cfg.helm_maxit)
    o.poisson_rhs(dt)
    o.poisson_solve(rtol=cfg.pois_rtol, maxit=cfg.pois_maxit)
    o.project(dt)
    stage_1 = time.perf_counter() - t0
    _omp_threads(threads)
    log(f"cpu baseline: one RK stage on 1 thread {stage_1:.1f}s")
    o.close()
    return float(np.median(times)), times, stage_1


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--workload", default="bump512x256")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the V-cycle kernels one by one (for tracing tools that do not follow HIP graphs)")
    ap.add_argument("--settle", type=int, default=30,
                    help="untimed FP64 steps from the seeded start before the snapshot both paths run from "
                         "(the start's non-solenoidal transient decays over the first ~25 steps)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--cpu-warmup", type=int, default=1)
    ap.add_argument("--cpu-reps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the x-slab "
              "decomposition over several GPUs is designed, not built or measured", file=sys.stderr)
        sys.exit(2)

    from turbins_amd import Solver, abi, workload as W
    c = W.case(args.workload)
    cfg = W.config_for(c, pois_rtol=1e-8, helm_rtol=1e-12)
    g = Solver(cfg)
    if args.no_graphs:
        g.set_graphs(False)
    t = time.perf_counter()
    g.setup_geometry(W.topography(c))
    log(f"geometry setup {time.perf_counter() - t:.2f}s, reinit iterations {g.reinit_iterations()}")
    flags = [g.flags(q) for q in range(4)]
    for k, v in W.initial_fields(c, flags).items():
        g.upload(k, v)

    # past the seeded start's transient (its non-solenoidal part decays over
    # the first steps), then one snapshot for both paths
    g.set_poisson_mode(abi.POIS_FP64_PCG, 1e-8)
    for _ in range(args.settle):
        g.step(c["dt"])
    snap = snapshot(g)
    t_plain, st_plain = run_path(g, c, snap, args.steps, args.warmup, abi.POIS_FP64_PCG, 1e-8)
    plain_sps = args.steps / t_plain
    from turbins_amd import Net
    net = None
    if os.path.exists(WEIGHTS):
        net = Net(seed=0)
        net.load(WEIGHTS)
        g.attach_net(net)
    learned = net is not None
    mode = abi.POIS_LEARNED_IR
    t_head, st_head = run_path(g, c, snap, args.steps, args.warmup, mode, 1e-8)
    head_sps = args.steps / t_head
    prof, st_prof = profile_path(g, c, max(2, min(args.steps, 4)))
    gate = g.net_gate() if learned else None
    if gate is not None and gate.get("net_ms_per_call", 0) > 0:
        # MFMA work of one inference: 2 * 27 * Cin * Cout flops per output
        # voxel of every convolution (the head's 16 -> 1 is vector code)
        n0 = c["nx"] * c["ny"] * c["nz"]
        flops = 54 * (16 * 16 + 48 * 16) * n0 + 54 * (16 * 32 + 32 * 32 + 96 * 32) * n0 / 8 \
            + 54 * (32 * 64 + 64 * 64) * n0 / 64
        tf = flops / (gate["net_ms_per_call"] * 1e-3) / 1e12
        gate["mfma_tflops"] = round(tf, 1)
        gate["mfma_frac_of_bf16_dense_peak"] = round(tf / 2500.0, 4)
    mixed_err = mixed_vs_fp64(g, c, mode)

    def pois_summary(stats):
        its = [p.iterations for st in stats for p in st.pois]
        res = [p.res for st in stats for p in st.pois]
        fb = sum(1 for st in stats for p in st.pois if p.fallback & 1)
        r0 = [p.res0 for st in stats for p in st.pois]
        return {"iterations_mean": round(float(np.mean(its)), 2), "max_rel_residual": float(np.max(res)),
                "initial_residual_median": float(np.median(r0)), "fallbacks": fb}

    helm_its = [h.iterations for st in st_head for stage in st.helm for h in stage]

    def phase_ms(stats):
        helm = sum(h.ms for st in stats for stage in st.helm for h in stage) / len(stats)
        pois = sum(p.ms for st in stats for p in st.pois) / len(stats)
        return {"helmholtz": round(helm, 2), "poisson": round(pois, 2)}
    log(f"plain: {plain_sps:.3f} steps/s ({1e3 / plain_sps:.1f} ms/step)  headline: {head_sps:.3f} steps/s "
        f"({1e3 / head_sps:.1f} ms/step)  ratio {head_sps / plain_sps:.3f}")
    log("poisson plain", pois_summary(st_plain), "headline", pois_summary(st_head))
    log("device ms per step: plain", phase_ms(st_plain), "headline", phase_ms(st_head))
    for name, (ms, n) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        log(f"  {name:10s} {n:6d} launches  {ms:9.2f} ms  {1e3 * ms / n:8.1f} us/launch")
    covered = sum(ms for ms, _ in prof.values()) / len(st_prof)
    log(f"  instrumented kernels: {covered:.2f} ms per step of {1e3 / head_sps:.2f} ms wall per timed step")

    rl = roofline(prof, st_prof, c)
    result = {
        "metric": METRIC,
        "value": round(head_sps, 4),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / head_sps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lock-exchange initial state, Gaussian bump topography)",
        "config": {
            "workload": f"{args.workload}: lock exchange over an immersed Gaussian bump, "
                        f"{c['nx']}x{c['ny']}x{c['nz']} cells, domain {c['lx']}x{c['ly']}x{c['lz']}, "
                        f"Re {c['re']}, ENO order 2, dt {c['dt']}",
            "cells": c["nx"] * c["ny"] * c["nz"],
            "cell_updates_per_s": round(head_sps * c["nx"] * c["ny"] * c["nz"], 1),
            "poisson_rtol": 1e-8,
            "helmholtz_rtol": 1e-12,
            "headline_path": "Galerkin guess from the last eight solutions"
                             + (" + network direction while its gate is open" if learned else "")
                             + ", FP32 CG + FP32 V-cycles inside FP64 iterative refinement",
            "learned_guess_loaded": learned,
            # what the timed headline steps ran: the network's direction
            # only while its run-time gate was open
            "headline_effective": ("Galerkin + network + mixed" if gate and gate.get("open")
                                   else "Galerkin + mixed (network gate closed)" if learned
                                   else "Galerkin + mixed (no network loaded)"),
            "network_gate": gate,
            "plain_path": "FP64 GMG-PCG",
            "plain_timesteps_per_s": round(plain_sps, 4),
            "plain_ms_per_step": round(1e3 / plain_sps, 3),
            "headline_over_plain": round(head_sps / plain_sps, 4),
            "poisson_headline": pois_summary(st_head),
            "poisson_plain": pois_summary(st_plain),
            "helmholtz_iterations_mean": round(float(np.mean(helm_its)), 2),
            "helmholtz_method": "red-black Gauss-Seidel sweeps (BiCGStab fallback), rtol 1e-12",
            "solve_ms_per_step_plain": phase_ms(st_plain),
            "headline_vs_fp64_step_rel_l2": mixed_err,
            "solve_ms_per_step_headline": phase_ms(st_head),
            "precision": "FP64 fields/ENO/Helmholtz/outer Poisson; FP32 multigrid in the headline path",
        },
        "roofline": rl,
    }
    if not args.no_cpu:
        secs, times, stage_1 = cpu_baseline(c, g, snap, args.cpu_threads, args.cpu_warmup, args.cpu_reps)
        result["cpu_baseline"] = {
            "value": round(1.0 / secs, 6), "unit": "timesteps/s", "cores": args.cpu_threads, "kind": "port",
            "sample": f"{args.workload} ({c['nx']}x{c['ny']}x{c['nz']}) on the FP64 C oracle (unpreconditioned "
                      f"BiCGStab Helmholtz, GMG-PCG Poisson, OpenMP, OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND', 'unset')}) "
                      f"from the bench's snapshot, device geometry installed: {args.cpu_warmup} untimed and "
                      f"{args.cpu_reps} timed SSP-RK2 steps on {args.cpu_threads} threads, median per step; "
                      "one_core: one RK stage on one thread, doubled",
            "reps_s": [round(v, 3) for v in times],
            "one_core": round(1.0 / (2.0 * stage_1), 6),
            "host_cpus": os.cpu_count(),
        }
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
