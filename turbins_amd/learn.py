"""Training pipeline of the learned Poisson initial guess.

The solver dumps its own (b_n, phi_{n-1}, phi_{n-2}) -> phi_n pairs while it
runs lock-exchange-over-a-bump cases with seeded bump heights U[0.1, 0.4] and
lock positions U[0.5, 2]; the network is then trained on random crops of
those snapshots with the device's own forward/backward kernels and judged by
the Poisson iterations it saves on held-out runs.

    python -m turbins_amd.learn --case learn256 --runs 8 --steps 125 \
        --train-steps 1500 --batch 8 --lr 3e-4 --out turbins_amd/weights/poisson_guess.f32

(the committed weights were made by exactly this command: 1984 snapshots at
256x128x128, 1500 Adam steps of 8 random 64^3 crops, ~33 s of training on
one MI355X)
"""
import argparse
import os
import time

import numpy as np

from . import Net, Solver, abi, workload as W

DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "poisson_guess.f32")


def run_params(seed, r):
    """Bump height U[0.1, 0.4] and lock position U[0.5, 2] of run r."""
    u = 0.5 * (W.uniform_pm1(seed, 2 * r, 2) + 1.0)
    return 0.1 + 0.3 * u[0], 0.5 + 1.5 * u[1]


def make_solver(case, bump, x_lock, mode, seed):
    c = W.case(case, bump_height=bump, x_lock=x_lock)
    g = Solver(W.config_for(c, pois_mode=mode))
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)], seed=seed).items():
        g.upload(k, v)
    return c, g


def generate(net, case, runs, steps, seed, log=print):
    c = W.case(case)
    net.reserve(runs * (2 * steps - 2), c["nx"], c["ny"], c["nz"])
    for r in range(runs):
        bump, x_lock = run_params(seed, r)
        c, g = make_solver(case, bump, x_lock, abi.POIS_MIXED_IR, seed + 100 + r)
        g.record_snapshots(net)
        t0 = time.time()
        for _ in range(steps):
            g.step(c["dt"])
        g.record_snapshots(None)
        g.close()
        log(f"run {r}: bump {bump:.3f} lock {x_lock:.3f}: {net.snapshot_count()} snapshots "
            f"({time.time() - t0:.1f}s)")
    return net.snapshot_count()


def compare_iterations(net, case, steps, seed, runs=1, log=print):
    """Poisson iterations and initial residuals on held-out runs (same seeds
    for every variant): 'previous' starts from phi_{n-1} (the reference's
    guess), 'subspace' from the Galerkin combination of the last solutions
    (learned mode without a network: up to eight), 'learned' adds the network's
    correction as a third direction."""
    res = {}
    variants = (("previous", abi.POIS_MIXED_IR, None), ("subspace", abi.POIS_LEARNED_IR, None),
                ("learned", abi.POIS_LEARNED_IR, net))
    for key, mode, use in variants:
        its, r0, rej, ms = [], [], 0, []
        for r in range(runs):
            bump, x_lock = run_params(seed, 1000 + r)
            c, g = make_solver(case, bump, x_lock, mode, seed + 5000 + r)
            if use is not None:
                g.attach_net(use)
            for k in range(steps):
                st = g.step(c["dt"])
                if k < 2:
                    continue        # no history yet
                for p in st.pois:
                    its.append(p.iterations)
                    r0.append(p.res0)
                    rej += p.fallback == 2
                    ms.append(p.ms)
            gate = g.net_gate() if use is not None else None
            g.close()
        res[key] = {"iterations_mean": float(np.mean(its)), "res0_median": float(np.median(r0)),
                    "solve_ms_mean": float(np.mean(ms)), "guess_rejected": int(rej), "solves": len(its)}
        if gate is not None:
            res[key]["gate"] = gate
        log(f"{key}: {res[key]}")
    res["iterations_saved_by_network"] = res["subspace"]["iterations_mean"] - res["learned"]["iterations_mean"]
    res["iterations_saved_vs_previous"] = res["previous"]["iterations_mean"] - res["learned"]["iterations_mean"]
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="learn256")
    ap.add_argument("--runs", type=int, default=8)
    ap.add_argument("--steps", type=int, default=125)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, nargs=3, default=[64, 64, 64])
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--eval-steps", type=int, default=12)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    a = ap.parse_args(argv)
    p = lambda *s: print(*s, flush=True)
    net = Net(seed=a.seed)
    t0 = time.time()
    n = generate(net, a.case, a.runs, a.steps, a.seed, log=p)
    p(f"{n} snapshots in {time.time() - t0:.1f}s")
    hold = max(1, n // 20)
    base = net.eval_snapshots(n - hold, hold)
    done, chunk = 0, 100
    while done < a.train_steps:
        k = min(chunk, a.train_steps - done)
        lr = a.lr * (0.3 if done >= 0.7 * a.train_steps else 1.0)
        t1 = time.time()
        l0, l1 = net.train(k, a.batch, a.crop, lr=lr, seed=a.seed + done)
        done += k
        p(f"train {done}/{a.train_steps}: loss {l0:.4e} -> {l1:.4e} ({(time.time() - t1) / k * 1e3:.1f} ms/step)")
    after = net.eval_snapshots(n - hold, hold)
    p(f"whole-snapshot loss on the last {hold}: untrained {base:.4e} -> trained {after:.4e}")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out)
    p(f"weights -> {a.out}")
    compare_iterations(net, a.case, a.eval_steps, a.seed, log=p)
    net.close()


if __name__ == "__main__":
    main()
