"""Learned Poisson initial guess on the GPU: the BF16 MFMA forward and
backward kernels against the CPU restatement (oracle/ora_net.c, itself
pinned against torch autograd and finite differences on the CPU), training
on the solver's own snapshots, and the guess inside the Poisson solve.

Tolerances: the device accumulates in FP32 in a different order than the
FP64 oracle, so an activation lying next to a BF16 rounding boundary can
round the other way (one BF16 ulp, 2^-8 relative).  Outputs and per-layer
gradients are compared to 1e-4 relative L2 (observed: 2e-6 to 4e-6 for the
output, <= 1.4e-7 for the gradients)."""
import numpy as np
import pytest

import oracle
from turbins_amd import Net, Solver, abi, workload as W

pytestmark = pytest.mark.gpu

LAYERS = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (48, 16), (16, 1)]


def _layer_slices():
    out, o = [], 0
    for ci, co in LAYERS:
        out.append(slice(o, o + 27 * ci * co))
        o += 27 * ci * co
        out.append(slice(o, o + co))
        o += co
    return out


@pytest.fixture(scope="module")
def net():
    n = Net(seed=11)
    yield n
    n.close()


@pytest.mark.parametrize("shape", [(8, 8, 16), (12, 8, 8)])
def test_forward_matches_oracle(net, shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(7)
    in3 = rng.standard_normal((nz, ny, nx, 3)).astype(np.float32)
    w = net.get_weights()
    out_g = net.forward(in3)
    out_o = oracle.net_forward(w.astype(np.float64), in3.astype(np.float64), round_bf16=True)
    err = np.linalg.norm(out_g - out_o) / np.linalg.norm(out_o)
    print(f"forward {shape}: rel L2 {err:.2e}")
    assert err < 1e-4


def test_gradient_matches_oracle(net):
    nz, ny, nx = 8, 8, 16
    rng = np.random.default_rng(8)
    in3 = rng.standard_normal((nz, ny, nx, 3)).astype(np.float32)
    tgt = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    w = net.get_weights()
    loss_g, g_g = net.gradient(in3, tgt)
    loss_o, g_o = oracle.net_gradient(w.astype(np.float64), in3.astype(np.float64), tgt.astype(np.float64),
                                      round_bf16=True)
    assert abs(loss_g - loss_o) < 1e-5 * loss_o
    for k, sl in enumerate(_layer_slices()):
        a, b = g_g[sl].astype(np.float64), g_o[sl]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) == 0
            continue
        err = np.linalg.norm(a - b) / nb
        print(f"param block {k}: rel L2 {err:.2e}")
        assert err < 1e-4, (k, err)
    # the three used input channels get gradient, the zero-padded ones none
    wi = g_g[_layer_slices()[0]].reshape(27, 16, 16)
    assert np.abs(wi[:, :3]).max() > 0 and np.abs(wi[:, 3:]).max() == 0


def test_weights_roundtrip(net):
    w = net.get_weights()
    w2 = w * 0.5
    net.set_weights(w2)
    assert np.array_equal(net.get_weights(), w2)
    net.set_weights(w)


def _small_run(net_rec=None, mode=None, net_use=None, steps=6, seed=1, bump=0.3, x_lock=1.0):
    c = W.case("bump_small", nx=64, ny=32, nz=16, bump_height=bump, x_lock=x_lock)
    cfg = W.config_for(c, pois_mode=mode if mode is not None else abi.POIS_MIXED_IR)
    g = Solver(cfg)
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)], seed=seed).items():
        g.upload(k, v)
    if net_rec is not None:
        g.record_snapshots(net_rec)
    if net_use is not None:
        g.attach_net(net_use)
    stats = [g.step(c["dt"]) for _ in range(steps)]
    return g, stats


def test_snapshots_training_and_guess():
    n = Net(seed=3)
    n.reserve(24, 64, 32, 16)
    g, _ = _small_run(net_rec=n, steps=12, seed=2)
    g.close()
    assert n.snapshot_count() == 22          # 24 solves, the first two lack a history
    before = n.eval_snapshots(16, 6)
    l0, l1 = n.train(steps=150, batch=2, crop=(32, 16, 16), lr=2e-3, seed=5)
    after = n.eval_snapshots(16, 6)
    print(f"train loss {l0:.3e} -> {l1:.3e}; snapshot loss {before:.3e} -> {after:.3e}")
    assert np.isfinite(l1) and l1 < l0
    assert after < before
    # the guess inside the solve: same tolerance reached, history respected
    g, st = _small_run(mode=abi.POIS_LEARNED_IR, net_use=n, steps=4, seed=4)
    for s in st:
        for p in s.pois:
            assert p.converged and p.res <= 1e-8
    guess = g.predict_guess("SCRATCH0")
    assert np.isfinite(guess).all()
    g.close()
    n.close()


def test_trained_network_gain():
    """What the committed network gains over the method without it, on a
    held-out 256x128x128 run (bump height / lock position drawn with seeds
    the training never used): Poisson iterations and initial residuals
    starting from phi_{n-1} (the reference's guess), from the Galerkin
    combination of phi_{n-1}, phi_{n-2}, and from that combination plus the
    network's direction.  The learned path must never cost iterations, and
    the run-time gate must close exactly when the saved iterations are worth
    less than the inference time."""
    import os
    from turbins_amd import learn
    n = Net(seed=0)
    n.load(learn.DEFAULT_WEIGHTS)
    res = learn.compare_iterations(n, "learn256", steps=10, seed=2024)
    n.close()
    prev, sub, lrn = res["previous"], res["subspace"], res["learned"]
    print("iterations previous / subspace / learned:", prev["iterations_mean"], sub["iterations_mean"],
          lrn["iterations_mean"], "gate", lrn["gate"])
    # the gain the network adds beyond the Galerkin subspace of past
    # solutions, reported as measured (it may be nothing: the gate then
    # closes and the headline path is Galerkin + mixed)
    gain = sub["iterations_mean"] - lrn["iterations_mean"]
    print(f"iterations saved beyond the subspace guess: {gain:+.3f} per solve "
          f"(subspace guess alone saves {prev['iterations_mean'] - sub['iterations_mean']:+.3f})")
    assert lrn["iterations_mean"] <= prev["iterations_mean"]
    assert lrn["res0_median"] <= prev["res0_median"]
    assert gain >= -0.5          # the network's direction never costs a noticeable iteration
    g = lrn["gate"]
    assert g["trials"] == 4
    # time the saved iterations are worth, priced at the subspace run's own ms per iteration
    worth = g["saved_iterations_per_solve"] * sub["solve_ms_mean"] / max(sub["iterations_mean"], 1.0)
    if g["open"]:
        assert worth > 0.5 * g["net_ms_per_call"]
    else:
        assert worth < 1.5 * g["net_ms_per_call"]
