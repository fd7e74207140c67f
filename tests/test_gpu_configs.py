"""The full-size workloads run end to end on one GPU with every solve
converged:
  512x128x128 lock exchange over a bump, FP64 GMG-PCG Poisson, 200 steps;
  1024x256x128 particle-laden turbidity current over 8 bumps (settling
  velocity 0.02), learned guess + FP32-in-FP64 refinement with the FP64
  fallback, 100 steps.
No oracle comparison at these sizes (the oracle needs ~10 s per step); the
checks are the solver's own: every Poisson solve at its 1e-8 true relative
residual, every Helmholtz solve at 1e-12, finite fields, divergence of the
projected velocity at the level the Poisson residual allows."""
import os
import time

import numpy as np
import pytest

from turbins_amd import Net, Solver, abi, workload as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "turbins_amd", "weights",
                       "poisson_guess.f32")


def _run(name, mode, steps, net=None):
    c = W.case(name)
    g = Solver(W.config_for(c, pois_rtol=1e-8, helm_rtol=1e-12, pois_mode=mode))
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)]).items():
        g.upload(k, v)
    if net is not None:
        g.attach_net(net)
    t0 = time.perf_counter()
    worst_p, worst_h, divs, fb = 0.0, 0.0, [], 0
    for s in range(steps):
        st = g.step(c["dt"])
        for p in st.pois:
            assert p.converged, (name, s, p.res)
            worst_p = max(worst_p, p.res)
            fb += p.fallback & 1
        for stage in st.helm:
            for h in stage:
                assert h.converged, (name, s, h.res)
                worst_h = max(worst_h, h.res)
        divs.append(st.div_max)
    t = time.perf_counter() - t0
    for n in ("U", "V", "W", "C"):
        assert np.isfinite(g.download(n)).all()
    bytes_ = g.device_bytes()
    g.close()
    print(f"{name}: {steps} steps in {t:.1f}s ({steps / t:.2f} steps/s), max Poisson res {worst_p:.1e}, "
          f"max Helmholtz res {worst_h:.1e}, div max first/last step {divs[0]:.2e}/{divs[-1]:.2e}, "
          f"FP64 fallbacks {fb}, device {bytes_ / 2**30:.1f} GiB")
    return divs


def _check_div(divs):
    # the seeded initial velocity is not solenoidal; the RK average halves
    # its divergence every step, after which the Poisson residual bounds it
    assert divs[-1] < 1e-3 * divs[0] and divs[-1] < 1e-4


def test_bump512x128_plain_200_steps():
    _check_div(_run("bump512x128", abi.POIS_FP64_PCG, 200))


def test_turbidity1024_learned_mixed_100_steps():
    net = Net(seed=0)
    net.load(WEIGHTS)
    divs = _run("turbidity1024", abi.POIS_LEARNED_IR, 100, net)
    net.close()
    _check_div(divs)
