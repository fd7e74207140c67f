"""Failure handling and the reference's own Helmholtz method, through the
C-ABI: BiCGStab parity, the Gauss-Seidel -> BiCGStab handover, a Poisson
solve that runs out of iterations, a rejected initial guess, a mixed solve
handed over to FP64 GMG-PCG, and the NaN scan.  Every failure is forced from
the host (configuration or inputs); every kernel exits normally."""
import numpy as np
import pytest

from turbins_amd import abi, workload as W

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)


def _pair(name, **over):
    from turbins_amd import Solver
    c = W.case(name)
    cfg = W.config_for(c, **over)
    g = Solver(cfg)
    o = oracle.Oracle(cfg)
    h = W.topography(c)
    g.setup_geometry(h)
    o.setup_geometry(h)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    return c, cfg, g, o


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _helmholtz_case(g, o, dt):
    g.eno_conv(-1)
    for q in range(4):
        o.eno_conv(q)
        g.set_rhs(q, dt)
        o.set_rhs(q, dt)


@pytest.mark.parametrize("name", ["bump_small", "turbidity_small"])
def test_helmholtz_bicgstab_matches_oracle(name):
    """TB_HELM_BICGSTAB, the reference's KSPIBCGS without preconditioner
    (Solver.c:39-86, 269-322): every quantity converges to rtol on the true
    residual and matches the oracle's solution."""
    c, cfg, g, o = _pair(name, helm_rtol=1e-12, helm_maxit=500)
    g.set_helmholtz_method(abi.HELM_BICGSTAB)
    dt = c["dt"]
    _helmholtz_case(g, o, dt)
    for q in range(4):
        sg = g.helmholtz_solve(q, dt)
        so = o.helmholtz_solve(q, dt, rtol=1e-15, maxit=500)
        nm = "UVWC"[q]
        xg = g.download(nm)
        b = o.download("RHS_" + nm)
        ores = np.linalg.norm(b - o.helmholtz_apply(q, dt, xg)) / np.linalg.norm(b)
        err = _rel(xg, o.download(nm))
        print(name, "bicgstab", nm, "its", sg.iterations, "res", sg.res, "oracle-operator res", ores, "rel err", err)
        assert sg.converged and sg.fallback == 0 and g.last_rc == 0
        assert sg.res <= 1e-12 and ores <= 1e-11
        assert err <= 1e-6
    g.close()
    o.close()


def test_helmholtz_gauss_seidel_hands_over_to_bicgstab():
    """With dt = 10 the 1/dt term no longer dominates the rows (the
    off-diagonal sum is ~1.3 against 1/dt = 0.1): the sweep pairs contract
    by more than 0.5 per pair, the Gauss-Seidel path stops and BiCGStab
    finishes the solve (fallback = 1) to the same tolerance."""
    c, cfg, g, o = _pair("bump_small", helm_rtol=1e-12, helm_maxit=800)
    dt = 10.0
    _helmholtz_case(g, o, dt)
    for q in range(4):
        sg = g.helmholtz_solve(q, dt)
        nm = "UVWC"[q]
        xg = g.download(nm)
        b = o.download("RHS_" + nm)
        ores = np.linalg.norm(b - o.helmholtz_apply(q, dt, xg)) / np.linalg.norm(b)
        print("handover", nm, "its", sg.iterations, "fallback", sg.fallback, "res", sg.res, "oracle-operator res", ores)
        assert sg.fallback == 1
        assert sg.converged and sg.res <= 1e-12 and ores <= 1e-11
    g.close()
    o.close()


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR])
def test_poisson_iteration_cap_reports_nonconvergence(mode):
    """pois_maxit = 1: the solve returns status 1 with converged = 0 and the
    residual it reached (the reference warns and continues, Solver.c:326-348);
    a step with the same cap returns status 1, not an error."""
    c, cfg, g, o = _pair("bump_small", pois_rtol=1e-10, pois_maxit=1, pois_mode=mode)
    dt = c["dt"]
    g.poisson_rhs(dt)
    st = g.poisson_solve()
    print("maxit 1 mode", mode, "its", st.iterations, "res", st.res, "rc", g.last_rc)
    assert g.last_rc == 1 and st.converged == 0
    assert 0.0 < st.res < st.res0
    g.step(dt)
    assert g.last_rc == 1
    g.close()
    o.close()


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR])
def test_rejected_guess_falls_back_to_the_previous_solution(mode):
    """tb_poisson_solve_guess: a guess whose residual exceeds that of the
    previous phi is dropped (fallback bit 2), FP64 GMG-PCG solves from the
    previous phi, and the result matches the oracle; a good guess (the
    previous solve's answer for the same right-hand side, perturbed) is kept
    and saves iterations."""
    c, cfg, g, o = _pair("bump_small", pois_rtol=1e-10, pois_mode=mode)
    dt = c["dt"]
    g.poisson_rhs(dt)
    o.poisson_rhs(dt)
    so = o.poisson_solve(rtol=1e-13)
    ref = o.download("PHI")
    first = g.poisson_solve()
    assert first.converged
    # the same right-hand side again from a guess of noise: rejected
    rng = np.random.default_rng(7)
    bad = 10.0 * rng.standard_normal(g.NT) * np.max(np.abs(ref))
    st = g.poisson_solve_guess(bad)
    print("bad guess mode", mode, "fallback", st.fallback, "its", st.iterations, "res", st.res)
    assert st.fallback & 2 and st.converged
    assert _rel(g.download("PHI"), ref) <= 1e-8
    # a guess close to the solution: kept, fewer iterations than from zero
    g.upload("PHI", np.zeros(g.NT))
    zero = g.poisson_solve()
    good = ref * (1.0 + 1e-9 * rng.standard_normal(g.NT))
    g.upload("PHI", np.zeros(g.NT))
    st = g.poisson_solve_guess(good)
    print("good guess mode", mode, "fallback", st.fallback, "its", st.iterations, "from zero", zero.iterations)
    assert st.fallback == 0 and st.converged
    assert st.iterations < zero.iterations
    assert _rel(g.download("PHI"), ref) <= 1e-8
    g.close()
    o.close()


def test_mixed_solve_hands_over_to_fp64():
    """A mixed solve whose refinement passes cannot reduce the residual
    (no FP32 inner iterations allowed) stalls on its first pass; FP64
    GMG-PCG finishes it (fallback bit 1) to the tolerance, and the solution
    matches the oracle."""
    c, cfg, g, o = _pair("bump_small", pois_rtol=1e-10, pois_mode=abi.POIS_MIXED_IR)
    g.set_mixed_inner(1e-5, 0)
    dt = c["dt"]
    g.poisson_rhs(dt)
    o.poisson_rhs(dt)
    st = g.poisson_solve()
    o.poisson_solve(rtol=1e-13)
    print("handover its", st.iterations, "fallback", st.fallback, "res", st.res)
    assert st.fallback & 1 and st.converged and st.res <= 1e-10
    assert _rel(g.download("PHI"), o.download("PHI")) <= 1e-8
    g.close()
    o.close()


def test_nan_scan():
    """A NaN in u (or an Inf in c) at an interior cell: tb_div_max reports
    the cell and fails (Debugger_check_nan, Debugger.c:286-324); a step from
    that state fails too instead of returning numbers."""
    from turbins_amd import TBError
    c, cfg, g, o = _pair("bump_small")
    NX, NY = c["nx"] + 1, c["ny"] + 1
    i, j, k = 20, 25, 7
    u = g.download("U")
    u[(k * NY + j) * NX + i] = np.nan
    g.upload("U", u)
    with pytest.raises(TBError, match=r"NaN or Inf .*\(20,25,7\)"):
        g.div_max()
    with pytest.raises(TBError):
        g.step(c["dt"])
    g.close()
    o.close()
    c, cfg, g, o = _pair("bump_small")
    cc = g.download("C")
    cc[(3 * NY + 30) * NX + 40] = np.inf
    g.upload("C", cc)
    with pytest.raises(TBError, match=r"\(40,30,3\)"):
        g.div_max()
    g.close()
    o.close()
