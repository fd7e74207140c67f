"""Parity of the gfx950 path (through the C-ABI) against the CPU oracle on
the same seeded inputs: flags and IB tables bit-exact, ENO convective terms
and operators to 1e-12 relative, solves to their tolerances, and 20-step
runs of the lock-exchange case to 1e-6 relative L2."""
import numpy as np
import pytest

from turbins_amd import abi, workload as W

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)


def _pair(name, dims=None, case_over=None, **over):
    from turbins_amd import Solver
    c = W.case(name, **(case_over or {})) if dims is None else \
        W.case(name, nx=dims[0], ny=dims[1], nz=dims[2], **(case_over or {}))
    cfg = W.config_for(c, **over)
    g = Solver(cfg)
    o = oracle.Oracle(cfg)
    h = W.topography(c)
    g.setup_geometry(h)
    o.setup_geometry(h)
    return c, cfg, g, o


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _maxrel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


@pytest.fixture(scope="module", params=["bump_small", "turbidity_small"])
def geo_pair(request):
    c, cfg, g, o = _pair(request.param)
    yield request.param, c, cfg, g, o
    g.close()
    o.close()


def test_geometry_flags_bit_exact(geo_pair):
    name, c, cfg, g, o = geo_pair
    print(name, "reinit iterations gpu/oracle", g.reinit_iterations(), o.reinit_iterations())
    assert g.reinit_iterations() == o.reinit_iterations()
    for nm in ["SDF_C", "SDF_U", "SDF_V", "SDF_W"]:
        a, b = g.download(nm), o.download(nm)
        assert np.array_equal(a, b), (nm, np.max(np.abs(a - b)))
    for q in range(4):
        fg, fo = g.flags(q), o.flags(q)
        assert np.array_equal(fg, fo), (q, np.sum(fg != fo))
        assert np.sum(fg == abi.IMMERSED) > 0


def test_ib_tables_bit_exact(geo_pair):
    name, c, cfg, g, o = geo_pair
    for q in range(4):
        tg, to = g.ib_table(q), o.ib_table(q)
        assert len(tg) == len(to) > 0
        for a, b in zip(tg, to):
            assert (a.i, a.j, a.k, a.n_fluid) == (b.i, b.j, b.k, b.n_fluid)
            assert list(a.fi) == list(b.fi) and list(a.fj) == list(b.fj) and list(a.fk) == list(b.fk)
            assert list(a.coef) == list(b.coef), (q, a.i, a.j, a.k, list(a.coef), list(b.coef))
            assert a.boundary_coef == b.boundary_coef
            assert list(a.image) == list(b.image) and list(a.control) == list(b.control)


@pytest.mark.parametrize("guard", [abi.REINIT_SCALED, abi.REINIT_REFERENCE])
def test_reinit_fine_grid_bit_exact(guard):
    """128x256x16 of the multi-bump topography (2 dy < 0.01): both
    stencil-width guards of the near-interface rule, SDF bit-exact."""
    from turbins_amd import Solver
    c = W.case("turbidity1024", nx=128, ny=256, nz=16)
    cfg = W.config_for(c, reinit_guard=guard)
    g, o = Solver(cfg), oracle.Oracle(cfg)
    h = W.topography(c)
    g.setup_geometry(h)
    o.setup_geometry(h)
    assert g.reinit_iterations() == o.reinit_iterations()
    a, b = g.download("SDF_C"), o.download("SDF_C")
    assert np.array_equal(a, b), np.max(np.abs(a - b))
    if guard == abi.REINIT_SCALED:
        for q in range(4):
            assert np.array_equal(g.flags(q), o.flags(q))
    g.close()
    o.close()


def _random_state(c, seed=11):
    NT = (c["nx"] + 1) * (c["ny"] + 1) * (c["nz"] + 1)
    f = {}
    for a, name in enumerate(["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ"]):
        f[name] = 0.1 * W.uniform_pm1(seed, a * NT, NT)
    # a smooth component so that stencil choices are not all noise-driven
    k, j, i = np.meshgrid(np.arange(c["nz"] + 1), np.arange(c["ny"] + 1), np.arange(c["nx"] + 1), indexing="ij")
    f["U"] += np.sin(0.3 * i).reshape(-1)
    f["C"] += np.tanh(0.2 * (i - 10.0)).reshape(-1)
    return f


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("name,walls,box", [("bump_small", (1, 1, 1, 1, 1, 1), None),
                                            ("bump_small", (1, 1, 0, 0, 0, 0), None),
                                            ("turbidity_small", (1, 1, 1, 1, 1, 1), None),
                                            ("bump_small", (1, 1, 1, 1, 1, 1), (12.0, 1.0, 0.9))])
def test_eno_conv(order, name, walls, box):
    """Power-of-two spacings (16/64, 1/32, 1/16) take the kernel's uniform
    window form inside the box; the 12 x 1 x 0.9 box has spacings 3/16 and
    0.05625 along x and z, which keep the general form there."""
    over = None if box is None else dict(lx=box[0], ly=box[1], lz=box[2])
    c, cfg, g, o = _pair(name, case_over=over, eno_order=order, wall_noslip=walls)
    st = _random_state(c)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    g.eno_conv(-1)
    for q in range(4):
        o.eno_conv(q)
        nm = "CONV_" + "UVWC"[q]
        a, b = g.download(nm), o.download(nm)
        same = np.mean(a == b)
        print(name, order, nm, "bit-identical fraction", same, "max rel", _maxrel(a, b))
        assert _maxrel(a, b) <= 1e-12
        # same operations in the same order (no contraction): bit-identical
        assert same == 1.0, (nm, same)
    g.close()
    o.close()


@pytest.mark.parametrize("name,order", [("turbidity_small", 2), ("bump_small", 3), ("lock64", 1)])
def test_fused_eno_rhs_bit_identical(name, order):
    """tb_eno_conv_rhs (the pass tb_step uses) against tb_eno_conv followed by
    tb_set_rhs for the four quantities: convective terms and right-hand
    sides identical bit for bit (buoyancy on, particles in turbidity)."""
    from turbins_amd import Solver
    c = W.case(name)
    cfg = W.config_for(c, eno_order=order)
    st = _random_state(c, seed=9)
    out = []
    for fused in (False, True):
        g = Solver(cfg)
        g.setup_geometry(W.topography(c))
        for k, v in st.items():
            g.upload(k, v)
        dt = 1e-3
        if fused:
            g.eno_conv_rhs(dt)
        else:
            g.eno_conv(-1)
            for q in range(4):
                g.set_rhs(q, dt)
        out.append({n: g.download(n) for n in ["CONV_U", "CONV_V", "CONV_W", "CONV_C", "RHS_U", "RHS_V", "RHS_W",
                                                  "RHS_C"]})
        g.close()
    for n in out[0]:
        assert np.array_equal(out[0][n], out[1][n]), n
        assert np.any(out[0][n] != 0.0), n


def test_rhs_and_operators():
    c, cfg, g, o = _pair("turbidity_small")
    st = _random_state(c, seed=5)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    dt = 1e-3
    g.eno_conv(-1)
    for q in range(4):
        o.eno_conv(q)
        g.set_rhs(q, dt)
        o.set_rhs(q, dt)
        nm = "RHS_" + "UVWC"[q]
        assert _maxrel(g.download(nm), o.download(nm)) <= 1e-12
        x = st["C"] + 0.5 * st["U"]
        yg, yo = g.helmholtz_apply(q, dt, x), o.helmholtz_apply(q, dt, x)
        print("helmholtz", q, "identical fraction", np.mean(yg == yo))
        assert _maxrel(yg, yo) <= 1e-14
    x = st["V"]
    assert _maxrel(g.poisson_apply(x), o.poisson_apply(x)) <= 1e-14
    g.close()
    o.close()


@pytest.mark.parametrize("name,dims", [("bump_small", None), ("turbidity_small", None),
                                       ("turbidity_small", (192, 64, 40))])
@pytest.mark.parametrize("pairs", [1, 3])
def test_helmholtz_gs_sweeps_match_oracle(name, dims, pairs):
    """The fused single-pass red-black sweep pair against the oracle's two
    in-place colour passes on the assembled rows; 65x33x17 points leave a
    one-column x tile, a one-row y tile and a one-plane k chunk, and
    193x65x41 has tiles whose halo lies wholly inside the box beside
    edge tiles."""
    c, cfg, g, o = _pair(name, dims)
    dt = 1e-3
    NT = o.NT
    for q in range(4):
        b = W.uniform_pm1(21 + q, 0, NT)
        x0 = W.uniform_pm1(31 + q, 0, NT)
        xg = g.helmholtz_gs_sweeps(q, dt, b, x0, pairs)
        xo = o.helmholtz_gs_sweeps(q, dt, b, x0, pairs)
        err = _maxrel(xg, xo)
        print(name, "q", q, "pairs", pairs, "max rel diff", err, "identical fraction", np.mean(xg == xo))
        assert err <= 1e-13
    g.close()
    o.close()


def test_helmholtz_solve_matches_oracle():
    c, cfg, g, o = _pair("bump_small", helm_rtol=1e-12)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    dt = c["dt"]
    g.eno_conv(-1)
    for q in range(4):   # all explicit terms from the same state
        o.eno_conv(q)
    for q in range(4):
        g.set_rhs(q, dt)
        o.set_rhs(q, dt)
        sg = g.helmholtz_solve(q, dt)
        so = o.helmholtz_solve(q, dt, rtol=1e-15, maxit=500)
        nm = "UVWC"[q]
        xg = g.download(nm)
        err = _rel(xg, o.download(nm))
        # residual of the device solution in the oracle's assembled system
        b = o.download("RHS_" + nm)
        ores = np.linalg.norm(b - o.helmholtz_apply(q, dt, xg)) / np.linalg.norm(b)
        print("helmholtz", nm, "its", sg.iterations, "res", sg.res, "oracle-operator res", ores,
              "oracle res", so.res, "rel err", err)
        assert sg.converged and sg.res <= 1e-11 and ores <= 1e-11
        # ||x - x*|| <= cond(A) ||r|| / ||x||: the IB rows make this ~1e4-1e5 x res
        assert err <= 1e-6
    g.close()
    o.close()


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR])
def test_poisson_solve_matches_oracle(mode):
    c, cfg, g, o = _pair("bump_small", pois_rtol=1e-10, pois_mode=mode)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    dt = c["dt"]
    g.poisson_rhs(dt)
    o.poisson_rhs(dt)
    assert _maxrel(g.download("PRHS"), o.download("PRHS")) == 0.0
    sg = g.poisson_solve()
    so = o.poisson_solve(rtol=1e-13)
    err = _rel(g.download("PHI"), o.download("PHI"))
    print("poisson mode", mode, "its", sg.iterations, "vcycles", sg.vcycles, "res", sg.res, "rel err", err,
          "fallback", sg.fallback)
    assert sg.converged and sg.res <= 1e-10
    assert err <= 1e-8
    # projection and divergence on identical phi
    phi = o.download("PHI")
    g.upload("PHI", phi)
    g.project(dt)
    o.project(dt)
    for nm in ["U", "V", "W", "DPDX", "DPDY", "DPDZ"]:
        assert _maxrel(g.download(nm), o.download(nm)) == 0.0, nm
    dg, ig = g.div_max()
    do, io = o.div_max()
    assert dg == do and ig == io
    g.close()
    o.close()


def test_diagnostics_match_oracle():
    """Kinetic energy and suspended mass over FLUID cells after a step of
    the bump case (IB cells excluded): FP64 sums in another order, 1e-12."""
    c, cfg, g, o = _pair("turbidity_small", helm_rtol=1e-13, pois_rtol=1e-11)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    g.step(c["dt"])
    for n in ("U", "V", "W", "C"):
        o.upload(n, g.download(n))
    dg, do = g.diagnostics(), o.diagnostics()
    print("diagnostics", dg, do)
    for key in dg:
        assert do[key] > 0.0
        assert abs(dg[key] - do[key]) <= 1e-12 * abs(do[key]), key
    particle = cfg.conc_type == abi.PARTICLE
    sg, so = g.stokes_dissipation(0.7), o.stokes_dissipation(particle, cfg.v_settle, 0.7)
    print("stokes dissipation", particle, sg, so)
    assert particle and so > 0.0
    assert abs(sg - so) <= 1e-12 * so
    # sedimentation rate: per-column values and their ordered sum, bit for bit
    (tg, bg), (to, bo) = g.sedimentation_rate(), o.sedimentation_rate(particle, cfg.v_settle)
    print("sedimentation rate", tg, to)
    assert np.array_equal(bg, bo) and tg == to and to > 0.0
    g.close()
    o.close()


def test_real_pressure_matches_oracle():
    """Pressure_compute_real_pressure (Pressure.c:590-685): the right-hand
    side from the pressure gradients bit for bit, then the Poisson solve into
    phi to the tolerance, as the oracle's solve."""
    c, cfg, g, o = _pair("bump_small", pois_rtol=1e-10)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    # gradients that vanish on their own box-boundary faces: zero net flux,
    # so the Neumann problem is consistent (as for a physical pressure)
    rng = np.random.default_rng(21)
    for q, name in enumerate(("DPDX", "DPDY", "DPDZ")):
        a = 0.1 * rng.standard_normal(o.NT)
        a[o.flags(q) == abi.BOUNDARY] = 0.0
        g.upload(name, a)
        o.upload(name, a)
    g.real_pressure_rhs()
    o.real_pressure_rhs()
    assert np.array_equal(g.download("PRHS"), o.download("PRHS"))
    st = g.real_pressure()
    so = o.poisson_solve(rtol=1e-13)
    assert st.converged and st.res <= 1e-10
    err = _rel(g.download("PHI"), o.download("PHI"))
    print("real pressure its", st.iterations, "res", st.res, "rel err", err)
    assert err <= 1e-8
    g.close()
    o.close()


@pytest.mark.parametrize("name,steps", [("lock64", 20), ("bump_small", 4), ("turbidity_small", 3)])
def test_steps_match_oracle(name, steps):
    c, cfg, g, o = _pair(name, helm_rtol=1e-13, pois_rtol=1e-11)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    dt = c["dt"]
    for s in range(steps):
        sg = g.step(dt)
        so = o.step(dt)
    worst = 0.0
    for nm in ["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ", "PHI"]:
        e = _rel(g.download(nm), o.download(nm))
        print(name, nm, "rel L2 after", steps, "steps:", e)
        worst = max(worst, e)
    assert worst <= 1e-6
    print("div max gpu/oracle", sg.div_max, so.div_max)
    g.close()
    o.close()


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR])
def test_poisson_on_a_grid_that_stops_halving(mode):
    """60 x 30 x 15 cells over a bump: coarsening stops at 15 x 15 x 15
    (3375 cells, too many for the dense coarse inverse), which takes 50 + 50
    symmetric sweep pairs instead (the oracle's rule).  The solve reaches its
    tolerance and the oracle's solution; three steps follow the oracle."""
    c, cfg, g, o = _pair("bump_small", dims=(60, 30, 15), case_over=dict(lx=8.0), pois_rtol=1e-10, pois_mode=mode,
                         helm_rtol=1e-13)
    assert o.gmg_levels()[-1] == (15, 15, 15)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    dt = c["dt"]
    g.poisson_rhs(dt)
    o.poisson_rhs(dt)
    sg = g.poisson_solve()
    o.poisson_solve(rtol=1e-13)
    err = _rel(g.download("PHI"), o.download("PHI"))
    print("odd grid, mode", mode, "its", sg.iterations, "vcycles", sg.vcycles, "res", sg.res, "rel err", err)
    assert sg.converged and sg.res <= 1e-10
    assert err <= 1e-8
    for k, v in init.items():
        g.upload(k, v)
        o.upload(k, v)
    for _ in range(3):
        st = g.step(dt)
        assert all(p.converged for p in st.pois)
        o.step(dt)
    for nm in ["U", "V", "W", "C"]:
        e = _rel(g.download(nm), o.download(nm))
        print(nm, "rel L2 after 3 steps", e)
        assert e <= 1e-6
    g.close()
    o.close()
