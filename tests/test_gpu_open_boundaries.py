"""Open x faces on the GPU, through the C-ABI, against the oracle on the
same inputs: inflow profiles and planes bit for bit (the influx, a sum,
to 1e-14); the convective outflow update bit for bit for v, w and the
concentration fields and to 1e-14 for u, whose mass correction sums in
another order; ENO convective terms and right-hand sides with random planes
bit for bit (every order, the fused second-order pass included); the
Helmholtz rows and sweeps, the Poisson right-hand side and the divergence at
the open faces; open-channel steps over the bump to 1e-6."""
import numpy as np
import pytest

from turbins_amd import abi, workload as W

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)

IN, OUT = abi.PLANE_INFLOW, abi.PLANE_OUTFLOW


def _pair(name, open_kw, **over):
    from turbins_amd import Solver
    c = W.case(name)
    cfg = W.config_for(c, **over)
    g, o = Solver(cfg), oracle.Oracle(cfg)
    h = W.topography(c)
    g.setup_geometry(h)
    o.setup_geometry(h)
    for x in (g, o):
        x.set_open_boundaries(**open_kw)
    return c, cfg, g, o


def _maxrel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def _random_state(c, seed):
    NT = (c["nx"] + 1) * (c["ny"] + 1) * (c["nz"] + 1)
    f = {}
    for a, name in enumerate(["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ"]):
        f[name] = 0.1 * W.uniform_pm1(seed, a * NT, NT)
    k, j, i = np.meshgrid(np.arange(c["nz"] + 1), np.arange(c["ny"] + 1), np.arange(c["nx"] + 1), indexing="ij")
    f["U"] += np.sin(0.3 * i).reshape(-1)
    f["C"] += np.tanh(0.2 * (i - 10.0)).reshape(-1)
    return f


def _planes(c, g, o, seed):
    """random planes on both faces of every quantity (and every field)"""
    NP = (c["ny"] + 1) * (c["nz"] + 1)
    rng = np.random.default_rng(seed)
    vals = {}
    for q in range(4):
        for side in (IN, OUT):
            if side == IN and q in (abi.Q_V, abi.Q_W):
                continue
            vals[(side, q)] = rng.uniform(-1, 1, NP) + (1.0 if q == abi.Q_U else 0.0)
    for x in (g, o):
        for (side, q), v in vals.items():
            x.set_plane(side, q, v)
    return vals


def test_inflow_profile_and_planes_bit_identical():
    c, cfg, g, o = _pair("bump_small", dict(inflow=True, outflow=True, transient_inflow=True, time_c=0.035))
    print("inflow profile", flush=True)
    fields = [dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.01), dict(type=abi.SALINITY, pe=cfg.pe)]
    for x in (g, o):
        x.set_conc_fields(fields)
    for t in (0.0, 0.02, 1.0):
        fg, fo = g.inflow_profile(t), o.inflow_profile(t)
        assert abs(fg - fo) <= 1e-14 * abs(fo) and fo > 0
        assert np.array_equal(g.get_plane(IN, abi.Q_U), o.get_plane(IN, abi.Q_U))
        for n in (0, 1):
            g.select_conc(n)
            o.select_conc(n)
            assert np.array_equal(g.get_plane(IN, abi.Q_C), o.get_plane(IN, abi.Q_C))
        g.select_conc(0)
        o.select_conc(0)
    p = np.linspace(0.0, 2.0, (c["ny"] + 1) * (c["nz"] + 1))
    g.set_plane(IN, abi.Q_U, p)
    o.set_plane(IN, abi.Q_U, p)
    assert abs(g.influx() - o.influx()) <= 1e-14 * abs(o.influx())
    g.close()
    o.close()


def test_outflow_convective_update():
    c, cfg, g, o = _pair("bump_small", dict(inflow=True, outflow=True, u_conv=0.8))
    print("outflow convective", flush=True)
    fields = [dict(type=abi.SALINITY, pe=cfg.pe), dict(type=abi.SALINITY, pe=0.5 * cfg.pe)]
    for x in (g, o):
        x.set_conc_fields(fields)
        x.inflow_profile(0.0)
    st = _random_state(c, 3)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    for x in (g, o):
        x.select_conc(1)
        x.upload("C", 0.5 * st["C"] + 0.25)
        x.select_conc(0)
    _planes(c, g, o, 5)
    for dt in (1e-3, 2e-3):
        g.outflow_convective(dt)
        o.outflow_convective(dt)
        for q in (abi.Q_V, abi.Q_W):
            assert np.array_equal(g.get_plane(OUT, q), o.get_plane(OUT, q)), q
        for n in (0, 1):
            g.select_conc(n)
            o.select_conc(n)
            assert np.array_equal(g.get_plane(OUT, abi.Q_C), o.get_plane(OUT, abi.Q_C)), n
        g.select_conc(0)
        o.select_conc(0)
        ug, uo = g.get_plane(OUT, abi.Q_U), o.get_plane(OUT, abi.Q_U)
        print("u plane max diff", np.max(np.abs(ug - uo)))
        assert np.max(np.abs(ug - uo)) <= 1e-14
    g.close()
    o.close()


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("name", ["bump_small", "turbidity_small"])
def test_open_eno_and_rhs_bit_identical(order, name):
    """ENO with the open-face ghosts, transposes and cell centres, and the
    right-hand sides with the open faces' terms: tb_eno_conv_rhs (the
    fused second-order pass at order 2) against the oracle's eno_conv and
    set_rhs, bit for bit."""
    c, cfg, g, o = _pair(name, dict(inflow=True, outflow=True), eno_order=order)
    print("open eno", name, order, flush=True)
    st = _random_state(c, 7)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    _planes(c, g, o, 11)
    dt = 1e-3
    g.eno_conv_rhs(dt)
    for q in range(4):
        o.eno_conv(q)
    for q in range(4):
        o.set_rhs(q, dt)
    for q in range(4):
        for pre in ("CONV_", "RHS_"):
            nm = pre + "UVWC"[q]
            a, b = g.download(nm), o.download(nm)
            assert np.array_equal(a, b), (nm, np.mean(a == b), _maxrel(a, b))
    g.close()
    o.close()


def test_open_closed_eno_differ_only_next_to_the_faces():
    """The planes change the convective terms only within the ENO
    stencil's reach of the x faces (and the right-hand sides only there)."""
    c, cfg, g, o = _pair("bump_small", dict(inflow=True, outflow=True), eno_order=2)
    from turbins_amd import Solver
    closed = Solver(cfg)
    closed.setup_geometry(W.topography(c))
    st = _random_state(c, 2)
    for k, v in st.items():
        g.upload(k, v)
        closed.upload(k, v)
    _planes(c, g, o, 4)
    g.eno_conv_rhs(1e-3)
    closed.eno_conv_rhs(1e-3)
    shape = (c["nz"] + 1, c["ny"] + 1, c["nx"] + 1)
    for nm in ("CONV_U", "CONV_V", "CONV_W", "CONV_C", "RHS_U", "RHS_V", "RHS_W", "RHS_C"):
        d = (g.download(nm) != closed.download(nm)).reshape(shape)
        cols = np.nonzero(d.any(axis=(0, 1)))[0]
        assert len(cols) > 0 and np.all((cols <= 2) | (cols >= c["nx"] - 2)), (nm, cols)
    closed.close()
    g.close()
    o.close()


def test_open_helmholtz_rows_and_sweeps():
    c, cfg, g, o = _pair("turbidity_small", dict(inflow=True, outflow=True))
    print("open helmholtz", flush=True)
    dt = 1e-3
    NT = o.NT
    for q in range(4):
        x = W.uniform_pm1(41 + q, 0, NT)
        yg, yo = g.helmholtz_apply(q, dt, x), o.helmholtz_apply(q, dt, x)
        assert _maxrel(yg, yo) <= 1e-14, q
        b = W.uniform_pm1(51 + q, 0, NT)
        xg = g.helmholtz_gs_sweeps(q, dt, b, x, 2)
        xo = o.helmholtz_gs_sweeps(q, dt, b, x, 2)
        assert _maxrel(xg, xo) <= 1e-13, q
    g.close()
    o.close()


def test_open_poisson_rhs_divergence_and_dissipation():
    c, cfg, g, o = _pair("bump_small", dict(inflow=True, outflow=True))
    print("open poisson rhs", flush=True)
    st = _random_state(c, 13)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    _planes(c, g, o, 17)
    g.poisson_rhs(1e-3)
    o.poisson_rhs(1e-3)
    assert np.array_equal(g.download("PRHS"), o.download("PRHS"))
    dg, ig = g.div_max()
    do, io = o.div_max()
    assert dg == do and tuple(ig) == tuple(io)
    vg, vo = g.viscous_dissipation(), o.viscous_dissipation()
    assert abs(vg - vo) <= 1e-12 * abs(vo)
    g.close()
    o.close()


@pytest.mark.parametrize("name,steps", [("bump_small", 4), ("turbidity_small", 3)])
def test_open_channel_steps_match_oracle(name, steps):
    """Inflow at x = 0 (the steady profile), convective outflow at x = Lx,
    from a plug flow u = 1 over the bump: fields, pressure gradients, the
    outflow planes and the divergence after each step (the RK average
    halves the start's divergence at the bump every step) follow the
    oracle."""
    c, cfg, g, o = _pair(name, dict(inflow=True, outflow=True, u_conv=1.0), helm_rtol=1e-13, pois_rtol=1e-11)
    print("open channel", name, flush=True)
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    fl = o.flags(abi.Q_U)
    init["U"] = np.where(fl == abi.FLUID, 1.0, 0.0)
    for x in (g, o):
        for k, v in init.items():
            x.upload(k, v)
        x.inflow_profile(0.0)
        x.set_plane(OUT, abi.Q_U, x.get_plane(IN, abi.Q_U))
        x.set_plane(IN, abi.Q_C, np.full((c["ny"] + 1) * (c["nz"] + 1), 0.5))
    dt = c["dt"]
    for s in range(steps):
        sg = g.step(dt)
        so = o.step(dt)
        print("step", s, "div", sg.div_max, so.div_max)
        assert abs(sg.div_max - so.div_max) <= 1e-6 * so.div_max + 1e-10
    worst = 0.0
    for nm in ["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ"]:
        a, b = g.download(nm), o.download(nm)
        e = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        print(name, nm, "rel L2", e)
        worst = max(worst, e)
    assert worst <= 1e-6
    for q in range(4):
        a, b = g.get_plane(OUT, q), o.get_plane(OUT, q)
        assert np.max(np.abs(a - b)) <= 1e-8 * max(np.max(np.abs(b)), 1.0), q
    g.close()
    o.close()
