"""Closed-form pins of the oracle's output-stage diagnostics (CPU).

Each expected value is worked out from the reference lines cited and from
mathematics that does not repeat the code under test:
  * viscous dissipation (Extract.c:453-725) of constant velocities on a
    uniform box: the Laplacian vanishes inside and each wall adds
    (k_wall - 1)/h^2 times the velocity (the wall-normal component: -1/h^2),
    so the total is a count of wall cells; for fields that vanish near the
    walls, summation by parts turns sum u lap(u) dV into minus the sum of
    squared differences over the edges;
  * the energy conversion and potential energies (Extract.c:727-802,
    Conc.c:1934-2058) of constant fields: products with the box volume and
    the midpoint rule's exact first moment Ly^2/2;
  * span-averaged height and front (Conc.c:1524-1673) of a profile linear
    in x: the height is the profile, the front its root;
  * deposits (Conc.c:1675-1933) of constant fields: dt |V_s0| w (c + c_old)/2
    per step on the columns with an interface cell;
  * wall shear (Extract.c:18-452): a uniform stream over the flat box floor
    gives |U|/(h/2 Re); a linear shear S (y - h0) over an immersed flat
    floor gives S/Re in both variants (trilinear interpolation and the
    quadratic fit are exact on it); a uniform downward v shows the
    quadratic variant's |V_n| (Extract.c:298-303): v_tan = 2 v there and 0
    in the linear variant."""
import numpy as np
import pytest

import oracle
from turbins_amd import abi, workload as W


def _box(nx=8, ny=6, nz=4, lx=2.0, ly=1.5, lz=1.0, walls=(1, 1, 1, 1, 1, 1), **kw):
    args = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, re=250.0, pe=250.0, eno_order=2,
                wall_noslip=walls, buoyancy=0.0, conc_type=abi.SALINITY, v_settle=0.0, sedimentation=0)
    args.update(kw)
    cfg = abi.make_config(**args)
    o = oracle.Oracle(cfg)
    o.setup_geometry(None)
    return cfg, o


def _set(o, **fields):
    for n in "UVWC":
        o.upload(n, np.asarray(fields.get(n, np.zeros(o.shape)), dtype=float) * np.ones(o.shape))


@pytest.mark.parametrize("walls", [(1, 1, 1, 1, 1, 1), (0, 0, 0, 0, 0, 0), (1, 0, 1, 0, 0, 1), (0, 1, 0, 1, 1, 0)])
def test_viscous_dissipation_constant_velocity(walls):
    cfg, o = _box(walls=walls)
    NI, NJ, NK = cfg.nx, cfg.ny, cfg.nz
    hx, hy, hz = cfg.lx / NI, cfg.ly / NJ, cfg.lz / NK
    k = [-1.0 if w else 1.0 for w in walls]   # left, right, bottom, top, back, front
    U0, V0, W0 = 0.7, -0.4, 0.25
    # cells on each wall: x walls NJ*NK, y walls NI*NK, z walls NI*NJ
    sx = NJ * NK / hx**2
    sy = NI * NK / hy**2
    sz = NI * NJ / hz**2
    lap_u = -2 * sx + ((k[2] - 1) + (k[3] - 1)) * sy + ((k[4] - 1) + (k[5] - 1)) * sz
    lap_v = ((k[0] - 1) + (k[1] - 1)) * sx - 2 * sy + ((k[4] - 1) + (k[5] - 1)) * sz
    lap_w = ((k[0] - 1) + (k[1] - 1)) * sx + ((k[2] - 1) + (k[3] - 1)) * sy - 2 * sz
    dV = hx * hy * hz
    for U, V, Wv, want in [(U0, 0, 0, U0 * U0 * lap_u), (0, V0, 0, V0 * V0 * lap_v), (0, 0, W0, W0 * W0 * lap_w),
                           (U0, V0, W0, U0 * U0 * lap_u + V0 * V0 * lap_v + W0 * W0 * lap_w)]:
        _set(o, U=U, V=V, W=Wv)
        got = o.viscous_dissipation()
        want_total = want * dV / cfg.re
        assert abs(got - want_total) <= 1e-12 * abs(want_total), (walls, U, V, Wv, got, want_total)
    o.close()


def test_viscous_dissipation_summation_by_parts():
    """Fields that vanish within two cells of every wall: sum u lap(u) dV =
    -sum over interior edges of (u_a - u_b)^2 (cell volume / h^2), per
    component and direction (uniform grid)."""
    cfg, o = _box(nx=12, ny=10, nz=8, lx=3.0, ly=2.5, lz=2.0)
    hx, hy, hz = cfg.lx / cfg.nx, cfg.ly / cfg.ny, cfg.lz / cfg.nz
    rng = np.random.default_rng(4)
    fields = {}
    for n in "UVW":
        f = np.zeros(o.shape)
        f[2:-3, 2:-3, 2:-3] = rng.uniform(-1, 1, size=f[2:-3, 2:-3, 2:-3].shape)
        fields[n] = f
    _set(o, **fields)
    got = o.viscous_dissipation()
    dV = hx * hy * hz
    want = 0.0
    for f in fields.values():
        for axis, h in ((2, hx), (1, hy), (0, hz)):
            want -= np.sum(np.diff(f, axis=axis) ** 2) * dV / h**2
    want /= cfg.re
    assert abs(got - want) <= 1e-12 * abs(want)
    o.close()


def test_conversion_and_potential_energies_constant_fields():
    cfg, o = _box(conc_type=abi.PARTICLE, v_settle=-0.02)
    o.set_conc_fields([dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.02, alpha=0.8)])
    V0, C0 = -0.3, 0.6
    _set(o, V=V0, C=C0)
    vol = cfg.lx * cfg.ly * cfg.lz
    conv = o.kinetic_potential_conversion()
    # one field: c[0]->data, no alpha (Extract.c:737-741)
    assert abs(conv - (-V0 * C0 * vol)) <= 1e-13 * abs(V0 * C0 * vol)
    e = o.potential_energies()
    # sum over cells of dV (y_c - y_v[0]) is the exact moment Lx Lz Ly^2 / 2
    ep = 0.8 * C0 * cfg.lx * cfg.lz * cfg.ly**2 / 2
    assert abs(e["ep"] - ep) <= 1e-13 * ep
    assert abs(e["stokes"] - 0.02 * 0.8 * C0 * vol) <= 1e-13 * vol
    assert abs(e["conversion"] - 0.8 * V0 * C0 * vol) <= 1e-13 * vol
    assert e["ep_passive"] == 0.0 and e["ep_passive_rate"] == 0.0
    assert e["ep_rate"] == e["conversion"] + e["ep_passive_rate"] - e["stokes"]
    # two fields: the weighted total drives the conversion
    o.set_conc_fields([dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.02, alpha=0.8),
                       dict(type=abi.SALINITY, pe=cfg.pe, alpha=0.2)])
    o.select_conc(1)
    o.upload("C", np.full(o.shape, 1.5))
    o.select_conc(0)
    conv2 = o.kinetic_potential_conversion()
    want = -V0 * (0.8 * C0 + 0.2 * 1.5) * vol
    assert abs(conv2 - want) <= 1e-13 * abs(want)
    o.select_conc(1)
    assert o.potential_energies()["stokes"] == 0.0      # salinity: v_settl0 = 0
    o.close()


def test_passive_potential_energy_of_a_deposit():
    cfg, o = _box(conc_type=abi.PARTICLE, v_settle=-0.02)
    dep = np.zeros((cfg.nz + 1, cfg.nx + 1))
    dep[:-1, :-1] = 0.125               # columns with a FLUID row 0
    dep[-1, :] = 9.0                    # columns without an interface cell: ignored
    dep[:, -1] = 9.0
    o.deposit_set(dep)
    e = o.potential_energies()
    hy = cfg.ly / cfg.ny
    want = 0.125 * cfg.lx * cfg.lz * (hy / 2)
    assert abs(e["ep_passive"] - want) <= 1e-14
    assert abs(e["ep_passive_rate"] - want * 0.02) <= 1e-16
    o.close()


def test_ave_height_and_front_of_a_linear_profile():
    cfg, o = _box(nx=32, lx=8.0)
    xc = o.grid("xc")
    prof = 0.9 - 0.2 * xc                # crosses 0.01 at x = 4.45
    C = np.broadcast_to(prof[None, None, :], o.shape)
    _set(o, C=C)
    h = o.ave_height_x()
    np.testing.assert_allclose(h[:-1], prof[:-1], rtol=1e-14, atol=1e-15)
    assert h[-1] == 0.0                  # the last plane is BOUNDARY
    assert abs(o.front_location(0.01) - (0.9 - 0.01) / 0.2) <= 1e-12
    assert o.front_location(2.0) == -1.0
    # two fields: the alpha-weighted heights cross where 0.5 p + 0.5 (p/2) = 0.01
    o.set_conc_fields([dict(type=abi.SALINITY, pe=cfg.pe, alpha=0.5), dict(type=abi.SALINITY, pe=cfg.pe, alpha=0.5)])
    o.upload("C", C)
    o.select_conc(1)
    o.upload("C", C * 0.5)
    o.select_conc(0)
    assert abs(o.front_location(0.01) - (0.9 - 0.01 / 0.75) / 0.2) <= 1e-12
    o.close()


def test_deposit_integration_and_dump():
    cfg, o = _box(conc_type=abi.PARTICLE, v_settle=-0.02)
    _set(o, C=0.5)
    o.store_old()
    o.upload("C", np.full(o.shape, 0.3))
    dt, wgt = 1e-3, 0.75
    o.deposit_integrate(wgt, dt)
    o.deposit_integrate(wgt, dt)
    d = o.deposit_height()
    one = dt * 0.02 * wgt * 0.5 * (0.3 + 0.5)
    np.testing.assert_allclose(d[:-1, :-1], 2 * one, rtol=1e-15)
    assert np.all(d[-1, :] == 0.0) and np.all(d[:, -1] == 0.0)
    dumped = o.deposit_dumped(wgt)
    np.testing.assert_allclose(dumped[:-1, :-1], 2 * one + wgt * 0.3 * cfg.ly, rtol=1e-14)
    o.close()


def test_wall_shear_uniform_stream_over_the_box_floor():
    cfg, o = _box()
    _set(o, U=0.3, W=-0.4)
    hy = cfg.ly / cfg.ny
    for quad in (False, True):
        s = o.wall_shear(quad)
        np.testing.assert_allclose(s["tau"][:-1, :-1], 0.5 / (hy / 2 * cfg.re), rtol=1e-14)
        np.testing.assert_allclose(s["u_tan"][:-1, :-1], 0.3, rtol=1e-15)
        np.testing.assert_allclose(s["w_tan"][:-1, :-1], -0.4, rtol=1e-15)
        assert np.all(s["v_tan"] == 0.0)
        assert np.all(s["tau"][-1, :] == 0.0) and np.all(s["tau"][:, -1] == 0.0)
    o.close()


def test_wall_shear_over_an_immersed_flat_floor():
    c = W.case("bump_small", nx=32, ny=32, nz=8)
    cfg = W.config_for(c)
    o = oracle.Oracle(cfg)
    o.setup_geometry(np.full((c["nz"] + 1, c["nx"] + 1), 0.23))
    ib = o.ib_table(abi.Q_C)
    assert len(ib) > 0 and all(abs(n.n[1] - 1.0) < 1e-12 for n in ib)
    # the floor the level set sees: 0.23 lifted by 2.001 dy_c (Grid.c:913)
    h0 = 0.23 + 2.001 * cfg.ly / cfg.ny
    assert all(abs(n.control[1] - h0) < 1e-12 for n in ib)
    S = 0.8
    yu = o.grid("yu")
    U = np.broadcast_to(S * (yu - h0)[None, :, None], o.shape)
    _set(o, U=U)
    cols = np.array([[o.interface_y_index(i, k) for i in range(cfg.nx + 1)] for k in range(cfg.nz + 1)])
    ok = cols >= 0
    assert ok.sum() > 100
    for quad in (False, True):
        s = o.wall_shear(quad)
        np.testing.assert_allclose(s["tau"][ok], S / cfg.re, rtol=1e-9)
    # a uniform downward v: the linear variant removes the signed normal
    # component, the quadratic one its magnitude
    _set(o, V=-0.05)
    lin, quad = o.wall_shear(False), o.wall_shear(True)
    np.testing.assert_allclose(lin["v_tan"][ok], 0.0, atol=1e-15)
    np.testing.assert_allclose(quad["v_tan"][ok], -0.1, rtol=1e-12)
    o.close()
