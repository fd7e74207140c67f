"""Several concentration fields and hindered settling in the oracle (CPU).

Pins:
  * Conc_settling_speed_function (Conc.c:815-836) against the exact rational
    value of V_s0 (1 - c phi)^4 and against pow();
  * the hindered particle velocity of Conc_compute_particle_settling_speed /
    Conc_set_particle_v_velocity (Conc.c:734-812) through the convective
    term of a linear profile, which ENO differentiates exactly: conv =
    v_p dc/dy with v_p the hindered speed on FLUID cells and 0 on the others;
  * the total concentration of Conc_compute_total_concentration
    (Conc.c:1291-1314) through the buoyancy of the v right-hand side
    (Velocity.c:1798-1815) on constant fields: -B sum alpha_i k_i with
    several fields, -B k_0 (no alpha) with one;
  * the field bookkeeping of the step: two identical fields of weight 1/2
    are one field (bit for bit), and a field that the velocity does not see
    (buoyancy 0) evolves as a one-field run with its own Pe and settling
    speed, immersed-boundary rows included;
  * the multigrain parameter rescaling of Input.c:453-481."""
from fractions import Fraction

import numpy as np
import pytest

import oracle
from turbins_amd import abi, workload as W


def _ulps(a, b):
    return abs(a - b) / np.spacing(abs(b)) if b != 0 else abs(a)


def test_hindered_speed_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(2000):
        c, phi, vs = rng.uniform(0, 1.2), rng.uniform(0, 0.7), -rng.uniform(1e-4, 0.1)
        got = oracle.hindered_speed(c, phi, vs)
        cphi = Fraction(c) * Fraction(phi)
        t = 1 - cphi
        exact = float(Fraction(vs) * t ** 4)
        # roundings: c phi (relative u, amplified by c phi / t in 1 - c phi),
        # 1 - c phi, two squarings and the product; the fourth power scales
        # the relative error of t by 4
        bound = 8 + 4 * float(cphi / t)
        assert _ulps(got, exact) <= bound, (c, phi, vs)
        assert _ulps(got, vs * pow(1.0 - c * phi, 4.0)) <= 4
    # special cases: no particles, a packed bed (c phi = 1), no volume fraction
    assert oracle.hindered_speed(0.0, 0.4, -0.02) == -0.02
    assert oracle.hindered_speed(2.0, 0.5, -0.02) == 0.0
    assert oracle.hindered_speed(0.7, 0.0, -0.02) == -0.02
    # (1 - 0.5)^4 = 1/16 exactly
    assert oracle.hindered_speed(1.0, 0.5, -0.02) == -0.02 / 16


def _particle_box(hindered, phi=0.5, geometry=True):
    c = W.case("turbidity_small", nx=32, ny=32, nz=8)
    cfg = W.config_for(c, eno_order=2)
    o = oracle.Oracle(cfg)
    o.setup_geometry(W.topography(c) if geometry else None)
    o.set_conc_fields([dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=cfg.v_settle, phi_max=phi)], hindered=hindered)
    return c, cfg, o


@pytest.mark.parametrize("hindered", [False, True])
def test_hindered_particle_velocity_through_a_linear_profile(hindered):
    """u = v = w = 0 and c = a + b y: ENO of a linear profile is exact away
    from the walls, so CONV_C = v_p b, with v_p = V_s0 everywhere (plain
    settling) or the hindered speed on FLUID cells and 0 elsewhere."""
    c, cfg, o = _particle_box(hindered)
    a, b = 0.25, 0.5
    yc = o.grid("yc")
    C = np.broadcast_to((a + b * yc)[None, :, None], o.shape).copy()
    for n in "UVW":
        o.upload(n, np.zeros(o.shape))
    o.upload("C", C)
    o.eno_conv(abi.Q_C)
    conv = o.download("CONV_C").reshape(o.shape)
    flags = o.flags(abi.Q_C).reshape(o.shape)
    vs, phi = cfg.v_settle, 0.5
    NY = o.shape[1]
    checked = {"fluid": 0, "other": 0}
    for k in range(o.shape[0] - 1):
        for j in range(3, NY - 5):
            for i in range(o.shape[2] - 1):
                s = flags[k, j, i]
                if s == abi.BOUNDARY:
                    continue
                if not hindered:
                    vp = vs
                elif s == abi.FLUID:
                    vp = vs * (1.0 - C[k, j, i] * phi) ** 4
                else:
                    vp = 0.0
                assert abs(conv[k, j, i] - vp * b) <= 1e-12 * abs(vs * b), (i, j, k, s)
                checked["fluid" if s == abi.FLUID else "other"] += 1
    assert checked["fluid"] > 1000 and checked["other"] > 20
    o.close()


@pytest.mark.parametrize("nconc", [1, 3])
def test_total_concentration_buoyancy_constant_fields(nconc):
    c = W.case("lock64", nx=16, ny=8, nz=8)
    cfg = W.config_for(c, buoyancy=1.25)
    o = oracle.Oracle(cfg)
    o.setup_geometry(None)
    alpha = [0.3, 0.5, 0.2][:nconc]
    k = [0.75, 2.0, -1.5][:nconc]
    o.set_conc_fields([dict(type=abi.SALINITY, pe=cfg.pe, alpha=a) for a in alpha])
    for n in "UVW":
        o.upload(n, np.zeros(o.shape))
    for n in range(nconc):
        o.select_conc(n)
        o.upload("C", np.full(o.shape, k[n]))
    o.select_conc(0)
    o.total_concentration()
    o.eno_conv(abi.Q_V)
    o.set_rhs(abi.Q_V, 1e-3)
    rhs = o.download("RHS_V").reshape(o.shape)
    fv = o.flags(abi.Q_V).reshape(o.shape)
    if nconc == 1:
        want = -1.25 * 0.75              # c[0]->data, no alpha (Velocity.c:1810-1815)
    else:
        want = -1.25 * (0.3 * 0.75 + 0.5 * 2.0 + 0.2 * -1.5)
    fl = rhs[fv == abi.FLUID]
    assert fl.size > 100
    assert np.all(np.abs(fl - want) <= 2 * np.spacing(abs(want))), (fl.min(), fl.max(), want)
    o.close()


def _lock_state(c, o):
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for kk, v in init.items():
        o.upload(kk, v)
    return init


def test_two_identical_fields_are_one_field():
    c = W.case("lock64", nx=32, ny=16, nz=8)
    cfg = W.config_for(c, helm_rtol=1e-13, pois_rtol=1e-11)
    one, two = oracle.Oracle(cfg), oracle.Oracle(cfg)
    for o in (one, two):
        o.setup_geometry(None)
    init = _lock_state(c, one)
    f = dict(type=abi.SALINITY, pe=cfg.pe, alpha=0.5)
    two.set_conc_fields([f, f])
    for kk, v in init.items():
        two.upload(kk, v)
    two.select_conc(1)
    two.upload("C", init["C"])
    two.select_conc(0)
    for _ in range(2):
        one.step(c["dt"])
        two.step(c["dt"])
    for n in ["U", "V", "W", "C", "DPDX", "PHI"]:
        assert np.array_equal(one.download(n), two.download(n)), n
    two.select_conc(1)
    assert np.array_equal(one.download("C"), two.download("C"))
    one.close()
    two.close()


def test_decoupled_field_matches_its_own_one_field_run():
    """Buoyancy 0: the velocity does not see the concentration, so field 1
    (particle, Pe 500, V_s0 -0.03, hindered) of a two-field run equals the
    one field of a run configured with those parameters; over a bump, so
    the concentration grid's immersed rows (scaled by Pe's ap_solid) are
    rebuilt for it."""
    c = W.case("turbidity_small", nx=32, ny=16, nz=8)
    cfg = W.config_for(c, buoyancy=0.0, helm_rtol=1e-13, pois_rtol=1e-11)
    f0 = dict(type=abi.SALINITY, pe=cfg.pe)
    f1 = dict(type=abi.PARTICLE, pe=500.0, v_settle=-0.03, phi_max=0.4)
    h = W.topography(c)
    two = oracle.Oracle(cfg)
    two.setup_geometry(h)
    assert two.L.ora_ib_count(two.h, abi.Q_C) > 0
    two.set_conc_fields([f0, f1], hindered=True)
    init = _lock_state(c, two)
    rng = np.random.default_rng(5)
    c1 = init["C"] * 0.5 + 0.1 * rng.uniform(size=init["C"].size)
    two.select_conc(1)
    two.upload("C", c1)
    two.select_conc(0)
    one = oracle.Oracle(cfg)
    one.setup_geometry(h)
    one.set_conc_fields([f1], hindered=True)
    for kk, v in init.items():
        one.upload(kk, v)
    one.upload("C", c1)
    for _ in range(2):
        two.step(c["dt"])
        one.step(c["dt"])
    two.select_conc(1)
    a, b = two.download("C"), one.download("C")
    assert np.array_equal(a, b), np.max(np.abs(a - b))
    assert not np.array_equal(a, c1)
    two.select_conc(0)
    assert not np.array_equal(two.download("C"), a)
    one.close()
    two.close()


def test_multigrain_properties():
    alpha, phi, vs = abi.multigrain_properties([2.0, 1.0, 1.0], 0.02, [0.01, 0.005, 0.0025])
    assert alpha == [0.5, 0.25, 0.25] and sum(alpha) == 1.0
    assert phi == [0.02, 0.01, 0.01]
    assert vs == [-0.01, -0.005, -0.0025]
    # one grain size: the weight becomes 1 whatever it was
    a1, p1, v1 = abi.multigrain_properties([0.37], 0.05, [0.02])
    assert a1 == [1.0] and p1 == [0.05] and v1 == [-0.02]
