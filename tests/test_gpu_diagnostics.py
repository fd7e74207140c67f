"""Output-stage diagnostics on the GPU, through the C-ABI, against the
oracle on the same inputs: per-column quantities (wall shear of both
variants, deposits, the passive potential energy) bit for bit; sums over
cells (viscous dissipation, energy conversion, potential energies, the
span-averaged height) to 1e-12 relative, their summation order being the
only difference; the front location from those heights."""
import numpy as np
import pytest

from turbins_amd import abi, workload as W

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)


def _pair(name, **over):
    from turbins_amd import Solver
    c = W.case(name)
    cfg = W.config_for(c, **over)
    g, o = Solver(cfg), oracle.Oracle(cfg)
    h = W.topography(c)
    g.setup_geometry(h)
    o.setup_geometry(h)
    return c, cfg, g, o


def _state(c, o, seed):
    """the lock-exchange start plus a smooth random velocity"""
    init = W.initial_fields(c, [o.flags(q) for q in range(4)])
    NT = o.NT
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(c["nz"] + 1), np.arange(c["ny"] + 1), np.arange(c["nx"] + 1), indexing="ij")
    for a, n in enumerate("UVW"):
        init[n] = (0.05 * np.sin(0.2 * i + 0.3 * a) * np.cos(0.15 * j) * np.cos(0.4 * k)).reshape(-1) + \
            0.01 * rng.uniform(-1, 1, NT)
    return init


def _close(a, b, rtol, what):
    assert abs(a - b) <= rtol * max(abs(b), 1e-300), (what, a, b)


@pytest.mark.parametrize("name,walls", [("bump_small", (1, 1, 1, 1, 1, 1)), ("turbidity_small", (1, 1, 1, 0, 0, 1))])
def test_cell_sums_match_oracle(name, walls):
    c, cfg, g, o = _pair(name, wall_noslip=walls)
    print("cell sums", name, walls, flush=True)
    st = _state(c, o, 3)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    _close(g.viscous_dissipation(), o.viscous_dissipation(), 1e-12, "viscous dissipation")
    _close(g.kinetic_potential_conversion(), o.kinetic_potential_conversion(), 1e-12, "conversion")
    eg, eo = g.potential_energies(), o.potential_energies()
    for n in eo:
        _close(eg[n], eo[n], 1e-12, n)
    hg, ho = g.ave_height_x(), o.ave_height_x()
    assert np.max(np.abs(hg - ho)) <= 1e-13 * np.max(np.abs(ho))
    fg, fo = g.front_location(0.01), o.front_location(0.01)
    print("front gpu/oracle", fg, fo)
    assert fo > 0 and abs(fg - fo) <= 1e-10
    g.close()
    o.close()


@pytest.mark.parametrize("name", ["bump_small", "turbidity_small"])
def test_wall_shear_bit_identical(name):
    c, cfg, g, o = _pair(name)
    print("wall shear", name, flush=True)
    st = _state(c, o, 7)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    for quad in (False, True):
        sg, so = g.wall_shear(quad), o.wall_shear(quad)
        for n in so:
            assert np.array_equal(sg[n], so[n]), (quad, n, np.max(np.abs(sg[n] - so[n])))
        assert np.count_nonzero(so["tau"]) > 100
    g.close()
    o.close()


def test_deposits_and_passive_energy_bit_identical():
    c, cfg, g, o = _pair("turbidity_small")
    st = _state(c, o, 11)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    g.store_old()
    o.store_old()
    c1 = np.clip(st["C"] * 0.9 + 0.05, 0.0, 1.0)
    g.upload("C", c1)
    o.upload("C", c1)
    for s in range(3):
        g.deposit_integrate(0.7, 1e-3 * (s + 1))
        o.deposit_integrate(0.7, 1e-3 * (s + 1))
    dg, do = g.deposit_height(), o.deposit_height()
    assert np.array_equal(dg, do) and np.count_nonzero(do) > 100
    assert np.array_equal(g.deposit_dumped(0.7), o.deposit_dumped(0.7))
    eg, eo = g.potential_energies(), o.potential_energies()
    assert eg["ep_passive"] == eo["ep_passive"] and eo["ep_passive"] > 0
    assert eg["ep_passive_rate"] == eo["ep_passive_rate"]
    # set and read back; a second field has its own deposit
    g.deposit_set(dg * 2)
    assert np.array_equal(g.deposit_height(), dg * 2)
    g.set_conc_fields([dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=cfg.v_settle),
                       dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.01)])
    assert np.array_equal(g.deposit_height(), dg * 2)
    g.select_conc(1)
    assert not np.any(g.deposit_height())
    g.close()
    o.close()


def test_diagnostics_of_several_fields():
    c, cfg, g, o = _pair("turbidity_small")
    fields = [dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.02, alpha=0.6),
              dict(type=abi.SALINITY, pe=cfg.pe, alpha=0.4)]
    g.set_conc_fields(fields)
    o.set_conc_fields(fields)
    st = _state(c, o, 5)
    for k, v in st.items():
        g.upload(k, v)
        o.upload(k, v)
    for x in (g, o):
        x.select_conc(1)
        x.upload("C", st["C"] * 0.5)
        x.select_conc(0)
    _close(g.kinetic_potential_conversion(), o.kinetic_potential_conversion(), 1e-12, "conversion, two fields")
    fg, fo = g.front_location(0.01), o.front_location(0.01)
    assert fo > 0 and abs(fg - fo) <= 1e-10
    for n in (0, 1):
        g.select_conc(n)
        o.select_conc(n)
        eg, eo = g.potential_energies(), o.potential_energies()
        for key in eo:
            _close(eg[key], eo[key], 1e-12, (n, key))
    g.close()
    o.close()
