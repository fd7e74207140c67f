"""Known-answer pins of the oracle's open x faces (Inflow.c, Outflow.c and
the inflow / outflow branches of the operators).

The boxes have unit cell sizes or binary-fraction ones, so positions,
metrics and interpolation weights are exact and most expected values below
are exact; each is worked out from the reference lines cited beside it.  A
plausible slip -- a ghost taken from the wall rule instead of the plane, a
plane read on the wrong side or index, a wall rule kept at an open face, a
dropped right-hand-side term, a mass correction with the wrong sign or
area -- changes at least one of them."""
import math

import numpy as np
import pytest

import oracle
from turbins_amd import abi

U, V, W, Cq = abi.Q_U, abi.Q_V, abi.Q_W, abi.Q_C
IN, OUT = abi.PLANE_INFLOW, abi.PLANE_OUTFLOW


def _cfg(nx, ny, nz, lx, ly, lz, **kw):
    args = dict(nx=nx, ny=ny, nz=nz, lx=lx, ly=ly, lz=lz, re=1000.0, pe=500.0, eno_order=1,
                wall_noslip=(1, 1, 1, 1, 1, 1), buoyancy=0.0, conc_type=abi.SALINITY, v_settle=0.0,
                sedimentation=0)
    args.update(kw)
    return abi.make_config(**args)


def _box(nx=8, ny=4, nz=4, **kw):
    """flat box with unit cells, both x faces open"""
    o = oracle.Oracle(_cfg(nx, ny, nz, float(nx), float(ny), float(nz), **kw))
    o.setup_geometry(None)
    o.set_open_boundaries(inflow=True, outflow=True)
    return o


def _plane(o, val):
    return np.full((o.cfg.nz + 1) * (o.cfg.ny + 1), float(val))


def _at(arr, o, i, j, k):
    return arr.reshape(o.shape)[k, j, i]


def _upload_all(o, fields):
    for name in ["U", "V", "W", "C", "CONV_U", "CONV_V", "CONV_W", "CONV_C", "DPDX", "DPDY", "DPDZ"]:
        o.upload(name, fields.get(name, np.zeros(o.NT)))


def _masked(o, q, val):
    """val on q's non-BOUNDARY nodes, 0 elsewhere"""
    return np.where(o.flags(q) != abi.BOUNDARY, float(val), 0.0)


def test_inflow_profile_closed_form():
    """Inflow_u_velocity_profile (Inflow.c:85-139): u_in = 1 times the
    transient factor 1 - exp(-(t + 0.01)/time_c) on the FLUID u nodes of
    i = 1 (every j < ny, k < nz in a flat box), influx = sum u_in dy dz =
    factor * Ly * Lz; Inflow_conc_profile (Inflow.c:219-252): c_in = factor on
    every non-SOLID cell of i = 0 (all of them in a flat box)."""
    nx, ny, nz = 8, 4, 2
    o = oracle.Oracle(_cfg(nx, ny, nz, 2.0, 1.0, 0.5))
    o.setup_geometry(None)
    o.set_open_boundaries(inflow=True, transient_inflow=True, time_c=0.5)
    t = 0.24
    fl = o.inflow_profile(t)
    tf = 1.0 - math.exp(-(t + 0.01) / 0.5)
    uin = o.get_plane(IN, U).reshape(nz + 1, ny + 1)
    assert np.all(uin[:nz, :ny] == tf) and np.all(uin[nz, :] == 0.0) and np.all(uin[:, ny] == 0.0)
    assert abs(fl - tf * 1.0 * 0.5) <= 1e-15 and o.influx() == fl
    assert np.all(o.get_plane(IN, Cq) == tf)
    # steady inflow: factor 1; setting the plane recomputes the influx
    o.set_open_boundaries(inflow=True)
    assert o.inflow_profile(3.0) == 0.5
    p = np.zeros((nz + 1) * (ny + 1))
    p[0] = 2.0          # (j, k) = (0, 0): dy dz = 0.25 * 0.25
    p[ny] = 7.0         # (j, k) = (ny, 0): a BOUNDARY u node, not counted
    o.set_plane(IN, U, p)
    assert o.influx() == 2.0 * 0.0625
    o.close()


def test_inflow_profile_skips_solid_and_immersed_rows():
    """Over the bump the u nodes of i = 1 under the topography are not FLUID
    and keep 0; the influx sums dy dz over the FLUID ones only."""
    from turbins_amd import workload as W
    c = W.case("bump_small", nx=32, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    h = np.full((c["nz"] + 1, c["nx"] + 1), 0.23)
    o.setup_geometry(h)
    o.set_open_boundaries(inflow=True)
    fl = o.inflow_profile(0.0)
    st = o.flags(U).reshape(o.shape)[:, :, 1]          # [k, j]
    uin = o.get_plane(IN, U).reshape(c["nz"] + 1, c["ny"] + 1)
    fluid = st == abi.FLUID
    assert 0 < fluid.sum() < fluid.size
    assert np.all(uin[fluid] == 1.0) and np.all(uin[~fluid] == 0.0)
    dy = np.diff(o.grid("yv"))
    dz = np.diff(o.grid("zw"))
    want = sum(dy[j] * dz[k] for k in range(c["nz"] + 1) for j in range(c["ny"] + 1) if fluid[k, j])
    assert abs(fl - want) <= 1e-14 * want
    o.close()


def test_outflow_convective_update_and_mass_correction():
    """Outflow_vel_impose_convective_boundary (Outflow.c:44-87): on rows
    whose node at i = NI-1 is FLUID, out = (1 - c) out + c q[NI-1] with
    c = dt U / dx; v, w and every concentration field exactly that; u then
    shifts by (influx - outflux)/area on those rows (Outflow.c:186-289), so
    its outflux equals the influx."""
    o = _box(nx=8, ny=4, nz=4)
    o.set_conc_fields([dict(type=abi.SALINITY, pe=500.0), dict(type=abi.SALINITY, pe=250.0)])
    o.set_open_boundaries(inflow=True, outflow=True, u_conv=0.75)
    o.inflow_profile(0.0)
    influx = o.influx()
    assert influx == 16.0
    rng = np.random.default_rng(4)
    NP = 5 * 5
    p0 = {q: rng.uniform(-1, 1, NP) for q in range(5)}
    fld = {q: rng.uniform(-1, 1, o.NT) for q in range(5)}
    for q, n in enumerate("UVW"):
        o.upload(n, fld[q])
        o.set_plane(OUT, q, p0[q])
    for s in (0, 1):
        o.select_conc(s)
        o.upload("C", fld[3 + s])
        o.set_plane(OUT, Cq, p0[3 + s])
    o.select_conc(0)
    dt = 0.25
    coef = dt * 0.75 / 1.0
    o.outflow_convective(dt)
    last = 7          # NI - 1 = NX - 2 with NX = nx + 1 = 9
    for q in range(5):
        qq = min(q, 3)
        st = o.flags(qq).reshape(o.shape)[:, :, last].reshape(-1)
        open_rows = (st == abi.FLUID) | (st == abi.IMMERSED)
        assert open_rows.sum() == (12 if q in (1, 2) else 16)   # v: j = 0, w: k = 0 are BOUNDARY
        data = fld[q].reshape(o.shape)[:, :, last].reshape(-1)
        want = np.where(open_rows, (1.0 - coef) * p0[q] + coef * data, p0[q])
        if q >= 3:
            o.select_conc(q - 3)
        got = o.get_plane(OUT, qq)
        if q == 0:
            d = got - want
            corr = d[open_rows]
            assert np.all(d[~open_rows] == 0.0)
            assert np.ptp(corr) <= 1e-15
            outflux = want[open_rows].sum()
            assert abs(corr[0] - (influx - outflux) / 16.0) <= 1e-14
            assert abs(got[open_rows].sum() - influx) <= 1e-13
        else:
            assert np.array_equal(got, want), q
    o.close()


def test_open_face_eno_ghosts_transposes_and_cell_centres():
    """First-order ENO at the open faces (upwind differences, unit spacing):

    u along x (Velocity.c:2193-2233): u = 2 on every interior u node and on
    both planes gives no x term anywhere; the closed box's wall ghost 0
    gives 2 * (2 - 0) = 4 at i = 1.

    v along x, advected by u = -1 (planes too): the transpose to the v nodes
    of i = NX-2 reads u_out (Velocity.c:3502-3506), so u_v = -1 there (the
    BOUNDARY u node would give -1/2); the right ghost is v_out
    (Velocity.c:2442-2444) at x = 8.5 for the node at 7.5, so with v = 3
    inside and v_out = 5: -1 * (5 - 3) / 1 = -2.  w the same (3599-3603,
    2669-2671).

    c along x: the cell-centre u of cell 0 interpolates u_in and u[1]
    (Velocity.c:1544-1558); with u = 1 and c = 1 inside, c_in = 3 (ghost at
    x = -0.5, Conc.c:1099-1101): 1 * (1 - 3) / 1 = -2 at i = 0; with u = -1
    and c_out = 4 (Conc.c:1120-1122): -1 * (4 - 1) / 1 = -3 at i = NX-2."""
    o = _box()
    NX = 9
    # u: no x term with matching planes
    for side in (IN, OUT):
        o.set_plane(side, U, _plane(o, 2.0))
    _upload_all(o, {"U": _masked(o, U, 2.0)})
    o.eno_conv(U)
    cu = o.download("CONV_U")
    assert np.all(cu == 0.0)
    # v and w: transposes and right ghosts
    for q, n in ((V, "V"), (W, "W")):
        for side in (IN, OUT):
            o.set_plane(side, U, _plane(o, -1.0))
        o.set_plane(OUT, q, _plane(o, 5.0))
        _upload_all(o, {"U": _masked(o, U, -1.0), n: _masked(o, q, 3.0)})
        o.eno_conv(q)
        cv = o.download("CONV_" + n)
        # a node away from the y and z walls
        assert _at(cv, o, NX - 2, 2, 2) == -2.0, n
        assert _at(cv, o, NX - 3, 2, 2) == 0.0
    # c: left and right
    o.set_plane(IN, U, _plane(o, 1.0))
    o.set_plane(OUT, U, _plane(o, 1.0))
    o.set_plane(IN, Cq, _plane(o, 3.0))
    _upload_all(o, {"U": _masked(o, U, 1.0), "C": _masked(o, Cq, 1.0)})
    o.eno_conv(Cq)
    cc = o.download("CONV_C")
    assert _at(cc, o, 0, 2, 2) == -2.0 and _at(cc, o, 1, 2, 2) == 0.0
    o.set_plane(IN, U, _plane(o, -1.0))
    o.set_plane(OUT, U, _plane(o, -1.0))
    o.set_plane(OUT, Cq, _plane(o, 4.0))
    _upload_all(o, {"U": _masked(o, U, -1.0), "C": _masked(o, Cq, 1.0)})
    o.eno_conv(Cq)
    cc = o.download("CONV_C")
    assert _at(cc, o, NX - 2, 2, 2) == -3.0 and _at(cc, o, NX - 3, 2, 2) == 0.0
    # the closed box keeps its wall ghosts
    o.set_open_boundaries()
    _upload_all(o, {"U": _masked(o, U, 2.0)})
    o.eno_conv(U)
    assert _at(o.download("CONV_U"), o, 1, 2, 2) == 4.0
    o.close()


def test_open_face_helmholtz_rows():
    """A constant field K = 1 through the assembled rows: an open face drops
    its wall rule (Velocity.c:989-998, 1336-1343; Conc.c:582-607) and the
    column, so the row sum there is 1/dt + kappa^-1 / dx^2 (the missing
    neighbour); the closed no-slip wall gives 1/dt + 2/(Re dx^2) for v, w and
    the Neumann wall 1/dt for c; u's rows do not change."""
    dt = 0.5
    closed = oracle.Oracle(_cfg(8, 4, 4, 8.0, 4.0, 4.0))
    closed.setup_geometry(None)
    o = _box()
    one = np.ones(o.NT)
    for q, kappa in ((U, 1000.0), (V, 1000.0), (W, 1000.0), (Cq, 500.0)):
        yc = closed.helmholtz_apply(q, dt, one).reshape(o.shape)
        yo = o.helmholtz_apply(q, dt, one).reshape(o.shape)
        if q == U:
            assert np.array_equal(yc, yo)
            continue
        e = yo[2, 2, 7]          # i = NX-2, away from the y and z walls
        assert abs(e - (1.0 / dt + 1.0 / kappa)) <= 1e-13
        ec = yc[2, 2, 7]
        assert abs(ec - (1.0 / dt + (2.0 / kappa if q != Cq else 0.0))) <= 1e-13
        if q == Cq:
            assert abs(yo[2, 2, 0] - (1.0 / dt + 1.0 / kappa)) <= 1e-13
            assert abs(yc[2, 2, 0] - 1.0 / dt) <= 1e-13
        # the interior rows agree
        assert np.array_equal(yc[1:3, 1:3, 2:6], yo[1:3, 1:3, 2:6])
    closed.close()
    o.close()


def test_open_face_rhs_terms():
    """With data, convective term and pressure gradient zero, the rows next
    to an open face carry the plane value times the dropped coefficient:
    u at i = 1: aw u_in with aw = -Cx metric_xu[1] metric_xc[0]
    (Velocity.c:1691-1705), u at i = NX-2: ae u_out (1707-1721), v and w at
    i = NX-2: ae out (1885-1897, 2011-2023), c at i = 0 and NX-2 (Conc.c:
    938-963, whose outflow term uses metric_xu[i]); on unit cells each
    coefficient is 1/(kappa dx^2) = 1/kappa.  Every other row stays 0."""
    o = _box()
    rng = np.random.default_rng(9)
    NP = 25
    for q, kappa, planes in ((U, 1000.0, (IN, OUT)), (V, 1000.0, (OUT,)), (W, 1000.0, (OUT,)),
                             (Cq, 500.0, (IN, OUT))):
        vals = {side: rng.uniform(1, 2, NP) for side in planes}
        for side in (IN, OUT):
            if q in (U, Cq) or side == OUT:
                o.set_plane(side, q, vals.get(side, np.zeros(NP)))
        _upload_all(o, {})
        o.set_rhs(q, 0.5)
        r = o.download("RHS_" + "UVWC"[q]).reshape(o.shape)
        fl = o.flags(q).reshape(o.shape)
        row = (fl == abi.FLUID) | ((fl == abi.BUFFER_SOLID) & (q < 3))
        want = np.zeros(o.shape)
        i_in = 1 if q == U else 0
        for side, i in ((IN, i_in), (OUT, 7)):
            if side in vals:
                pv = vals[side].reshape(5, 5)
                want[:, :, i] += np.where(row[:, :, i], pv / kappa, 0.0)
        np.testing.assert_allclose(r, want, rtol=1e-15, atol=0.0)
        assert np.count_nonzero(r) > 0
    o.close()


def test_open_face_divergence_and_poisson_rhs():
    """Pressure_set_RHS / Pressure_compute_velocity_divergence (Pressure.c:
    242-249, 541-548): at cell 0 du/dx = (u[1] - u_in)/dx, at cell NX-2
    (u_out - u[NX-2])/dx.  u = v = w = 0 inside, u_in = 1, u_out = 2 on unit
    cells: rhs = -(du/dx)/dt = 1/dt and -2/dt, |div| 1 and 2."""
    o = _box()
    o.set_plane(IN, U, _plane(o, 1.0))
    o.set_plane(OUT, U, _plane(o, 2.0))
    _upload_all(o, {})
    dt = 0.25
    o.poisson_rhs(dt)
    r = o.download("PRHS").reshape(o.shape)
    assert np.all(r[:4, :4, 0] == 4.0) and np.all(r[:4, :4, 7] == -8.0)
    assert np.all(r[:4, :4, 1:7] == 0.0)
    dm, ijk = o.div_max()
    assert dm == 2.0 and ijk[0] == 7
    o.close()


def test_open_channel_steps_conserve_mass():
    """Steps of an open channel (inflow at x = 0, convective outflow at
    x = Lx, flat bottom, no-slip side walls) from the plug flow u = 1 that
    matches the inflow: after every stage the outflow plane carries the
    influx (the mass correction), so the Neumann Poisson problem stays
    compatible and every step ends divergence-free to the solver tolerance
    (a missing or wrong correction leaves the projected field with a
    divergence of the size of the flux defect); the RK average of two
    corrected planes carries the influx too."""
    from turbins_amd import workload as W
    c = W.case("lock64", nx=16, ny=8, nz=8, lx=2.0)
    cfg = W.config_for(c, pois_rtol=1e-12, helm_rtol=1e-13)
    o = oracle.Oracle(cfg)
    o.setup_geometry(None)
    o.set_open_boundaries(inflow=True, outflow=True, u_conv=1.0)
    fl = o.inflow_profile(0.0)
    assert fl == 1.0
    o.upload("U", _masked(o, U, 1.0))
    o.set_plane(OUT, U, o.get_plane(IN, U))
    assert o.div_max()[0] == 0.0
    dy = np.diff(o.grid("yv"))
    dz = np.diff(o.grid("zw"))
    for s in range(4):
        st = o.step(5e-3)
        assert st.div_max <= 1e-9, (s, st.div_max)
        uo = o.get_plane(OUT, U).reshape(o.shape[0], o.shape[1])
        flux = sum(uo[k, j] * dy[j] * dz[k] for k in range(cfg.nz) for j in range(cfg.ny))
        assert abs(flux - fl) <= 1e-12, (s, flux)
    u = o.download("U").reshape(o.shape)
    assert 0.9 < u[4, 4, 8] < 1.2 and u[0, 0, 8] < u[4, 4, 8]   # core and a wall-slowed corner
    o.close()


def test_open_face_viscous_dissipation_linear_in_the_planes():
    """Extract_viscous_dissipation_rate with open faces (Extract.c:577-590,
    630-645): the planes replace the wall ghosts of cells 0 and NX-2.  With
    the outflow planes equal to the no-slip ghosts (0, -v, -w of the last
    cells) the rate equals the closed box's; it is affine in the outflow
    planes and moves with them.  The inflow plane enters u's Laplacian at
    cell 0 only, where it is weighted by u[0], the BOUNDARY node, which the
    solves keep at 0: it leaves the rate unchanged (the reference's own note
    at Extract.c:452 that in/outflow cases are not handled properly)."""
    rng = np.random.default_rng(2)
    o = _box()
    closed = oracle.Oracle(_cfg(8, 4, 4, 8.0, 4.0, 4.0))
    closed.setup_geometry(None)
    f = {n: np.where(o.flags(q) != abi.BOUNDARY, rng.uniform(-1, 1, o.NT), 0.0) for q, n in enumerate("UVW")}
    for x in (o, closed):
        _upload_all(x, f)
    last = lambda n: f[n].reshape(o.shape)[:, :, 7].reshape(-1)
    o.set_plane(OUT, V, -last("V"))
    o.set_plane(OUT, W, -last("W"))
    ref = closed.viscous_dissipation()
    assert abs(o.viscous_dissipation() - ref) <= 1e-13 * abs(ref)
    o.set_plane(IN, U, rng.uniform(-1, 1, 25))
    assert abs(o.viscous_dissipation() - ref) <= 1e-13 * abs(ref)
    for q in (U, V, W):
        base = o.viscous_dissipation()
        p = o.get_plane(OUT, q)
        dp = rng.uniform(-1, 1, 25)
        o.set_plane(OUT, q, p + dp)
        d1 = o.viscous_dissipation() - base
        o.set_plane(OUT, q, p + 2 * dp)
        d2 = o.viscous_dissipation() - base
        assert d1 != 0.0 and abs(d2 - 2 * d1) <= 1e-11 * abs(d1), q
        o.set_plane(OUT, q, p)
    closed.close()
    o.close()
