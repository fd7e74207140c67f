"""Pins of the CPU oracle against mathematics and library routines.

The reference ships no tests, golden vectors or fixtures and cannot be built
here (PETSc 3.1 / hypre / MPI are absent, the driver was never published), so
every check below is chosen so that a plausible restatement mistake (a
dropped term, a wrong sign or index, a transposed operand) fails:
closed forms, textbook/library routines (numpy, scipy), invariants and
brute force on tiny inputs."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle
from turbins_amd import abi, workload as W


# ---------------------------------------------------------------------------
# ENO_compute_convective_terms (ENO.c:103-245)

def _line(n, order, seed):
    rng = np.random.default_rng(seed)
    x = np.cumsum(0.5 + rng.random(n))          # strictly increasing, non-uniform
    return x


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_eno_exact_on_polynomials(order, sign):
    """Divided-difference ENO differentiates polynomials of degree <= order
    exactly on any positions, whichever stencil it picks."""
    n = 24
    x = _line(n, order, 3)
    coef = [0.3, -1.2, 0.7, 0.25][: order + 1]
    D0 = sum(c * x ** p for p, c in enumerate(coef))
    dD = sum(p * c * x ** (p - 1) for p, c in enumerate(coef) if p > 0)
    V = sign * (1.0 + 0.1 * np.sin(x))
    out = oracle.eno_line(order, D0, x, V)
    s = slice(order, n - order)
    np.testing.assert_allclose(out[s], (V * dD)[s], rtol=1e-9, atol=1e-9)


def test_eno_order_is_not_higher_than_claimed():
    """A cubic is not reproduced at order 2 (guards against a silently used D3)."""
    n = 24
    x = _line(n, 2, 5)
    D0 = x ** 3
    out = oracle.eno_line(2, D0, x, np.ones(n))
    assert np.max(np.abs(out[2:-2] - 3 * x[2:-2] ** 2)) > 1e-3


def test_eno_upwinding_and_stencil_choice():
    """Upwind first difference (V>0 uses the left pair) and the smaller-|D2|
    stencil: on a step the ENO stencil avoids the jump."""
    x = np.arange(16, dtype=float)
    step = np.where(x >= 8, 1.0, 0.0)
    # order 1: V>0 -> backward difference, V<0 -> forward difference
    out_p = oracle.eno_line(1, step, x, np.ones(16))
    out_m = oracle.eno_line(1, step, x, -np.ones(16))
    assert out_p[8] == 1.0 and out_p[9] == 0.0
    assert out_m[7] == -1.0 and out_m[8] == 0.0
    # order 2 at i=9 (V>0, k=8): D2[8] sees the jump, D2[9] does not -> no overshoot
    out2 = oracle.eno_line(2, step, x, np.ones(16))
    assert out2[10] == 0.0 and out2[11] == 0.0


def test_eno_ties_go_left():
    """|D2[k]| == |D2[k+1]| picks D2[k] (the '<=' of ENO.c:202)."""
    x = np.arange(10, dtype=float)
    D0 = np.array([0, 0, 0, 0, 1, 0, 0, 0, 0, 0], dtype=float)   # symmetric bump
    # at i=4, V<0 -> k=4: D2[4] = -1, D2[5] = 0.5 ; at i = 3, V>0 -> k = 2 ...
    out = oracle.eno_line(2, D0, x, -np.ones(10))
    # hand evaluation of i=5 with V<0: k=5, D2[5]=0.5 (=(D1[5]-D1[4])/2=(0-(-1))/2),
    # D2[6]=0 -> right; c=0 -> dQ = D1[5] = 0
    assert out[5] == 0.0
    # i=3, V<0: k=3, D2[3]=0.5, D2[4]=-1 -> left (0.5); dQ = D1[3] + 0.5*((3-3)+(3-4)) = 1 - 0.5
    assert out[3] == -(1.0 - 0.5)


def test_eno_mirror_symmetry():
    """Reflecting x -> -x and V -> -V reflects the convective term."""
    n, order = 20, 3
    x = _line(n, order, 9)
    rng = np.random.default_rng(1)
    D0 = np.sin(x) + 0.1 * rng.random(n)
    V = 0.3 + rng.random(n)
    a = oracle.eno_line(order, D0, x, V)
    b = oracle.eno_line(order, D0[::-1], -x[::-1], -V[::-1])[::-1]
    s = slice(order, n - order)
    np.testing.assert_allclose(a[s], b[s], rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_eno_conv_linear_fields_closed_form(order):
    """With every velocity component and the scalar linear in (x, y, z), the
    convective term at each node is u(x_q).grad(q) exactly away from the
    walls.  This pins the face/centre transposes (ENO.c interpolation of the
    other components to the node), the cell-centre averages and the sign and
    axis of every directional term of ENO_compute_convective_terms."""
    c = W.case("lock64", nx=16, ny=12, nz=10, lx=2.0)
    o = oracle.Oracle(W.config_for(c, eno_order=order))
    o.setup_geometry(None)
    sh = o.shape
    K, J, I = np.indices(sh)
    A = np.array([[0.4, 0.5, -0.3, 0.2],      # u = a0 + a1 x + a2 y + a3 z
                  [-0.1, 0.2, 0.6, -0.4],
                  [0.3, -0.5, 0.1, 0.7],
                  [1.0, 2.0, 3.0, -1.0]])     # c
    pos = {0: ("xu", "yu", "zu"), 1: ("xv", "yv", "zv"), 2: ("xw", "yw", "zw"), 3: ("xc", "yc", "zc")}
    X = {q: [o.grid(n)[ix] for n, ix in zip(pos[q], (I, J, K))] for q in range(4)}

    def lin(row, q):
        x, y, z = X[q]
        return A[row, 0] + A[row, 1] * x + A[row, 2] * y + A[row, 3] * z

    for q, name in enumerate("UVWC"):
        o.upload(name, lin(q, q).ravel())
    m = order + 2
    inner = (I >= m) & (I < sh[2] - 1 - m) & (J >= m) & (J < sh[1] - 1 - m) & (K >= m) & (K < sh[0] - 1 - m)
    for q, name in enumerate(["CONV_U", "CONV_V", "CONV_W", "CONV_C"]):
        o.eno_conv(q)
        got = o.download(name).reshape(sh)
        want = sum(lin(d, q) * A[q, 1 + d] for d in range(3))
        np.testing.assert_allclose(got[inner], want[inner], rtol=1e-11, atol=1e-12)


# ---------------------------------------------------------------------------
# MyMath_GaussJordan_matrix_inv (MyMath.c:117-209)

def test_gauss_jordan_matches_numpy():
    rng = np.random.default_rng(2)
    for _ in range(5):
        a = rng.standard_normal((8, 8)) + 4 * np.eye(8)
        np.testing.assert_allclose(oracle.gauss_jordan(a), np.linalg.inv(a), rtol=1e-10, atol=1e-12)
    # a pivot-demanding matrix (zero diagonal)
    p = np.eye(8)[::-1] + 0.01 * rng.standard_normal((8, 8))
    np.testing.assert_allclose(oracle.gauss_jordan(p) @ p, np.eye(8), atol=1e-10)


# ---------------------------------------------------------------------------
# grid (Grid.c:470-708)

def test_grid_coordinates_and_metrics_closed_form():
    c = W.case("bump_small")
    o = oracle.Oracle(W.config_for(c))
    nx, lx = c["nx"], c["lx"]
    dx = lx / nx
    np.testing.assert_allclose(o.grid("xu"), np.arange(nx + 1) * dx, rtol=0, atol=1e-13)
    np.testing.assert_allclose(o.grid("xc"), (np.arange(nx + 1) + 0.5) * dx, rtol=0, atol=1e-13)
    # (dx/deta)^-1 = dEta/dx = 1/Lx on a uniform grid, at centres and faces
    np.testing.assert_allclose(o.grid("mxc"), 1.0 / lx, rtol=1e-12)
    np.testing.assert_allclose(o.grid("mxu"), 1.0 / lx, rtol=1e-12)
    np.testing.assert_allclose(o.grid("myv"), 1.0 / c["ly"], rtol=1e-12)
    np.testing.assert_allclose(o.grid("dy_c"), c["ly"] / c["ny"], rtol=1e-12)


# ---------------------------------------------------------------------------
# operators

def _state(o, c, seed=3):
    """Random fields; velocities vanish on their own box-boundary nodes so the
    net boundary flux is zero and the Neumann Poisson problem is consistent."""
    NT = o.NT
    for a, name in enumerate(["U", "V", "W", "C", "DPDX", "DPDY", "DPDZ"]):
        f = 0.1 * W.uniform_pm1(seed, a * NT, NT)
        if a < 3:
            f[o.flags(a) == abi.BOUNDARY] = 0.0
        o.upload(name, f)


def test_poisson_operator_structure_and_quadratic():
    """Symmetric on the cells, zero row sums except the pinned corner, and
    -Laplacian of a quadratic exact in the interior."""
    c = W.case("lock64", nx=16, ny=8, nz=8, lx=2.0)
    o = oracle.Oracle(W.config_for(c))
    A = o.poisson_matrix().tocsr()
    fl = o.flags(abi.Q_C)
    cells = np.where(fl != abi.BOUNDARY)[0]
    Ac = A[cells][:, cells]
    assert abs(Ac - Ac.T).max() < 1e-10 * abs(Ac).max()
    rs = np.asarray(Ac.sum(axis=1)).ravel()
    corner = np.argmax(np.abs(rs))
    assert np.sum(np.abs(rs) > 1e-9) == 1      # only the pinned corner
    NX, NY = c["nx"] + 1, c["ny"] + 1
    kk, jj, ii = np.unravel_index(cells[corner], (c["nz"] + 1, NY, NX))
    assert (ii, jj, kk) == (c["nx"] - 1, c["ny"] - 1, c["nz"] - 1)
    xc, yc, zc = o.grid("xc"), o.grid("yc"), o.grid("zc")
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    phi = (X ** 2 + 2 * Y ** 2 - Z ** 2).ravel()
    y = o.poisson_apply(phi)
    k, j, i = np.meshgrid(np.arange(c["nz"] + 1), np.arange(NY), np.arange(NX), indexing="ij")
    inner = ((i > 0) & (i < c["nx"] - 1) & (j > 0) & (j < c["ny"] - 1) & (k > 0) & (k < c["nz"] - 1)).ravel()
    np.testing.assert_allclose(y[inner], -(2 + 4 - 2), rtol=1e-9)


def test_poisson_solve_matches_scipy():
    c = W.case("bump_small", nx=16, ny=8, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    _state(o, c)
    o.poisson_rhs(1e-3)
    b = o.download("PRHS")
    st = o.poisson_solve(rtol=1e-13)
    A = o.poisson_matrix().tocsc()
    x_ref = sla.spsolve(A, b)
    x = o.download("PHI")
    assert st.converged
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-9


@pytest.mark.parametrize("q", [0, 1, 2, 3])
def test_helmholtz_solve_matches_scipy(q):
    c = W.case("turbidity_small", nx=16, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    _state(o, c)
    dt = 1e-3
    for qq in range(4):
        o.eno_conv(qq)
    o.set_rhs(q, dt)
    b = o.download("RHS_" + "UVWC"[q])
    st = o.helmholtz_solve(q, dt, rtol=1e-14)
    A = o.helmholtz_matrix(q, dt).tocsc()
    x_ref = sla.spsolve(A, b)
    x = o.download("UVWC"[q])
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-9, st.res


@pytest.mark.parametrize("q", [0, 1, 2, 3])
def test_helmholtz_gs_sweeps_converge_to_the_solution(q):
    """The red-black sweep pairs (with the Jacobi update of the IB rows) are a
    convergent splitting of the assembled Helmholtz rows: the direct solution
    is their fixed point and repeated pairs reach it from a random start."""
    c = W.case("turbidity_small", nx=16, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    dt = 1e-3
    rng = np.random.default_rng(3 + q)
    b = rng.standard_normal(o.NT)
    A = o.helmholtz_matrix(q, dt).tocsc()
    x_ref = sla.spsolve(A, b)
    one = o.helmholtz_gs_sweeps(q, dt, b, x_ref, 1)
    assert np.max(np.abs(one - x_ref)) <= 1e-13 * np.max(np.abs(x_ref))
    x = o.helmholtz_gs_sweeps(q, dt, b, rng.standard_normal(o.NT), 30)
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-12
    # and one pair of an unconverged state is not the identity
    x3 = o.helmholtz_gs_sweeps(q, dt, b, np.zeros(o.NT), 3)
    x4 = o.helmholtz_gs_sweeps(q, dt, b, x3, 1)
    assert np.linalg.norm(x4 - x_ref) < 0.2 * np.linalg.norm(x3 - x_ref)


def test_helmholtz_rows_on_constant_field():
    """Interior FLUID rows: (1/dt) I - (1/Re) Laplacian applied to a constant
    gives const/dt (no-slip wall rows differ, IB rows give ap_solid*(1+sum)."""
    c = W.case("lock64", nx=16, ny=8, nz=8, lx=2.0)
    cfg = W.config_for(c)
    o = oracle.Oracle(cfg)
    dt = 1e-3
    y = o.helmholtz_apply(abi.Q_C, dt, np.full(o.NT, 2.0))
    fl = o.flags(abi.Q_C)
    # concentration: Neumann walls everywhere -> every FLUID row gives 2/dt
    np.testing.assert_allclose(y[fl == abi.FLUID], 2.0 / dt, rtol=1e-12)


def test_projection_divergence_identity():
    """div(u - dt grad phi) = -dt (b - A phi): ties Pressure_set_RHS, the
    Poisson rows and Pressure_project_velocity together."""
    c = W.case("bump_small", nx=32, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    _state(o, c)
    dt = 1e-3
    o.poisson_rhs(dt)
    b = o.download("PRHS")
    rng = np.random.default_rng(4)
    fl = o.flags(abi.Q_C)
    phi = np.where(fl != abi.BOUNDARY, rng.standard_normal(o.NT), 0.0)
    corner = np.ravel_multi_index((c["nz"] - 1, c["ny"] - 1, c["nx"] - 1), (c["nz"] + 1, c["ny"] + 1, c["nx"] + 1))
    phi[fl != abi.BOUNDARY] -= phi[corner]
    o.upload("PHI", phi)
    r = b - o.poisson_apply(phi)
    o.project(dt)
    o.div_max()
    div = o.download("DIV")
    m = fl != abi.BOUNDARY
    np.testing.assert_allclose(div[m], np.abs(dt * r[m]), rtol=1e-7, atol=1e-9 * np.abs(dt * r).max())


def test_real_pressure_rhs_closed_form_and_poisson_rhs():
    """Pressure_compute_real_pressure's right-hand side: for the gradient of
    p = x^2 + 2 y^2 - z^2 sampled on the faces, -div grad p = -(2 + 4 - 2)
    on every non-BOUNDARY cell (central differences of linear functions are
    exact), 0 on BOUNDARY cells; and it is Pressure_set_RHS's expression on
    the same arrays with 1/dt = 1."""
    c = W.case("bump_small", nx=32, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    xu, yv, zw = o.grid("xu"), o.grid("yv"), o.grid("zw")
    Z, Y, X = np.meshgrid(zw, yv, xu, indexing="ij")
    o.upload("DPDX", (2.0 * X).ravel())
    o.upload("DPDY", (4.0 * Y).ravel())
    o.upload("DPDZ", (-2.0 * Z).ravel())
    o.real_pressure_rhs()
    b = o.download("PRHS")
    fl = o.flags(abi.Q_C)
    m = fl != abi.BOUNDARY
    np.testing.assert_allclose(b[m], -4.0, rtol=1e-9)
    assert np.all(b[~m] == 0.0)
    rng = np.random.default_rng(11)
    f = [rng.standard_normal(o.NT) for _ in range(3)]
    for name, a in zip(["DPDX", "DPDY", "DPDZ"], f):
        o.upload(name, a)
    o.real_pressure_rhs()
    b1 = o.download("PRHS")
    for name, a in zip(["U", "V", "W"], f):
        o.upload(name, a)
    o.poisson_rhs(1.0)
    assert np.array_equal(b1, o.download("PRHS"))


def test_diagnostics_closed_form():
    """Total kinetic energy and suspended mass over FLUID cells: without
    topography every interior cell is FLUID, so a uniform velocity (2, -1, 0.5)
    and c = 1 give 0.5 |U|^2 Lx Ly Lz and Lx Ly Lz; a velocity linear in x
    has exact midpoint cell-centre values, so the energy is the closed-form
    sum of 0.5 x_c^2 over the cells."""
    c = W.case("lock64", nx=16, ny=8, nz=8, lx=2.0)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(None)
    vol = c["lx"] * c["ly"] * c["lz"]
    for name, val in (("U", 2.0), ("V", -1.0), ("W", 0.5), ("C", 1.0)):
        o.upload(name, np.full(o.NT, val))
    d = o.diagnostics()
    np.testing.assert_allclose(d["suspended_mass"], vol, rtol=1e-12)
    np.testing.assert_allclose(d["kinetic_energy"], 0.5 * (4.0 + 1.0 + 0.25) * vol, rtol=1e-12)
    xu = o.grid("xu")
    NX, NY, NZ = c["nx"] + 1, c["ny"] + 1, c["nz"] + 1
    o.upload("U", np.tile(xu, NY * NZ))
    o.upload("V", np.zeros(o.NT))
    o.upload("W", np.zeros(o.NT))
    xc = 0.5 * (xu[:-1] + xu[1:])
    dv = (c["lx"] / c["nx"]) * (c["ly"] / c["ny"]) * (c["lz"] / c["nz"])
    want = 0.5 * np.sum(xc[: c["nx"]] ** 2) * dv * c["ny"] * c["nz"]
    np.testing.assert_allclose(o.diagnostics()["kinetic_energy"], want, rtol=1e-12)
    # Extract_Stokes_dissipation_rate: |v_s| alpha times the particle volume
    o.upload("C", np.full(o.NT, 0.25))
    np.testing.assert_allclose(o.stokes_dissipation(True, -0.02, 0.5), 0.02 * 0.5 * 0.25 * vol, rtol=1e-12)
    assert o.stokes_dissipation(False, -0.02, 0.5) == 0.0


def test_sedimentation_rate_brute_force():
    """Grid_set_interface_y_index + Conc_compute_sedimentation_rate: without
    topography every column's interface is row 0 (FLUID), so the rate is
    |v_s| times the integral of c over the bottom row; over a bump it is
    checked against a brute-force scan of the flags."""
    c = W.case("lock64", nx=16, ny=8, nz=8, lx=2.0)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(None)
    o.upload("C", np.full(o.NT, 0.5))
    tot, b = o.sedimentation_rate(True, -0.02)
    np.testing.assert_allclose(tot, 0.02 * 0.5 * c["lx"] * c["lz"], rtol=1e-12)
    assert o.sedimentation_rate(False, -0.02)[0] == 0.0
    c = W.case("bump_small", nx=32, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    rng = np.random.default_rng(5)
    cf = rng.random(o.NT)
    o.upload("C", cf)
    tot, b = o.sedimentation_rate(True, 0.03)
    NX, NY, NZ = c["nx"] + 1, c["ny"] + 1, c["nz"] + 1
    fl = o.flags(abi.Q_C).reshape(NZ, NY, NX)
    cc = cf.reshape(NZ, NY, NX)
    dx, dz = c["lx"] / c["nx"], c["lz"] / c["nz"]
    want = np.zeros((NZ, NX))
    for k in range(NZ):
        for i in range(NX):
            for j in range(NY):
                if fl[k, j, i] == abi.IMMERSED or (j == 0 and fl[k, j, i] == abi.FLUID):
                    want[k, i] = 0.03 * cc[k, j, i] * o.grid("dx_c")[i] * o.grid("dz_c")[k]
                    break
    assert np.count_nonzero(want) > 0
    np.testing.assert_allclose(b, want, rtol=1e-14, atol=0)
    np.testing.assert_allclose(tot, want.sum(), rtol=1e-13)


def test_vcycle_symmetric_and_contracting():
    c = W.case("bump_small", nx=32, ny=16, nz=16)
    o = oracle.Oracle(W.config_for(c))
    n = c["nx"] * c["ny"] * c["nz"]
    rng = np.random.default_rng(5)
    r1, r2 = rng.standard_normal(n), rng.standard_normal(n)
    r1 -= r1.mean(); r2 -= r2.mean()
    z1, z2 = o.vcycle(r1), o.vcycle(r2)
    assert abs(z1 @ r2 - r1 @ z2) < 1e-10 * abs(z1 @ r2)
    assert z1 @ r1 > 0


def test_gmg_hierarchy_rule():
    """Semi-coarsening of the directions at most half as coarse as the
    coarsest one, full coarsening once the spacings are within a factor 2,
    down to <= 64 cells (512x256x128 over 16x1x1: h = 1/32, 1/256, 1/128)."""
    c = W.case("bump512x256")
    o = oracle.Oracle(W.config_for(c))
    dims = [tuple(d) for d in o.gmg_levels()]
    assert dims == [(512, 256, 128), (512, 128, 64), (512, 64, 32), (512, 32, 32), (256, 16, 16), (128, 8, 8),
                    (64, 4, 4), (32, 2, 2), (16, 2, 2)]


# ---------------------------------------------------------------------------
# geometry: classification and signed distance

def test_flat_bottom_reinit_is_a_fixed_point_and_layer_is_analytic():
    """y - h is already a distance: reinitialisation leaves it unchanged, and
    the IMMERSED layer is the first solid row under y = 2.001 dy."""
    c = W.case("bump_small", nx=16, ny=16, nz=8)
    o = oracle.Oracle(W.config_for(c))
    h = np.zeros((c["nz"] + 1, c["nx"] + 1))
    o.setup_geometry(h)
    dy = c["ly"] / c["ny"]
    yc = o.grid("yc")
    sdf = o.download("SDF_C").reshape(c["nz"] + 1, c["ny"] + 1, c["nx"] + 1)
    want = np.broadcast_to((yc[:-1] - 2.001 * dy)[None, :, None], sdf[:-1, :-1, :-1].shape)
    np.testing.assert_allclose(sdf[:-1, :-1, :-1], want, atol=1e-12)
    fl = o.flags(abi.Q_C).reshape(sdf.shape)
    assert np.all(fl[:-1, 0, :-1] == abi.SOLID)
    assert np.all(fl[:-1, 1, :-1] == abi.IMMERSED)
    assert np.all(fl[:-1, 2:-1, :-1] == abi.FLUID)


def _reinit_sdf(ny, guard):
    c = W.case("turbidity1024", nx=128, ny=ny, nz=16)
    cfg = W.config_for(c)
    cfg.reinit_guard = guard
    o = oracle.Oracle(cfg)
    h = W.topography(c)
    o.setup_geometry(h)
    sh = o.shape
    sdf = o.download("SDF_C").reshape(sh)[:-1, :-1, :-1]
    dy = c["ly"] / ny
    vert = o.grid("yc")[None, :-1, None] - (h[:-1, None, :-1] + 2.001 * dy)
    return sdf, vert, dy


def test_reinit_guard_scaled_equals_reference_on_coarse_grids():
    """With every spacing >= 0.01/1.5 (ny = 128 over a unit height) the
    scaled stencil-width guard is the reference's absolute 0.01 exactly."""
    a, _, _ = _reinit_sdf(128, abi.REINIT_SCALED)
    b, _, _ = _reinit_sdf(128, abi.REINIT_REFERENCE)
    assert np.array_equal(a, b)


def test_reinit_guard_fine_grid():
    """At ny = 256 (2 dy < 0.01) the absolute guard drops the interior y
    difference next to a near-flat bottom, D = phi/0.01 overshoots and the
    march diverges from the bottom wall (Levelset.c:282,426-456 as written);
    the scaled guard keeps a distance function: the sign of y - h everywhere,
    and within a fraction of a cell of it at the interface."""
    sdf, vert, dy = _reinit_sdf(256, abi.REINIT_REFERENCE)
    assert np.abs(sdf).max() > 1e3
    sdf, vert, dy = _reinit_sdf(256, abi.REINIT_SCALED)
    assert np.abs(sdf).max() < 1.0
    far = np.abs(vert) > 1e-9
    assert np.array_equal(np.sign(sdf[far]), np.sign(vert[far]))
    band = np.abs(vert) < 2 * dy
    assert np.abs(sdf - vert)[band].max() < 0.5 * dy


@pytest.mark.parametrize("name", ["bump_small", "turbidity_small"])
def test_classification_brute_force(name):
    """IMMERSED <=> sdf < 0 with a sign change to a (clamped) 6-neighbour;
    BUFFER_SOLID <=> SOLID within Chebyshev distance 2 of a pre-pass IMMERSED."""
    c = W.case(name)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    shp = (c["nz"] + 1, c["ny"] + 1, c["nx"] + 1)
    for q, nm in enumerate(["SDF_U", "SDF_V", "SDF_W", "SDF_C"]):
        s = o.download(nm).reshape(shp)
        fl = o.flags(q).reshape(shp)
        inner = fl != abi.BOUNDARY
        K, J, I = np.indices(shp)
        nb = []
        for ax, d in [(2, 1), (2, -1), (1, 1), (1, -1), (0, 1), (0, -1)]:
            idx = [K, J, I]
            lim = shp[ax] - 2   # neighbours clamped to [0, N-2] (Grid.c:1553-1587)
            idx[ax] = np.clip(idx[ax] + d, 0, lim)
            nb.append(s[tuple(idx)])
        change = np.zeros(shp, bool)
        for v in nb:
            change |= (s * v) < 0
        imm = inner & (s < 0) & change
        assert np.array_equal(fl == abi.IMMERSED, imm)
        assert np.array_equal(fl == abi.FLUID, inner & (s >= 0))
        if q < 3:
            pre_imm = imm
            pad = np.pad(pre_imm, 2)
            near = np.zeros(shp, bool)
            for dk in range(5):
                for dj in range(5):
                    for di in range(5):
                        near |= pad[dk:dk + shp[0], dj:dj + shp[1], di:di + shp[2]]
            solid = inner & (s < 0) & ~change
            assert np.array_equal(fl == abi.BUFFER_SOLID, solid & near)


@pytest.mark.parametrize("name", ["bump_small", "turbidity_small"])
def test_ib_coefficient_invariants(name):
    c = W.case(name)
    o = oracle.Oracle(W.config_for(c))
    o.setup_geometry(W.topography(c))
    for q in range(4):
        nodes = o.ib_table(q)
        assert len(nodes) == int(np.sum(o.flags(q) == abi.IMMERSED))
        for nd in nodes:
            n = np.array(nd.n)
            assert abs(np.linalg.norm(n) - 1.0) < 1e-12
            im, ctl, img = np.array(nd.im), np.array(nd.control), np.array(nd.image)
            ds, df = np.linalg.norm(im - ctl), np.linalg.norm(img - ctl)
            if nd.n_fluid == 8:
                if q < 3:   # Dirichlet: trilinear weights sum to 1, scaled by ds/df
                    assert abs(sum(nd.coef) - ds / df) < 1e-12
                else:       # Neumann: a constant field satisfies the row
                    self_in = any(nd.fi[g] == nd.i and nd.fj[g] == nd.j and nd.fk[g] == nd.k for g in range(8))
                    if not self_in:
                        assert abs(sum(nd.coef) + 1.0) < 1e-9
                # image point inside the box of fluid corners
                assert nd.fi[7] == nd.fi[0] + 1 and nd.fj[7] == nd.fj[0] + 1 and nd.fk[7] == nd.fk[0] + 1


# ---------------------------------------------------------------------------

def test_splitmix64_published_vectors():
    z = W.splitmix64(0, 0, 3)
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_library_exports_header_symbols():
    """The C-ABI library loads (no GPU needed) and exports every entry point
    declared in include/turbins_gpu.h."""
    import ctypes
    import turbins_amd
    L = ctypes.CDLL(turbins_amd.LIB_PATH)
    names = turbins_amd.exported_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n


def test_oracle_step_runs_and_converges():
    c = W.case("lock64", nx=32, ny=8, nz=8, lx=4.0)
    o = oracle.Oracle(W.config_for(c, pois_rtol=1e-10))
    o.setup_geometry(None)
    for k, v in W.initial_fields(c, [o.flags(q) for q in range(4)]).items():
        o.upload(k, v)
    for _ in range(2):
        st = o.step(c["dt"])
    assert all(p.converged for p in st.pois)
    assert all(h.converged for stage in st.helm for h in stage)


# ---------------------------------------------------------------------------
# learned-guess network: oracle (exact mode) against torch autograd (FP64)
# and finite differences

NET_LAYERS = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (48, 16), (16, 1)]


def _net_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    w = []
    for ci, co in NET_LAYERS:
        a = np.sqrt(6.0 / (27 * ci)) * scale
        w.append(rng.uniform(-a, a, 27 * ci * co))
        w.append(rng.uniform(-0.1, 0.1, co))
    w = np.concatenate(w)
    assert w.size == oracle.net_param_count()
    return w


def _torch_net(w, in3):
    import torch
    import torch.nn.functional as F
    ws, o = [], 0
    for ci, co in NET_LAYERS:
        W = torch.tensor(w[o:o + 27 * ci * co].reshape(3, 3, 3, ci, co)).permute(4, 3, 0, 1, 2).contiguous()
        o += 27 * ci * co
        b = torch.tensor(w[o:o + co])
        o += co
        W.requires_grad_(True)
        b.requires_grad_(True)
        ws.append((W, b))
    x = torch.zeros((1, 16) + in3.shape[:3], dtype=torch.float64)
    x[0, :3] = torch.tensor(in3).permute(3, 0, 1, 2)

    def cv(l, t, relu=True):
        y = F.conv3d(t, ws[l][0], ws[l][1], padding=1)
        return F.relu(y) if relu else y

    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    x0 = cv(0, x)
    x1 = cv(1, F.avg_pool3d(x0, 2))
    x1b = cv(2, x1)
    x2 = cv(3, F.avg_pool3d(x1b, 2))
    x2b = cv(4, x2)
    y1 = cv(5, torch.cat([up(x2b), x1b], 1))
    y0 = cv(6, torch.cat([up(y1), x0], 1))
    out = cv(7, y0, relu=False)[0, 0]
    return out, ws


def test_net_oracle_matches_torch_forward_and_gradient():
    import torch
    rng = np.random.default_rng(1)
    nz, ny, nx = 8, 8, 12
    in3 = rng.standard_normal((nz, ny, nx, 3))
    tgt = rng.standard_normal((nz, ny, nx))
    w = _net_params(2)
    out_o = oracle.net_forward(w, in3, round_bf16=False)
    loss_o, g_o = oracle.net_gradient(w, in3, tgt, round_bf16=False)
    out_t, ws = _torch_net(w, in3)
    np.testing.assert_allclose(out_o, out_t.detach().numpy(), rtol=1e-11, atol=1e-12)
    loss_t = ((out_t - torch.tensor(tgt)) ** 2).mean()
    loss_t.backward()
    assert abs(loss_o - loss_t.item()) < 1e-12 * abs(loss_t.item())
    g_t = np.concatenate([np.concatenate([W.grad.permute(2, 3, 4, 1, 0).reshape(-1).numpy(), b.grad.numpy()])
                          for W, b in ws])
    np.testing.assert_allclose(g_o, g_t, rtol=1e-9, atol=1e-13)


def test_net_oracle_gradient_finite_differences():
    rng = np.random.default_rng(3)
    nz, ny, nx = 4, 8, 8
    in3 = rng.standard_normal((nz, ny, nx, 3))
    tgt = rng.standard_normal((nz, ny, nx))
    w = _net_params(4)
    _, g = oracle.net_gradient(w, in3, tgt, round_bf16=False)
    for q in rng.choice(w.size, 12, replace=False):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[q] += h
        wm[q] -= h
        lp, _ = oracle.net_gradient(wp, in3, tgt, round_bf16=False)
        lm, _ = oracle.net_gradient(wm, in3, tgt, round_bf16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-9, (q, fd, g[q])


def test_net_oracle_rounding_mode_is_close_to_exact():
    rng = np.random.default_rng(5)
    in3 = rng.standard_normal((8, 8, 8, 3))
    w = _net_params(6)
    a = oracle.net_forward(w, in3, round_bf16=False)
    b = oracle.net_forward(w, in3, round_bf16=True)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 3e-2
    assert not np.array_equal(a, b)
