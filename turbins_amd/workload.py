"""Synthetic workloads: the lock-exchange / bump / turbidity-current cases.

The same seeded generator feeds the device path and the CPU checker, so both
see bit-identical inputs.  Random numbers come from SplitMix64 (Steele,
Lea & Flood 2014), evaluated counter-style: the n-th draw of a stream with
seed s is mix(s + (n+1)*0x9E3779B97F4A7C15) -- exactly what the sequential
generator produces -- so any slice can be drawn without the ones before it.
The reference's unseeded rand() perturbation (MyMath.c:81-87,
Conc.c:1391-1405) is not used.
"""
import numpy as np

from . import abi

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n0, count):
    """Draws n0 .. n0+count-1 of the SplitMix64 stream with the given seed."""
    with np.errstate(over="ignore"):
        n = np.arange(n0 + 1, n0 + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + n * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform_pm1(seed, n0, count):
    """Uniform doubles in [-1, 1) from the top 53 bits."""
    z = splitmix64(seed, n0, count)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -52) - 1.0


# ---------------------------------------------------------------------------
# named cases (cells, domain, physics)

CASES = {
    # 3-D lock exchange, CPU-runnable size, no solid
    "lock64": dict(nx=64, ny=16, nz=16, lx=8.0, ly=1.0, lz=1.0, re=1000.0, sc=1.0,
                   dt=1e-3, steps=20, pois_rtol=1e-10, bump="none", x_lock=1.0),
    # lock exchange over an immersed Gaussian bump
    "bump512x128": dict(nx=512, ny=128, nz=128, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                        dt=1e-3, steps=200, pois_rtol=1e-8, bump="gauss1", x_lock=1.0),
    "bump512x256": dict(nx=512, ny=256, nz=128, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                        dt=1e-3, steps=200, pois_rtol=1e-8, bump="gauss1", x_lock=1.0),
    "learn256": dict(nx=256, ny=128, nz=128, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                     dt=1e-3, steps=200, pois_rtol=1e-8, bump="gauss1", x_lock=1.0),
    "turbidity1024": dict(nx=1024, ny=256, nz=128, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                          dt=1e-3, steps=100, pois_rtol=1e-8, bump="gauss8", x_lock=1.0,
                          particle=True, v_settle=0.02),
    # small cases with the same physics, for parity tests
    "bump_small": dict(nx=64, ny=32, nz=16, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                       dt=1e-3, steps=4, pois_rtol=1e-8, bump="gauss1", x_lock=1.0),
    "turbidity_small": dict(nx=64, ny=32, nz=16, lx=16.0, ly=1.0, lz=1.0, re=2000.0, sc=1.0,
                            dt=1e-3, steps=4, pois_rtol=1e-8, bump="gauss8", x_lock=1.0,
                            particle=True, v_settle=0.02),
}


def case(name, **over):
    c = dict(CASES[name])
    c.update(over)
    return c


def config_for(c, **kw):
    """tb_config for a case. buoyancy defaults to the physical 1.0 (the
    shipped reference multiplies it by 0.0 at Velocity.c:1883)."""
    args = dict(
        nx=c["nx"], ny=c["ny"], nz=c["nz"], lx=c["lx"], ly=c["ly"], lz=c["lz"],
        re=c["re"], pe=c["re"] * c["sc"], eno_order=c.get("eno_order", 2),
        wall_noslip=c.get("walls", (1, 1, 1, 1, 1, 1)),
        buoyancy=c.get("buoyancy", 1.0),
        conc_type=abi.PARTICLE if c.get("particle") else abi.SALINITY,
        # the reference flips the sign of the settling speed (Input.c:481)
        v_settle=-abs(c.get("v_settle", 0.0)),
        sedimentation=1 if c.get("particle") else 0,
        pois_rtol=c.get("pois_rtol", 1e-10),
    )
    args.update(kw)
    return abi.make_config(**args)


def cell_centres(n, L):
    d = L / n
    return (np.arange(n + 1) + 0.5) * d


def topography(c, seed=7):
    """Bottom height at cell centres, shape (nz+1, nx+1), or None."""
    kind = c.get("bump", "none")
    if kind == "none":
        return None
    xc = cell_centres(c["nx"], c["lx"])
    zc = cell_centres(c["nz"], c["lz"])
    X, Z = np.meshgrid(xc, zc)
    if kind == "gauss1":
        h0, xb, w = c.get("bump_height", 0.3), c.get("bump_x", 6.0), c.get("bump_width", 0.5)
        return h0 * np.exp(-(X - xb) ** 2 / (2.0 * w * w))
    if kind == "gauss8":
        r = 0.5 * (uniform_pm1(seed, 0, 32) + 1.0)   # U[0,1)
        h = np.zeros_like(X)
        for b in range(8):
            hb = 0.05 + 0.25 * r[4 * b]
            wb = 0.2 + 0.8 * r[4 * b + 1]
            xb = 3.0 + (c["lx"] - 4.0) * r[4 * b + 2]
            zb = c["lz"] * r[4 * b + 3]
            h += hb * np.exp(-((X - xb) ** 2 + (Z - zb) ** 2) / (2.0 * wb * wb))
        return h
    raise ValueError(kind)


def initial_fields(c, flags=None, seed=1):
    """u, v, w: 1e-3 * U(-1,1) per node; c: lock released at x_lock,
    0.5*(1 - tanh((x - x_lock)/(2 dx))).  Values sit on non-BOUNDARY nodes;
    if per-quantity flags are given, only FLUID nodes are set (as
    Conc_initialize does for c, Conc.c:1318-1440)."""
    nx, ny, nz = c["nx"], c["ny"], c["nz"]
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    NT = NX * NY * NZ
    out = {}
    shape = (NZ, NY, NX)
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    interior = {
        "U": (i >= 1) & (i <= NX - 2) & (j <= NY - 2) & (k <= NZ - 2),
        "V": (i <= NX - 2) & (j >= 1) & (j <= NY - 2) & (k <= NZ - 2),
        "W": (i <= NX - 2) & (j <= NY - 2) & (k >= 1) & (k <= NZ - 2),
        "C": (i <= NX - 2) & (j <= NY - 2) & (k <= NZ - 2),
    }
    for q, name in enumerate("UVW"):
        vals = 1e-3 * uniform_pm1(seed, q * NT, NT).reshape(shape)
        m = interior[name]
        if flags is not None:
            m = m & (flags[q].reshape(shape) == abi.FLUID)
        out[name] = np.where(m, vals, 0.0)
    dx = c["lx"] / nx
    xc = cell_centres(nx, c["lx"])
    prof = 0.5 * (1.0 - np.tanh((xc - c.get("x_lock", 1.0)) / (2.0 * dx)))
    cc = np.broadcast_to(prof[None, None, :], shape).copy()
    m = interior["C"]
    if flags is not None:
        m = m & (flags[3].reshape(shape) == abi.FLUID)
    out["C"] = np.where(m, cc, 0.0)
    for name in out:
        out[name] = np.ascontiguousarray(out[name].reshape(-1))
    return out
