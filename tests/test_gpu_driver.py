"""The plain-C host driver (turbins_amd/tb_driver, C99 over the C-ABI)
against the CPU oracle: the 64 x 16 x 16 lock exchange for 20 steps at a
fixed dt with an output stage (real pressure) after step 10, and a run whose
dt follows the CFL rule and changes every step.  The driver reads and writes
the reference's restart set (Resume_*.bin, Resume_primary.dat)."""
import os
import subprocess

import numpy as np
import pytest

from turbins_amd import abi, petsc_read, petsc_write, write_resume_primary, workload as W

pytestmark = pytest.mark.gpu

import oracle  # noqa: E402  (checker only)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "turbins_amd", "tb_driver")
RESUME = {"U": "Resume_u.bin", "V": "Resume_v.bin", "W": "Resume_w.bin", "DPDX": "Resume_dpdx.bin",
          "DPDY": "Resume_dpdy.bin", "DPDZ": "Resume_dpdz.bin", "C": "Resume_conc0.bin"}


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _write_start(d, fields, dt):
    os.makedirs(d, exist_ok=True)
    for name, fn in RESUME.items():
        petsc_write(os.path.join(d, fn), fields.get(name, np.zeros_like(fields["U"])))
    write_resume_primary(d, 0.0, dt, 0.0, 0, 0)


def _run(args):
    assert os.path.exists(DRIVER), "build the driver first (make -C turbins_amd/csrc)"
    print("driver:", " ".join(args), flush=True)
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    steps = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("step ")]
    outs = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("output ")]
    return steps, outs


def _oracle_start(c, cfg, fields):
    o = oracle.Oracle(cfg)
    o.setup_geometry(W.topography(c))
    for k, v in fields.items():
        o.upload(k, v)
    return o


def test_c_driver_lock_exchange_fixed_dt(tmp_path):
    c = W.case("lock64")
    cfg = W.config_for(c, helm_rtol=1e-13, pois_rtol=1e-11)
    o = _oracle_start(c, cfg, {})
    fields = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in fields.items():
        o.upload(k, v)
    din, dout = str(tmp_path / "in"), str(tmp_path / "out")
    _write_start(din, fields, c["dt"])
    steps, outs = _run(["--grid", "64", "16", "16", "--box", "8", "1", "1", "--re", "1000", "--sc", "1",
                        "--steps", "20", "--dt", "1e-3", "--helm-rtol", "1e-13", "--pois-rtol", "1e-11",
                        "--in", din, "--out", dout, "--output-every", "10"])
    assert len(steps) == 20 and len(outs) == 2
    for s in range(20):
        o.step(c["dt"])
        if s + 1 in (10, 20):
            # Output_flow_properties' real pressure (Output.c:2198-2204)
            o.real_pressure_rhs()
            o.poisson_solve(rtol=1e-13)
    worst = 0.0
    for name, fn in RESUME.items():
        e = _rel(petsc_read(os.path.join(dout, fn)), o.download(name))
        print("driver vs oracle after 20 steps", name, e)
        worst = max(worst, e)
    assert worst <= 1e-6
    # the output files of the second output stage hold the real pressure
    phi = petsc_read(os.path.join(dout, "output_00002", "PHI.bin"))
    assert _rel(phi, o.download("PHI")) <= 1e-7
    rec = open(os.path.join(dout, "Resume_primary.dat")).read().split()
    assert rec[rec.index("iteration=") + 1] == "20" and rec[rec.index("output_iter=") + 1] == "2"
    o.close()


def test_c_driver_cfl_time_step(tmp_path):
    """dt from the CFL rule (alpha 0.1, capped at 0.05) on the lock exchange:
    the first steps run at the cap, then the accelerating current takes dt
    below it from step 12 on (the diagonal follows each new dt).  The
    device's dt sequence equals the oracle's rule on its own fields to 1e-9
    and the fields after 20 steps agree to 1e-6."""
    c = W.case("lock64")
    cfg = W.config_for(c, helm_rtol=1e-13, pois_rtol=1e-11)
    o = _oracle_start(c, cfg, {})
    fields = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in fields.items():
        o.upload(k, v)
    din, dout = str(tmp_path / "in"), str(tmp_path / "out")
    _write_start(din, fields, 0.05)
    steps, _ = _run(["--grid", "64", "16", "16", "--box", "8", "1", "1", "--re", "1000", "--sc", "1",
                     "--steps", "20", "--cfl", "0.1", "--dt-max", "0.05", "--helm-rtol", "1e-13",
                     "--pois-rtol", "1e-11", "--in", din, "--out", dout])
    dts = [float(s[s.index("dt") + 1]) for s in steps]
    odts = []
    for s in range(20):
        dt, rate = o.cfl_dt(0.1, 0.05)
        odts.append(dt)
        o.step(dt)
    print("dt sequence", dts)
    np.testing.assert_allclose(dts, odts, rtol=1e-9)
    assert dts[0] == 0.05 and dts[-1] < 0.04 and all(a >= b for a, b in zip(dts, dts[1:]))
    worst = 0.0
    for name, fn in RESUME.items():
        e = _rel(petsc_read(os.path.join(dout, fn)), o.download(name))
        print("cfl driver vs oracle", name, e)
        worst = max(worst, e)
    assert worst <= 1e-6
    o.close()


def test_c_driver_open_channel(tmp_path):
    """--inflow --transient-inflow --outflow: the driver sets the inflow
    profile at each step's start time (the transient factor grows from
    1 - exp(-0.2)) and tb_step applies the convective outflow condition; the
    lock exchange then runs as an open channel and the fields after 12
    steps follow the oracle driven the same way to 1e-6."""
    c = W.case("lock64")
    cfg = W.config_for(c, helm_rtol=1e-13, pois_rtol=1e-11)
    o = _oracle_start(c, cfg, {})
    fields = W.initial_fields(c, [o.flags(q) for q in range(4)])
    for k, v in fields.items():
        o.upload(k, v)
    o.set_open_boundaries(inflow=True, outflow=True, transient_inflow=True, time_c=0.05, u_conv=1.0)
    din, dout = str(tmp_path / "in"), str(tmp_path / "out")
    _write_start(din, fields, c["dt"])
    steps, _ = _run(["--grid", "64", "16", "16", "--box", "8", "1", "1", "--re", "1000", "--sc", "1",
                     "--steps", "12", "--dt", "1e-3", "--helm-rtol", "1e-13", "--pois-rtol", "1e-11",
                     "--inflow", "--transient-inflow", "0.05", "--outflow", "1.0", "--in", din, "--out", dout])
    assert len(steps) == 12
    t = 0.0
    for s in range(12):
        o.inflow_profile(t)
        o.step(c["dt"])
        t += c["dt"]
    worst = 0.0
    for name, fn in RESUME.items():
        e = _rel(petsc_read(os.path.join(dout, fn)), o.download(name))
        print("open-channel driver vs oracle", name, e)
        worst = max(worst, e)
    assert worst <= 1e-6
    u = o.download("U").reshape(o.shape)
    assert u[8, 8, 1] > 0.05   # the inflow has entered
    o.close()
