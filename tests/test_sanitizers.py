"""The CPU oracle and the library's host code under AddressSanitizer and
UndefinedBehaviorSanitizer.

`make -C oracle asan` builds the oracle's C sources with
-fsanitize=address,undefined (host code only); a child process with the
sanitizer runtimes preloaded loads that build (TURBINS_ORACLE_LIB) and runs
every oracle entry point the parity tests rely on -- geometry with immersed
boundaries, ENO of every order, right-hand sides, Helmholtz and Poisson
operators and solves, multigrid, steps with several concentration fields and
hindered settling, open x faces, the diagnostics and the network's forward
and gradient --
on small boxes.  `make -C turbins_amd/csrc asan` builds the library with
the sanitizers on its host code (device code as usual); a child with
clang's ASan runtime preloaded runs the host paths that need no GPU: the
PETSc vector and restart-record formats with well-formed and malformed
files, and the argument checks of the entry points.  Any out-of-bounds
access, use after free or undefined behaviour aborts the child with a
report."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import numpy as np
import oracle
from turbins_amd import abi, workload as W

for name, dims in (("bump_small", (24, 16, 8)), ("turbidity_small", (16, 12, 8))):
    for order in (1, 2, 3):
        c = W.case(name, nx=dims[0], ny=dims[1], nz=dims[2])
        cfg = W.config_for(c, eno_order=order, helm_rtol=1e-12, pois_rtol=1e-10)
        o = oracle.Oracle(cfg)
        o.setup_geometry(W.topography(c))
        init = W.initial_fields(c, [o.flags(q) for q in range(4)])
        for k, v in init.items():
            o.upload(k, v)
        rng = np.random.default_rng(order)
        for n in "UVW":
            o.upload(n, init[n] + 0.01 * rng.uniform(-1, 1, o.NT))
        for q in range(4):
            o.eno_conv(q)
            o.set_rhs(q, 1e-3)
            o.helmholtz_apply(q, 1e-3, rng.uniform(size=o.NT))
            o.helmholtz_gs_sweeps(q, 1e-3, rng.uniform(size=o.NT), np.zeros(o.NT), 2)
        o.poisson_apply(rng.uniform(size=o.NT))
        o.vcycle(rng.uniform(size=dims[0] * dims[1] * dims[2]))
        fields = [dict(type=abi.PARTICLE, pe=cfg.pe, v_settle=-0.02, alpha=0.6, phi_max=0.3),
                  dict(type=abi.SALINITY, pe=500.0, alpha=0.4)]
        o.set_conc_fields(fields, hindered=True)
        o.select_conc(1)
        o.upload("C", init["C"] * 0.5)
        o.select_conc(0)
        for _ in range(2):
            o.step(c["dt"])
        o.real_pressure_rhs()
        o.poisson_solve(rtol=1e-10)
        o.div_max()
        o.cfl_dt(0.5, 0.1)
        o.diagnostics()
        o.viscous_dissipation()
        o.kinetic_potential_conversion()
        o.deposit_integrate(0.6, 1e-3)
        o.potential_energies()
        o.deposit_dumped(0.6)
        o.wall_shear(False)
        o.wall_shear(True)
        o.ave_height_x()
        o.front_location(0.01)
        o.sedimentation_rate(1, 0.02)
        o.stokes_dissipation(1, 0.02, 0.6)
        o.set_open_boundaries(inflow=True, outflow=True, transient_inflow=True, time_c=0.1)
        o.inflow_profile(0.05)
        o.set_plane(1, 0, o.get_plane(0, 0))
        o.outflow_convective(1e-3)
        for q in range(4):
            o.eno_conv(q)
            o.set_rhs(q, 1e-3)
            o.helmholtz_apply(q, 1e-3, rng.uniform(size=o.NT))
        o.step(c["dt"])
        o.viscous_dissipation()
        o.div_max()
        o.set_conc_fields(fields[:1])
        o.close()

npar = oracle.net_param_count()
w = np.random.default_rng(1).normal(0, 0.1, npar)
x = np.random.default_rng(2).normal(size=(8, 8, 8, 3))
oracle.net_forward(w, x, 1)
oracle.net_gradient(w, x, np.zeros((8, 8, 8)), 0)
print("sanitized oracle run complete")
"""


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_and_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TURBINS_ORACLE_LIB=os.path.join(ROOT, "oracle", "liboracle_asan.so"), OMP_NUM_THREADS="4",
               PYTHONPATH=ROOT + os.pathsep + env.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-c", SCRIPT], env=env, capture_output=True, text=True, timeout=900,
                       cwd=ROOT)
    assert r.returncode == 0 and "sanitized oracle run complete" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])


LIB_SCRIPT = r"""
import os, tempfile
import numpy as np
import turbins_amd as T
from turbins_amd import abi, workload as W

assert T.lib()._name.endswith("libturbins_gpu_asan.so")
d = tempfile.mkdtemp()
rng = np.random.default_rng(0)
for n in (0, 1, 7, 4099, (1 << 17) + 3):
    a = rng.normal(size=n)
    p = os.path.join(d, "v.bin")
    T.petsc_write(p, a)
    assert np.array_equal(T.petsc_read(p), a)
    if n > 1:
        try:
            T.petsc_read(p, capacity=n - 1)
            raise SystemExit("oversized vector accepted")
        except OSError:
            pass
        raw = open(p, "rb").read()
        open(p, "wb").write(raw[:-5])            # truncated
        try:
            T.petsc_read(p, capacity=n)
            raise SystemExit("short file accepted")
        except OSError:
            pass
        open(p, "wb").write(b"\x00\x12\x7b\x4f" + raw[4:])   # wrong class id
        try:
            T.petsc_read(p, capacity=n)
            raise SystemExit("bad class id accepted")
        except OSError:
            pass
for dt in (1e-3, 1e-7):
    T.write_resume_primary(d, 12.5, dt, 1e-3, 40, 2)
    r = T.read_resume_primary(d)
    assert r["zero_step"] == (dt < 5e-7)
for text in ("", "time=\n", "time=\n1.0\ndt=\n0.1\n", "garbage" * 1000):
    open(os.path.join(d, "Resume_primary.dat"), "w").write(text)
    try:
        T.read_resume_primary(d)
        raise SystemExit("malformed record accepted")
    except OSError:
        pass
bad = W.config_for(W.case("lock64"), nx=2)
assert not T.lib().tb_create(bad)
import ctypes as C
assert T.lib().tb_div_max_noalloc(None, *([None] * 12)) == -1           # TB_NOALLOC_BAD_ARGUMENT
box = abi.Box(16, 8, 8, 10, 1 / 16, 1 / 8, 1 / 8)                       # pitch < nx + 1
fake = [C.c_void_p(64)] * 11
assert T.lib().tb_div_max_noalloc(C.byref(box), *fake, None) == -1
box = abi.Box(1 << 12, 1 << 10, 1 << 10, (1 << 12) + 16, 1.0, 1.0, 1.0)  # more than 2^31 points
assert T.lib().tb_div_max_noalloc(C.byref(box), *fake, None) == -2
print("sanitized library run complete")
"""


def test_library_host_code_under_asan_and_ubsan():
    rts = sorted(__import__("glob").glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rts:
        pytest.skip("clang's ASan runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "turbins_amd", "csrc"), "-j8", "asan"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=rts[-1], ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TURBINS_GPU_LIB=os.path.join(ROOT, "turbins_amd", "libturbins_gpu_asan.so"),
               PYTHONPATH=ROOT + os.pathsep + env.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-c", LIB_SCRIPT], env=env, capture_output=True, text=True, timeout=900,
                       cwd=ROOT)
    assert r.returncode == 0 and "sanitized library run complete" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
