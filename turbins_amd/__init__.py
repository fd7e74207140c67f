"""turbins_amd -- MI355X (gfx950) implementation of the TURBINS per-timestep
hot path behind the C-ABI declared in include/turbins_gpu.h.

`Solver` is a thin ctypes binding over libturbins_gpu.so.  There is no
fallback: if the library is missing or cannot be loaded, constructing a
Solver raises.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libturbins_gpu.so")
_LIB = None


def build(force=False):
    """Compile the HIP library for gfx950 in place (make -C csrc)."""
    args = ["make", "-s", "-C", os.path.join(_HERE, "csrc"), "-j8"]
    if force:
        subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), "clean"], check=True)
    subprocess.run(args, check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    # TURBINS_GPU_LIB selects another build of the same sources (the
    # host-sanitizer build of tests/test_sanitizers.py)
    path = os.environ.get("TURBINS_GPU_LIB") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run turbins_amd.build() first")
    L = C.CDLL(path)
    P = C.c_void_p
    L.tb_create.restype = P
    L.tb_create.argtypes = [C.POINTER(abi.Config)]
    L.tb_destroy.argtypes = [P]
    L.tb_last_error.restype = C.c_char_p
    L.tb_last_error.argtypes = [P]
    L.tb_device_bytes.argtypes = [P, C.POINTER(C.c_int64)]
    L.tb_setup_geometry.argtypes = [P, P]
    L.tb_download_flags.argtypes = [P, C.c_int, P]
    L.tb_ib_count.argtypes = [P, C.c_int]
    L.tb_download_ib_table.argtypes = [P, C.c_int, P, C.c_int]
    L.tb_reinit_iterations.argtypes = [P]
    L.tb_grid_array.argtypes = [P, C.c_int, P, C.c_int]
    L.tb_upload_field.argtypes = [P, C.c_int, P]
    L.tb_download_field.argtypes = [P, C.c_int, P]
    L.tb_eno_conv.argtypes = [P, C.c_int]
    L.tb_set_rhs.argtypes = [P, C.c_int, C.c_double]
    L.tb_eno_conv_rhs.argtypes = [P, C.c_double]
    L.tb_helmholtz_apply.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_int]
    L.tb_helmholtz_gs_sweeps.argtypes = [P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
    L.tb_helmholtz_solve.argtypes = [P, C.c_int, C.c_double, C.POINTER(abi.SolveStats)]
    L.tb_poisson_rhs.argtypes = [P, C.c_double]
    L.tb_poisson_apply.argtypes = [P, C.c_int, C.c_int]
    L.tb_poisson_solve.argtypes = [P, C.POINTER(abi.SolveStats)]
    L.tb_poisson_solve_guess.argtypes = [P, C.c_int, C.POINTER(abi.SolveStats)]
    L.tb_project.argtypes = [P, C.c_double]
    L.tb_div_max.argtypes = [P, C.POINTER(C.c_double), P]
    L.tb_store_old.argtypes = [P]
    L.tb_rk_average.argtypes = [P, C.c_double, C.c_double]
    L.tb_step.argtypes = [P, C.c_double, C.POINTER(abi.StepStats)]
    L.tb_cfl_dt.argtypes = [P, C.c_double, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tb_set_conc_fields.argtypes = [P, C.c_int, P, C.c_int]
    L.tb_set_open_boundaries.argtypes = [P, C.POINTER(abi.OpenBC)]
    L.tb_inflow_profile.argtypes = [P, C.c_double, C.POINTER(C.c_double)]
    L.tb_set_boundary_plane.argtypes = [P, C.c_int, C.c_int, P]
    L.tb_get_boundary_plane.argtypes = [P, C.c_int, C.c_int, P]
    L.tb_get_influx.argtypes = [P, C.POINTER(C.c_double)]
    L.tb_outflow_convective.argtypes = [P, C.c_double]
    L.tb_viscous_dissipation.argtypes = [P, C.POINTER(C.c_double)]
    L.tb_kinetic_potential_conversion.argtypes = [P, C.POINTER(C.c_double)]
    L.tb_potential_energies.argtypes = [P, C.POINTER(abi.Energies)]
    L.tb_wall_shear.argtypes = [P, C.c_int, P, P, P, P]
    L.tb_ave_height_x.argtypes = [P, P]
    L.tb_front_location.argtypes = [P, C.c_double, C.POINTER(C.c_double)]
    L.tb_deposit_integrate.argtypes = [P, C.c_double, C.c_double]
    L.tb_deposit_height.argtypes = [P, P]
    L.tb_deposit_set.argtypes = [P, P]
    L.tb_deposit_dumped.argtypes = [P, C.c_double, P]
    L.tb_select_conc.argtypes = [P, C.c_int]
    L.tb_total_concentration.argtypes = [P]
    L.tb_set_poisson_mode.argtypes = [P, C.c_int, C.c_double]
    L.tb_set_helmholtz_rtol.argtypes = [P, C.c_double]
    L.tb_set_mixed_inner.argtypes = [P, C.c_double, C.c_int]
    L.tb_set_guess_history.argtypes = [P, C.c_int]
    L.tb_set_helmholtz_method.argtypes = [P, C.c_int]
    L.tb_set_graphs.argtypes = [P, C.c_int]
    L.tb_set_mg_finest_sweeps.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
    L.tb_set_mg_smoother.argtypes = [P, C.c_int]
    L.tb_petsc_write_array.argtypes = [C.c_char_p, P, C.c_int64]
    L.tb_petsc_read_array.restype = C.c_int64
    L.tb_petsc_read_array.argtypes = [C.c_char_p, P, C.c_int64]
    L.tb_real_pressure_rhs.argtypes = [P]
    L.tb_diagnostics.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tb_stokes_dissipation.argtypes = [P, C.c_double, C.POINTER(C.c_double)]
    L.tb_sedimentation_rate.argtypes = [P, C.POINTER(C.c_double), P]
    L.tb_real_pressure.argtypes = [P, C.POINTER(abi.SolveStats)]
    L.tb_save_field.argtypes = [P, C.c_int, C.c_char_p]
    L.tb_load_field.argtypes = [P, C.c_int, C.c_char_p]
    L.tb_write_resume_primary.argtypes = [C.c_char_p, C.POINTER(abi.ResumeInfo)]
    L.tb_read_resume_primary.argtypes = [C.c_char_p, C.POINTER(abi.ResumeInfo)]
    L.tb_write_resume.argtypes = [P, C.c_char_p, C.POINTER(abi.ResumeInfo)]
    L.tb_read_resume.argtypes = [P, C.c_char_p, C.POINTER(abi.ResumeInfo)]
    L.tb_div_max_noalloc.argtypes = [C.POINTER(abi.Box)] + [P] * 11 + [P]
    L.tb_profile_enable.argtypes = [P, C.c_int]
    L.tb_profile_read.argtypes = [P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.tb_net_create.restype = P
    L.tb_net_create.argtypes = [C.c_uint64]
    L.tb_net_destroy.argtypes = [P]
    L.tb_net_last_error.restype = C.c_char_p
    L.tb_net_last_error.argtypes = [P]
    L.tb_net_param_count.restype = C.c_int64
    L.tb_net_get_weights.argtypes = [P, P, C.c_int64]
    L.tb_net_set_weights.argtypes = [P, P, C.c_int64]
    L.tb_net_attach.argtypes = [P, P]
    L.tb_net_gate.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]
    L.tb_net_predict_field.argtypes = [P, C.c_int]
    L.tb_net_reserve.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int]
    L.tb_net_add_snapshot.argtypes = [P, P]
    L.tb_net_record.argtypes = [P, P]
    L.tb_net_snapshot_count.argtypes = [P]
    L.tb_net_train.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.tb_net_eval_snapshots.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.tb_net_forward.argtypes = [P, P, C.c_int, C.c_int, C.c_int, P]
    L.tb_net_gradient.argtypes = [P, P, P, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), P]
    _LIB = L
    return L


def exported_symbols():
    """Entry points declared in include/turbins_gpu.h."""
    hdr = os.path.join(os.path.dirname(_HERE), "include", "turbins_gpu.h")
    import re
    names = []
    for line in open(hdr):
        m = re.match(r"^(?:int|int64_t|void|const char \*|tb_ctx \*|tb_net \*)\s*\*?(tb_\w+)\(", line.strip())
        if m:
            names.append(m.group(1))
    return names


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _io_check(rc, what, path):
    if rc < 0:
        raise OSError(f"{what} {path}: {abi.IO_ERRORS.get(int(rc), rc)}")
    return rc


def petsc_write(path, a):
    """Write a 1-D float64 array as a PETSc binary Vec (VecView on a binary
    viewer, Resume.c:102-103): big-endian class id 1211214, length, values."""
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    _io_check(lib().tb_petsc_write_array(os.fsencode(path), _ptr(a), a.size), "petsc_write", path)


def petsc_read(path, capacity=None):
    """Read a PETSc binary Vec (VecLoadIntoVector, Resume.c:134) into a
    float64 array; capacity defaults to what the file size allows."""
    if capacity is None:
        capacity = max(0, (os.path.getsize(path) - 8) // 8)
    out = np.empty(capacity, dtype=np.float64)
    n = _io_check(lib().tb_petsc_read_array(os.fsencode(path), _ptr(out), capacity), "petsc_read", path)
    return out[:n]


def write_resume_primary(directory, time, dt, dt_old, iteration, output_iter):
    info = abi.ResumeInfo(time, dt, dt_old, iteration, output_iter)
    _io_check(lib().tb_write_resume_primary(os.fsencode(directory), C.byref(info)), "write_resume_primary",
              directory)


def read_resume_primary(directory):
    """The Resume_primary.dat record; zero_step says that dt or dt_old read
    as 0 (the record's six decimals lose any step below 5e-7)."""
    info = abi.ResumeInfo()
    rc = _io_check(lib().tb_read_resume_primary(os.fsencode(directory), C.byref(info)), "read_resume_primary",
                   directory)
    return {"time": info.time, "dt": info.dt, "dt_old": info.dt_old, "iteration": info.iter,
            "output_iter": info.output_iter, "zero_step": rc == 1}


class TBError(RuntimeError):
    pass


def div_max_noalloc(box, mxc, myc, mzc, u, v, w, c, flags_c, div, out_max, out_index, stream=None):
    """tb_div_max_noalloc on device addresses (ints): enqueues the
    divergence check on `stream` (a hipStream_t address, None: default)."""
    rc = lib().tb_div_max_noalloc(C.byref(box), mxc, myc, mzc, u, v, w, c, flags_c, div, out_max, out_index,
                                  stream)
    if rc < 0:
        raise TBError(f"tb_div_max_noalloc: {abi.NOALLOC_ERRORS.get(int(rc), rc)}")
    return rc


class Solver:
    """One device context (one HIP stream, all fields resident in HBM)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.L = lib()
        self.h = self.L.tb_create(C.byref(cfg))
        if not self.h:
            raise TBError("tb_create failed (bad config or no usable GPU)")
        self.shape = (cfg.nz + 1, cfg.ny + 1, cfg.nx + 1)
        self.NT = int(np.prod(self.shape))

    def _ok(self, rc, what):
        # status of the last call: 0 ok, > 0 finished without convergence
        # or with a fallback taken (the stats say which)
        self.last_rc = rc
        if rc < 0:
            raise TBError(f"{what}: {self.L.tb_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.tb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_bytes(self):
        v = C.c_int64()
        self._ok(self.L.tb_device_bytes(self.h, C.byref(v)), "tb_device_bytes")
        return v.value

    def setup_geometry(self, h_xz):
        if h_xz is None:
            return self._ok(self.L.tb_setup_geometry(self.h, None), "tb_setup_geometry")
        h = np.ascontiguousarray(h_xz, dtype=np.float64)
        assert h.size == (self.cfg.nz + 1) * (self.cfg.nx + 1)
        return self._ok(self.L.tb_setup_geometry(self.h, _ptr(h)), "tb_setup_geometry")

    def reinit_iterations(self):
        return self.L.tb_reinit_iterations(self.h)

    def flags(self, q):
        out = np.zeros(self.NT, dtype=np.uint8)
        self._ok(self.L.tb_download_flags(self.h, q, _ptr(out)), "tb_download_flags")
        return out

    def ib_table(self, q):
        n = self.L.tb_ib_count(self.h, q)
        arr = (abi.IBNode * max(n, 1))()
        self._ok(self.L.tb_download_ib_table(self.h, q, C.cast(arr, C.c_void_p), n), "tb_download_ib_table")
        return list(arr)[:n]

    GRID_ARRAYS = ["xc", "yc", "zc", "xu", "yu", "zu", "xv", "yv", "zv", "xw", "yw", "zw",
                   "dx_c", "dy_c", "dz_c", "mxc", "myc", "mzc", "mxu", "myv", "mzw"]

    def grid(self, name):
        """One 1-D grid array (coordinates, spacings, metrics) of the context."""
        out = np.zeros(max(self.shape) + 1)
        n = self._ok(self.L.tb_grid_array(self.h, self.GRID_ARRAYS.index(name), _ptr(out), out.size), "tb_grid_array")
        return out[:n].copy()

    def upload(self, name, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        assert a.size == self.NT
        self._ok(self.L.tb_upload_field(self.h, abi.F[name], _ptr(a)), "tb_upload_field")

    def download(self, name):
        out = np.zeros(self.NT)
        self._ok(self.L.tb_download_field(self.h, abi.F[name], _ptr(out)), "tb_download_field")
        return out

    def eno_conv(self, q):
        return self._ok(self.L.tb_eno_conv(self.h, q), "tb_eno_conv")

    def set_rhs(self, q, dt):
        return self._ok(self.L.tb_set_rhs(self.h, q, dt), "tb_set_rhs")

    def eno_conv_rhs(self, dt):
        """eno_conv(-1) and set_rhs(q, dt) for all four quantities in one pass."""
        return self._ok(self.L.tb_eno_conv_rhs(self.h, dt), "tb_eno_conv_rhs")

    def helmholtz_apply(self, q, dt, x):
        self.upload("SCRATCH0", x)
        self._ok(self.L.tb_helmholtz_apply(self.h, q, dt, abi.F["SCRATCH0"], abi.F["SCRATCH1"]), "tb_helmholtz_apply")
        return self.download("SCRATCH1")

    def helmholtz_gs_sweeps(self, q, dt, b, x, pairs):
        """`pairs` red-black Gauss-Seidel sweep pairs of A_q(dt) x = b from x
        (the smoother of the Helmholtz solve), through SCRATCH0/SCRATCH1."""
        self.upload("SCRATCH0", b)
        self.upload("SCRATCH1", x)
        self._ok(self.L.tb_helmholtz_gs_sweeps(self.h, q, dt, abi.F["SCRATCH0"], abi.F["SCRATCH1"], pairs),
                 "tb_helmholtz_gs_sweeps")
        return self.download("SCRATCH1")

    def helmholtz_solve(self, q, dt):
        st = abi.SolveStats()
        self._ok(self.L.tb_helmholtz_solve(self.h, q, dt, C.byref(st)), "tb_helmholtz_solve")
        return st

    def poisson_rhs(self, dt):
        return self._ok(self.L.tb_poisson_rhs(self.h, dt), "tb_poisson_rhs")

    def poisson_apply(self, x):
        self.upload("SCRATCH0", x)
        self._ok(self.L.tb_poisson_apply(self.h, abi.F["SCRATCH0"], abi.F["SCRATCH1"]), "tb_poisson_apply")
        return self.download("SCRATCH1")

    def poisson_solve(self):
        st = abi.SolveStats()
        self._ok(self.L.tb_poisson_solve(self.h, C.byref(st)), "tb_poisson_solve")
        return st

    def poisson_solve_guess(self, guess):
        """Pressure_solve from a caller-supplied initial guess (uploaded to
        SCRATCH0); the previous PHI is kept if its residual is smaller."""
        self.upload("SCRATCH0", guess)
        st = abi.SolveStats()
        self._ok(self.L.tb_poisson_solve_guess(self.h, abi.F["SCRATCH0"], C.byref(st)), "tb_poisson_solve_guess")
        return st

    def diagnostics(self):
        """Total kinetic energy and suspended mass over the FLUID cells."""
        ke, m = C.c_double(), C.c_double()
        self._ok(self.L.tb_diagnostics(self.h, C.byref(ke), C.byref(m)), "tb_diagnostics")
        return {"kinetic_energy": ke.value, "suspended_mass": m.value}

    def sedimentation_rate(self):
        """Total sedimentation rate and the per-column bottom rates (NZ+1, NX+1)."""
        tot = C.c_double()
        b = np.zeros((self.cfg.nz + 1, self.cfg.nx + 1))
        self._ok(self.L.tb_sedimentation_rate(self.h, C.byref(tot), _ptr(b)), "tb_sedimentation_rate")
        return tot.value, b

    def stokes_dissipation(self, conc_alpha=1.0):
        r = C.c_double()
        self._ok(self.L.tb_stokes_dissipation(self.h, conc_alpha, C.byref(r)), "tb_stokes_dissipation")
        return r.value

    def real_pressure_rhs(self):
        return self._ok(self.L.tb_real_pressure_rhs(self.h), "tb_real_pressure_rhs")

    def real_pressure(self):
        """Pressure_compute_real_pressure: PHI <- the pressure whose gradient
        is closest to DPDX/Y/Z (a Poisson solve); it seeds the next solve."""
        st = abi.SolveStats()
        self._ok(self.L.tb_real_pressure(self.h, C.byref(st)), "tb_real_pressure")
        return st

    def project(self, dt):
        return self._ok(self.L.tb_project(self.h, dt), "tb_project")

    def div_max(self):
        v = C.c_double()
        ijk = np.zeros(3, dtype=np.int32)
        self._ok(self.L.tb_div_max(self.h, C.byref(v), _ptr(ijk)), "tb_div_max")
        return v.value, tuple(int(a) for a in ijk)

    def store_old(self):
        return self._ok(self.L.tb_store_old(self.h), "tb_store_old")

    def rk_average(self, a, b):
        return self._ok(self.L.tb_rk_average(self.h, a, b), "tb_rk_average")

    def step(self, dt):
        st = abi.StepStats()
        self._ok(self.L.tb_step(self.h, dt, C.byref(st)), "tb_step")
        return st

    def cfl_dt(self, alpha, dt_max):
        """(dt, rate): dt = min(dt_max, alpha / max over FLUID cells of
        |u_c|/dx + |v_c|/dy + |w_c|/dz)."""
        dt, r = C.c_double(), C.c_double()
        self._ok(self.L.tb_cfl_dt(self.h, alpha, dt_max, C.byref(dt), C.byref(r)), "tb_cfl_dt")
        return dt.value, r.value

    def set_conc_fields(self, fields, hindered=False):
        """Concentration fields from dicts (type, pe, v_settle, alpha,
        phi_max); field 0 replaces the config's; hindered is
        params->hindered_settling.  Selects field 0."""
        arr = abi.conc_fields(fields)
        return self._ok(self.L.tb_set_conc_fields(self.h, len(fields), C.cast(arr, C.c_void_p), 1 if hindered else 0),
                        "tb_set_conc_fields")

    def select_conc(self, n):
        """Field n becomes the one the concentration operators and the C,
        CONV_C, RHS_C transfers act on (the reference's c[iconc])."""
        return self._ok(self.L.tb_select_conc(self.h, n), "tb_select_conc")

    def set_open_boundaries(self, **kw):
        """Open or close the x faces: inflow, outflow, transient_inflow,
        u_conv (params->U), time_c (abi.open_bc); planes start at zero."""
        bc = abi.open_bc(**kw)
        return self._ok(self.L.tb_set_open_boundaries(self.h, C.byref(bc)), "tb_set_open_boundaries")

    def inflow_profile(self, time):
        """Inflow_u_velocity_profile + Inflow_conc_profile at time; returns
        the influx."""
        fl = C.c_double()
        self._ok(self.L.tb_inflow_profile(self.h, time, C.byref(fl)), "tb_inflow_profile")
        return fl.value

    def set_plane(self, which, q, plane):
        a = np.ascontiguousarray(plane, dtype=np.float64).reshape(-1)
        assert a.size == (self.cfg.ny + 1) * (self.cfg.nz + 1)
        return self._ok(self.L.tb_set_boundary_plane(self.h, which, q, _ptr(a)), "tb_set_boundary_plane")

    def get_plane(self, which, q):
        out = np.zeros((self.cfg.ny + 1) * (self.cfg.nz + 1))
        self._ok(self.L.tb_get_boundary_plane(self.h, which, q, _ptr(out)), "tb_get_boundary_plane")
        return out

    def influx(self):
        fl = C.c_double()
        self._ok(self.L.tb_get_influx(self.h, C.byref(fl)), "tb_get_influx")
        return fl.value

    def outflow_convective(self, dt):
        return self._ok(self.L.tb_outflow_convective(self.h, dt), "tb_outflow_convective")

    def total_concentration(self):
        """Conc_compute_total_concentration: the v buoyancy source of
        several fields, sum of alpha_i c_i."""
        return self._ok(self.L.tb_total_concentration(self.h), "tb_total_concentration")

    # output-stage diagnostics (column arrays have shape (nz+1, nx+1))
    def _cols(self):
        return np.zeros((self.cfg.nz + 1, self.cfg.nx + 1))

    def viscous_dissipation(self):
        r = C.c_double()
        self._ok(self.L.tb_viscous_dissipation(self.h, C.byref(r)), "tb_viscous_dissipation")
        return r.value

    def kinetic_potential_conversion(self):
        r = C.c_double()
        self._ok(self.L.tb_kinetic_potential_conversion(self.h, C.byref(r)), "tb_kinetic_potential_conversion")
        return r.value

    def potential_energies(self):
        e = abi.Energies()
        self._ok(self.L.tb_potential_energies(self.h, C.byref(e)), "tb_potential_energies")
        return {n: getattr(e, n) for n, _ in abi.Energies._fields_}

    def wall_shear(self, quadratic=False):
        out = {n: self._cols() for n in ("tau", "u_tan", "v_tan", "w_tan")}
        self._ok(self.L.tb_wall_shear(self.h, 1 if quadratic else 0, _ptr(out["tau"]), _ptr(out["u_tan"]),
                                      _ptr(out["v_tan"]), _ptr(out["w_tan"])), "tb_wall_shear")
        return out

    def ave_height_x(self):
        h = np.zeros(self.cfg.nx + 1)
        self._ok(self.L.tb_ave_height_x(self.h, _ptr(h)), "tb_ave_height_x")
        return h

    def front_location(self, limit=0.01):
        x = C.c_double()
        self._ok(self.L.tb_front_location(self.h, limit, C.byref(x)), "tb_front_location")
        return x.value

    def deposit_integrate(self, weight, dt):
        return self._ok(self.L.tb_deposit_integrate(self.h, weight, dt), "tb_deposit_integrate")

    def deposit_height(self):
        h = self._cols()
        self._ok(self.L.tb_deposit_height(self.h, _ptr(h)), "tb_deposit_height")
        return h

    def deposit_set(self, h):
        h = np.ascontiguousarray(h, dtype=np.float64).reshape(self.cfg.nz + 1, self.cfg.nx + 1)
        return self._ok(self.L.tb_deposit_set(self.h, _ptr(h)), "tb_deposit_set")

    def deposit_dumped(self, weight):
        h = self._cols()
        self._ok(self.L.tb_deposit_dumped(self.h, weight, _ptr(h)), "tb_deposit_dumped")
        return h

    def set_poisson_mode(self, mode, rtol):
        return self._ok(self.L.tb_set_poisson_mode(self.h, mode, rtol), "tb_set_poisson_mode")

    def set_helmholtz_rtol(self, rtol):
        return self._ok(self.L.tb_set_helmholtz_rtol(self.h, rtol), "tb_set_helmholtz_rtol")

    def set_helmholtz_method(self, method):
        return self._ok(self.L.tb_set_helmholtz_method(self.h, method), "tb_set_helmholtz_method")

    def set_graphs(self, on):
        return self._ok(self.L.tb_set_graphs(self.h, 1 if on else 0), "tb_set_graphs")

    def set_mixed_inner(self, reduction_floor, maxit=40):
        return self._ok(self.L.tb_set_mixed_inner(self.h, reduction_floor, maxit), "tb_set_mixed_inner")

    def set_guess_history(self, h):
        return self._ok(self.L.tb_set_guess_history(self.h, h), "tb_set_guess_history")

    def save_field(self, name, path):
        """Field -> PETSc binary Vec file (VecView of its G_data Vec)."""
        return self._ok(self.L.tb_save_field(self.h, abi.F[name], os.fsencode(path)), "tb_save_field")

    def load_field(self, name, path):
        """PETSc binary Vec file -> field (VecLoadIntoVector)."""
        return self._ok(self.L.tb_load_field(self.h, abi.F[name], os.fsencode(path)), "tb_load_field")

    def write_resume(self, directory, time, dt, dt_old, iteration, output_iter):
        """The reference's restart set (Resume_write_data + Resume_write_primary_data)."""
        info = abi.ResumeInfo(time, dt, dt_old, iteration, output_iter)
        return self._ok(self.L.tb_write_resume(self.h, os.fsencode(directory), C.byref(info)), "tb_write_resume")

    def read_resume(self, directory):
        """Load the restart set; returns the Resume_primary.dat record."""
        info = abi.ResumeInfo()
        self._ok(self.L.tb_read_resume(self.h, os.fsencode(directory), C.byref(info)), "tb_read_resume")
        return {"time": info.time, "dt": info.dt, "dt_old": info.dt_old, "iteration": info.iter,
                "output_iter": info.output_iter}

    def set_mg_finest_sweeps(self, pre32=0, post32=0, pre64=0, post64=0):
        """Sweeps of the finest multigrid level per V-cycle side (0: the
        config's mg_pre / mg_post, which the coarser levels keep)."""
        return self._ok(self.L.tb_set_mg_finest_sweeps(self.h, pre32, post32, pre64, post64),
                        "tb_set_mg_finest_sweeps")

    def set_mg_smoother(self, blocked=True):
        """Smoother of the levels between the finest and the LDS cycle: one
        blocked launch per V-cycle side (default) or one launch per half
        sweep; both give the same bits."""
        return self._ok(self.L.tb_set_mg_smoother(self.h, 1 if blocked else 0), "tb_set_mg_smoother")

    def profile(self, on=True):
        return self._ok(self.L.tb_profile_enable(self.h, 1 if on else 0), "tb_profile_enable")

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_int64()
        self._ok(self.L.tb_profile_read(self.h, kernel, C.byref(ms), C.byref(n)), "tb_profile_read")
        return ms.value, n.value

    def attach_net(self, net):
        """Use `net` (a Net) for the learned initial guess; None detaches."""
        self._net = net
        return self._ok(self.L.tb_net_attach(self.h, net.h if net is not None else None), "tb_net_attach")

    def net_gate(self):
        """(open, trials, iterations saved per solve, inference ms per call)"""
        o, t, sv, ms = C.c_int(), C.c_int(), C.c_double(), C.c_double()
        self._ok(self.L.tb_net_gate(self.h, C.byref(o), C.byref(t), C.byref(sv), C.byref(ms)), "tb_net_gate")
        return {"open": bool(o.value), "trials": t.value, "saved_iterations_per_solve": sv.value,
                "net_ms_per_call": ms.value}

    def record_snapshots(self, net):
        """Every later Poisson solve adds its snapshot to net's store."""
        return self._ok(self.L.tb_net_record(self.h, net.h if net is not None else None), "tb_net_record")

    def predict_guess(self, name="SCRATCH0"):
        """Learned guess for the current PRHS and pressure history -> field."""
        self._ok(self.L.tb_net_predict_field(self.h, abi.F[name]), "tb_net_predict_field")
        return self.download(name)


class Net:
    """Learned Poisson initial guess: weights, snapshot store and training,
    all on the device (tb_net_* entry points)."""

    def __init__(self, seed=0):
        self.L = lib()
        self.h = self.L.tb_net_create(seed)
        if not self.h:
            raise TBError("tb_net_create failed")
        self.np = int(self.L.tb_net_param_count())

    def _ok(self, rc, what):
        if rc < 0:
            raise TBError(f"{what}: {self.L.tb_net_last_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.tb_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_weights(self):
        w = np.zeros(self.np, dtype=np.float32)
        self._ok(self.L.tb_net_get_weights(self.h, _ptr(w), self.np), "tb_net_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.np:
            raise ValueError(f"expected {self.np} weights, got {w.size}")
        self._ok(self.L.tb_net_set_weights(self.h, _ptr(w), self.np), "tb_net_set_weights")

    def save(self, path):
        self.get_weights().tofile(path)

    def load(self, path):
        self.set_weights(np.fromfile(path, dtype=np.float32))

    def reserve(self, capacity, nx, ny, nz):
        self._ok(self.L.tb_net_reserve(self.h, capacity, nx, ny, nz), "tb_net_reserve")

    def add_snapshot(self, solver):
        self._ok(self.L.tb_net_add_snapshot(self.h, solver.h), "tb_net_add_snapshot")

    def snapshot_count(self):
        return int(self.L.tb_net_snapshot_count(self.h))

    def train(self, steps, batch, crop, lr=1e-3, seed=0):
        l0, l1 = C.c_double(), C.c_double()
        self._ok(self.L.tb_net_train(self.h, steps, batch, crop[0], crop[1], crop[2], lr, seed,
                                     C.byref(l0), C.byref(l1)), "tb_net_train")
        return l0.value, l1.value

    def eval_snapshots(self, first, count):
        v = C.c_double()
        self._ok(self.L.tb_net_eval_snapshots(self.h, first, count, C.byref(v)), "tb_net_eval_snapshots")
        return v.value

    def forward(self, in3):
        """in3: (nz, ny, nx, 3) float32 -> (nz, ny, nx)"""
        in3 = np.ascontiguousarray(in3, dtype=np.float32)
        nz, ny, nx, _ = in3.shape
        out = np.zeros((nz, ny, nx), dtype=np.float32)
        self._ok(self.L.tb_net_forward(self.h, _ptr(in3), nx, ny, nz, _ptr(out)), "tb_net_forward")
        return out

    def gradient(self, in3, target):
        in3 = np.ascontiguousarray(in3, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        nz, ny, nx, _ = in3.shape
        g = np.zeros(self.np, dtype=np.float32)
        loss = C.c_double()
        self._ok(self.L.tb_net_gradient(self.h, _ptr(in3), _ptr(target), nx, ny, nz, C.byref(loss), _ptr(g)),
                 "tb_net_gradient")
        return loss.value, g
