"""ctypes mirror of include/turbins_gpu.h (structs, constants, field ids)."""
import ctypes as C

Q_U, Q_V, Q_W, Q_C = 0, 1, 2, 3
OUTSIDE, SOLID, FLUID, IMMERSED, BOUNDARY, BUFFER_SOLID = 1, 2, 3, 5, 7, 11
SALINITY, PARTICLE = 1, 2
POIS_FP64_PCG, POIS_MIXED_IR, POIS_LEARNED_IR = 0, 1, 2
HELM_GS, HELM_BICGSTAB = 0, 1
REINIT_SCALED, REINIT_REFERENCE = 0, 1

KERNELS = ["ENO_MOM", "ENO_CONC", "HELM_OP", "BICG_UPD", "POIS_OP", "RBGS0", "RBGS0_F32",
           "PCG_UPD", "PROJECT", "RHS", "NET", "CG32_OP", "CG32_UPD", "HELM_GS",
           "MG_LEVELS", "MG_TRANSFER", "MG_SMALL", "GUESS", "IB"]
K = {name: i for i, name in enumerate(KERNELS)}

FIELDS = [
    "U", "V", "W", "C", "DPDX", "DPDY", "DPDZ", "PHI",
    "CONV_U", "CONV_V", "CONV_W", "CONV_C",
    "RHS_U", "RHS_V", "RHS_W", "RHS_C",
    "PRHS", "DIV", "SDF_U", "SDF_V", "SDF_W", "SDF_C",
    "SCRATCH0", "SCRATCH1",
]
F = {name: i for i, name in enumerate(FIELDS)}
F_COUNT = len(FIELDS)


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("re", C.c_double), ("pe", C.c_double),
        ("eno_order", C.c_int32),
        ("wall_noslip", C.c_int32 * 6),
        ("buoyancy", C.c_double),
        ("conc_type", C.c_int32),
        ("v_settle", C.c_double),
        ("sedimentation", C.c_int32),
        ("helm_rtol", C.c_double), ("helm_maxit", C.c_int32),
        ("pois_rtol", C.c_double), ("pois_maxit", C.c_int32),
        ("pois_mode", C.c_int32),
        ("bottom_offset", C.c_int32),
        ("mg_pre", C.c_int32), ("mg_post", C.c_int32),
        ("reinit_guard", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class IBNode(C.Structure):
    _fields_ = [
        ("i", C.c_int32), ("j", C.c_int32), ("k", C.c_int32),
        ("n_fluid", C.c_int32),
        ("fi", C.c_int32 * 8), ("fj", C.c_int32 * 8), ("fk", C.c_int32 * 8),
        ("coef", C.c_double * 8),
        ("boundary_coef", C.c_double),
        ("n", C.c_double * 3), ("im", C.c_double * 3),
        ("control", C.c_double * 3), ("image", C.c_double * 3),
    ]


class Box(C.Structure):
    """tb_box: a field layout for the context-free entry points."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pitch", C.c_int32),
        ("deta", C.c_double), ("dxi", C.c_double), ("dphi", C.c_double),
    ]


NOALLOC_ERRORS = {-1: "bad argument", -2: "box larger than 2^31 points", -3: "no device workspace",
                  -4: "kernel launch failed"}


NCONC_MAX = 4


class ConcField(C.Structure):
    """tb_conc_field: one concentration field (params->conc_type[i], Pe[i],
    V_s0[i] as used, Conc_alpha[i], Phi[i])."""
    _fields_ = [
        ("type", C.c_int32), ("reserved", C.c_int32),
        ("pe", C.c_double), ("v_settle", C.c_double), ("alpha", C.c_double), ("phi_max", C.c_double),
    ]


class OpenBC(C.Structure):
    """tb_open_bc: the open x faces (params->inflow, outflow,
    transient_inflow, U, time_c; Input.c:371-372, 424, 446-447)."""
    _fields_ = [
        ("inflow", C.c_int32), ("outflow", C.c_int32), ("transient_inflow", C.c_int32), ("reserved", C.c_int32),
        ("u_conv", C.c_double), ("time_c", C.c_double),
    ]


PLANE_INFLOW, PLANE_OUTFLOW = 0, 1


def open_bc(inflow=False, outflow=False, transient_inflow=False, u_conv=1.0, time_c=0.035):
    """tb_open_bc with the reference's defaults (U = 1, time_c = 0.035)."""
    return OpenBC(1 if inflow else 0, 1 if outflow else 0, 1 if transient_inflow else 0, 0, u_conv, time_c)


def conc_fields(fields):
    """ctypes array of tb_conc_field from dicts with keys type, pe,
    v_settle, alpha (default 1) and phi_max (default 0)."""
    arr = (ConcField * len(fields))()
    for a, f in zip(arr, fields):
        a.type = int(f["type"])
        a.pe = float(f["pe"])
        a.v_settle = float(f.get("v_settle", 0.0))
        a.alpha = float(f.get("alpha", 1.0))
        a.phi_max = float(f.get("phi_max", 0.0))
    return arr


def multigrain_properties(alpha, phi0, v_s0):
    """The host-side parameter rescaling of several grain sizes
    (RECALCULATE_MULTIGRAIN_PROPERTIES, Input.c:453-481): with
    E_i = alpha_i / alpha_0 and E_t = sum E_i, alpha_i <- E_i / E_t (the
    weights of the total concentration, summing to 1), Phi_i <- E_i Phi_0,
    and every settling speed V_s0[i] changes sign (negative = downward).
    Returns (alpha, phi, v_settle) lists."""
    a0 = float(alpha[0])
    et = 0.0
    for a in alpha:
        et += a / a0
    e = [a / a0 for a in alpha]
    return [ei / et for ei in e], [ei * phi0 for ei in e], [-v for v in v_s0]


class Energies(C.Structure):
    """tb_energies: Conc_compute_potential_energies of one field."""
    _fields_ = [(n, C.c_double) for n in ("ep", "ep_passive", "stokes", "conversion", "ep_passive_rate", "ep_rate")]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("converged", C.c_int32),
        ("fallback", C.c_int32), ("vcycles", C.c_int32),
        ("res0", C.c_double), ("res", C.c_double), ("ms", C.c_double),
    ]


class ResumeInfo(C.Structure):
    """tb_resume_info: the fields of Resume_primary.dat (Resume.c:247-319)."""
    _fields_ = [
        ("time", C.c_double), ("dt", C.c_double), ("dt_old", C.c_double),
        ("iter", C.c_int32), ("output_iter", C.c_int32),
    ]


# tb_petsc_* / tb_*_resume_primary status codes
IO_ERRORS = {-1: "bad argument", -2: "cannot open the file", -3: "not a PETSc binary vector",
             -4: "vector length exceeds the buffer", -5: "short read or write"}


class StepStats(C.Structure):
    _fields_ = [
        ("helm", (SolveStats * 4) * 2),
        ("pois", SolveStats * 2),
        ("div_max", C.c_double),
        ("ms", C.c_double),
    ]


IB_DTYPE_FIELDS = [f[0] for f in IBNode._fields_]


def make_config(**kw):
    """Build a Config from keyword arguments (defaults follow the reference's
    hard-coded internals, Input.c:405-509, except where the workload says
    otherwise)."""
    cfg = Config()
    defaults = dict(
        nx=64, ny=16, nz=16, lx=8.0, ly=1.0, lz=1.0, re=1000.0, pe=1000.0,
        eno_order=2, wall_noslip=(1, 1, 1, 1, 1, 1), buoyancy=0.0,
        conc_type=SALINITY, v_settle=0.0, sedimentation=0,
        helm_rtol=1e-12, helm_maxit=200, pois_rtol=1e-10, pois_maxit=200,
        pois_mode=POIS_FP64_PCG, bottom_offset=1, mg_pre=2, mg_post=2,
    )
    defaults.update(kw)
    for k, v in defaults.items():
        if k == "wall_noslip":
            for a in range(6):
                cfg.wall_noslip[a] = int(v[a])
        else:
            setattr(cfg, k, v)
    return cfg
