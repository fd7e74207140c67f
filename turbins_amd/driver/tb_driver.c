/*
 * tb_driver.c -- host time-step driver over the C-ABI (include/turbins_gpu.h),
 * in plain C99: the role of the reference's unpublished driver (gvg.c,
 * README.md:26), which calls the per-step module functions.
 *
 *   restart set in  -> SSP-RK2 steps (fixed dt or CFL dt) -> output cadence
 *   (real pressure, Output.c:2198-2204; diagnostics; field files) -> restart
 *   set out
 *
 * Usage:
 *   tb_driver --grid NX NY NZ --box LX LY LZ --re RE [--sc SC] --steps N
 *             (--dt DT | --cfl ALPHA --dt-max DTMAX)
 *             [--in DIR] [--out DIR] [--topography FILE] [--no-offset]
 *             [--eno-order K] [--walls W0 W1 W2 W3 W4 W5] [--buoyancy B]
 *             [--particle VS] [--pois-mode M] [--pois-rtol R] [--helm-rtol R]
 *             [--output-every N] [--inflow] [--transient-inflow TC] [--outflow U]
 *
 * --in reads a restart set (Resume_*.bin + Resume_primary.dat, the
 * reference's Resume_read_data, Resume.c:92-319) into the fields; without it
 * every field starts at zero.  --topography reads the bottom height
 * ((nz+1)*(nx+1) doubles, x fastest) from a PETSc binary vector; --no-offset
 * says those heights already carry the reference's 2.001*dy offset
 * (grid->interface_position, Grid.c:913,1118).  --cfl sets each step's dt
 * from the CFL rule of tb_cfl_dt (alpha, capped at --dt-max, the reference's
 * default_dt, Input.c:421-423).  --inflow opens the x = 0 face: before each
 * step the inflow profiles at the step's start time (Inflow.c:13-260; with
 * --transient-inflow the factor 1 - exp(-(t + 0.01)/TC)); --outflow opens
 * x = Lx with the convective condition at speed U (Outflow.c:12-304, applied
 * by tb_step at the start of each stage).  Every --output-every steps: the real
 * pressure, the kinetic energy and suspended mass, and the field files
 * U, V, W, C, PHI as PETSc binary vectors in OUT/output_NNNNN/.  At the end
 * the restart set is written to --out.  One line per step on stdout:
 *   step <iter> time <t> dt <dt> div <max|div u|> pois <its,its> ms <ms> rc <rc>
 * Exit status: 0, 1 on a hard error (message on stderr), 2 when some solve
 * missed its tolerance.
 */
#define _POSIX_C_SOURCE 200809L
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "../../include/turbins_gpu.h"

static void usage(void) {
    fprintf(stderr,
            "usage: tb_driver --grid NX NY NZ --box LX LY LZ --re RE [--sc SC] --steps N\n"
            "                 (--dt DT | --cfl ALPHA --dt-max DTMAX) [--in DIR] [--out DIR]\n"
            "                 [--topography FILE] [--no-offset] [--eno-order K] [--walls W0..W5]\n"
            "                 [--buoyancy B] [--particle VS] [--pois-mode M] [--pois-rtol R]\n"
            "                 [--helm-rtol R] [--output-every N] [--inflow] [--transient-inflow TC]\n"
            "                 [--outflow U]\n");
}

static int make_dir(const char *d) {
    if (mkdir(d, 0755) == 0 || errno == EEXIST) return 0;
    fprintf(stderr, "tb_driver: cannot create %s\n", d);
    return -1;
}

static int need(int i, int argc, int n, const char *opt) {
    if (i + n >= argc) {
        fprintf(stderr, "tb_driver: %s needs %d value(s)\n", opt, n);
        return -1;
    }
    return 0;
}

/* Output_flow_properties' pressure side effect and field dumps
 * (Output.c:2062-2216): phi becomes the real pressure, which the next
 * solve starts from, as in the reference */
static int output(tb_ctx *ctx, const char *out, int output_iter) {
    tb_solve_stats st;
    if (tb_real_pressure(ctx, &st) < 0) return -1;
    double ke = 0.0, mass = 0.0;
    if (tb_diagnostics(ctx, &ke, &mass) < 0) return -1;
    printf("output %d kinetic_energy %.17g suspended_mass %.17g pressure_its %d\n", output_iter, ke, mass,
           st.iterations);
    if (!out) return 0;
    char dir[4096], path[4200];
    snprintf(dir, sizeof(dir), "%s/output_%05d", out, output_iter);
    if (make_dir(dir)) return -1;
    const int fields[5] = {TB_F_U, TB_F_V, TB_F_W, TB_F_C, TB_F_PHI};
    const char *names[5] = {"U", "V", "W", "C", "PHI"};
    for (int a = 0; a < 5; a++) {
        snprintf(path, sizeof(path), "%s/%s.bin", dir, names[a]);
        if (tb_save_field(ctx, fields[a], path) < 0) return -1;
    }
    return 0;
}

int main(int argc, char **argv) {
    tb_config cfg;
    memset(&cfg, 0, sizeof(cfg));
    /* the reference's defaults (Input.c:405-509) where it has one */
    cfg.eno_order = 2;
    for (int w = 0; w < 6; w++) cfg.wall_noslip[w] = 1;
    cfg.buoyancy = 1.0;
    cfg.conc_type = TB_SALINITY;
    cfg.helm_rtol = 1e-12;
    cfg.helm_maxit = 1000;
    cfg.pois_rtol = 1e-8;
    cfg.pois_maxit = 300;
    cfg.pois_mode = TB_POIS_FP64_PCG;
    cfg.bottom_offset = 1;
    cfg.mg_pre = 2;
    cfg.mg_post = 2;
    double sc = 1.0, dt_fixed = 0.0, alpha = 0.0, dt_max = 0.0;
    int steps = -1, output_every = 0, have_grid = 0, have_box = 0, have_re = 0;
    const char *in = NULL, *out = NULL, *topo = NULL;
    tb_open_bc obc;
    memset(&obc, 0, sizeof(obc));
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "--grid")) {
            if (need(i, argc, 3, a)) return 1;
            cfg.nx = atoi(argv[i + 1]); cfg.ny = atoi(argv[i + 2]); cfg.nz = atoi(argv[i + 3]);
            i += 3;
            have_grid = 1;
        } else if (!strcmp(a, "--box")) {
            if (need(i, argc, 3, a)) return 1;
            cfg.lx = atof(argv[i + 1]); cfg.ly = atof(argv[i + 2]); cfg.lz = atof(argv[i + 3]);
            i += 3;
            have_box = 1;
        } else if (!strcmp(a, "--walls")) {
            if (need(i, argc, 6, a)) return 1;
            for (int w = 0; w < 6; w++) cfg.wall_noslip[w] = atoi(argv[i + 1 + w]);
            i += 6;
        } else if (!strcmp(a, "--inflow")) {
            obc.inflow = 1;
        } else if (!strcmp(a, "--no-offset")) {
            cfg.bottom_offset = 0;
        } else {
            if (need(i, argc, 1, a)) return 1;
            const char *v = argv[++i];
            if (!strcmp(a, "--re")) { cfg.re = atof(v); have_re = 1; }
            else if (!strcmp(a, "--sc")) sc = atof(v);
            else if (!strcmp(a, "--steps")) steps = atoi(v);
            else if (!strcmp(a, "--dt")) dt_fixed = atof(v);
            else if (!strcmp(a, "--cfl")) alpha = atof(v);
            else if (!strcmp(a, "--dt-max")) dt_max = atof(v);
            else if (!strcmp(a, "--in")) in = v;
            else if (!strcmp(a, "--out")) out = v;
            else if (!strcmp(a, "--topography")) topo = v;
            else if (!strcmp(a, "--eno-order")) cfg.eno_order = atoi(v);
            else if (!strcmp(a, "--buoyancy")) cfg.buoyancy = atof(v);
            else if (!strcmp(a, "--particle")) {
                /* settling speed as used (Input.c:481 flips the sign: < 0 downward) */
                cfg.conc_type = TB_PARTICLE;
                cfg.v_settle = -atof(v);
                cfg.sedimentation = 1;
            }
            else if (!strcmp(a, "--pois-mode")) cfg.pois_mode = atoi(v);
            else if (!strcmp(a, "--pois-rtol")) cfg.pois_rtol = atof(v);
            else if (!strcmp(a, "--helm-rtol")) cfg.helm_rtol = atof(v);
            else if (!strcmp(a, "--output-every")) output_every = atoi(v);
            else if (!strcmp(a, "--outflow")) { obc.outflow = 1; obc.u_conv = atof(v); }
            else if (!strcmp(a, "--transient-inflow")) { obc.transient_inflow = 1; obc.time_c = atof(v); }
            else {
                fprintf(stderr, "tb_driver: unknown option %s\n", a);
                usage();
                return 1;
            }
        }
    }
    if (!have_grid || !have_box || !have_re || steps < 0 || (dt_fixed > 0.0) == (alpha > 0.0) ||
        (alpha > 0.0 && !(dt_max > 0.0))) {
        usage();
        return 1;
    }
    cfg.pe = cfg.re * sc;
    tb_ctx *ctx = tb_create(&cfg);
    if (!ctx) {
        fprintf(stderr, "tb_driver: tb_create failed (bad configuration or no usable GPU)\n");
        return 1;
    }
    int status = 1;
    double *h = NULL;
    if (topo) {
        const int64_t n = (int64_t)(cfg.nx + 1) * (cfg.nz + 1);
        h = (double *)malloc(sizeof(double) * (size_t)n);
        if (!h || tb_petsc_read_array(topo, h, n) != n) {
            fprintf(stderr, "tb_driver: cannot read %lld heights from %s\n", (long long)n, topo);
            goto done;
        }
    }
    /* Grid_identify_geometry + the SDF chain + Immersed_setup_q_immersed_nodes */
    if (tb_setup_geometry(ctx, h) < 0) {
        fprintf(stderr, "tb_driver: tb_setup_geometry: %s\n", tb_last_error(ctx));
        goto done;
    }
    tb_resume_info info;
    memset(&info, 0, sizeof(info));
    if (in) {
        const int rc = tb_read_resume(ctx, in, &info);
        if (rc < 0) {
            fprintf(stderr, "tb_driver: tb_read_resume: %s\n", tb_last_error(ctx));
            goto done;
        }
        if (rc > 0) fprintf(stderr, "tb_driver: %s: dt or dt_old read as 0 (six-decimal record)\n", in);
    }
    if (out && make_dir(out)) goto done;
    if ((obc.inflow || obc.outflow) && tb_set_open_boundaries(ctx, &obc) < 0) {
        fprintf(stderr, "tb_driver: tb_set_open_boundaries: %s\n", tb_last_error(ctx));
        goto done;
    }
    int missed = 0;
    for (int s = 0; s < steps; s++) {
        double dt = dt_fixed;
        if (alpha > 0.0 && tb_cfl_dt(ctx, alpha, dt_max, &dt, NULL) < 0) {
            fprintf(stderr, "tb_driver: tb_cfl_dt: %s\n", tb_last_error(ctx));
            goto done;
        }
        if (obc.inflow && tb_inflow_profile(ctx, info.time, NULL) < 0) {
            fprintf(stderr, "tb_driver: tb_inflow_profile: %s\n", tb_last_error(ctx));
            goto done;
        }
        tb_step_stats st;
        const int rc = tb_step(ctx, dt, &st);
        if (rc < 0) {
            fprintf(stderr, "tb_driver: tb_step: %s\n", tb_last_error(ctx));
            goto done;
        }
        if (rc > 0) missed = 1;   /* the reference warns and goes on (Solver.c:326-348) */
        info.time += dt;
        info.dt_old = info.dt;
        info.dt = dt;
        info.iter++;
        printf("step %d time %.17g dt %.17g div %.6e pois %d,%d ms %.3f rc %d\n", info.iter, info.time, dt,
               st.div_max, st.pois[0].iterations, st.pois[1].iterations, st.ms, rc);
        fflush(stdout);
        if (output_every > 0 && info.iter % output_every == 0) {
            info.output_iter++;
            if (output(ctx, out, info.output_iter) < 0) {
                fprintf(stderr, "tb_driver: output: %s\n", tb_last_error(ctx));
                goto done;
            }
        }
    }
    if (out && tb_write_resume(ctx, out, &info) < 0) {
        fprintf(stderr, "tb_driver: tb_write_resume: %s\n", tb_last_error(ctx));
        goto done;
    }
    status = missed ? 2 : 0;
done:
    free(h);
    tb_destroy(ctx);
    return status;
}
