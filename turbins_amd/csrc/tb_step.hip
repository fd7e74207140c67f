// tb_step.hip -- the SSP-RK2 time step (the role of the reference's missing
// driver, gvg.c): per stage, explicit terms of every field from the
// stage-start state, then the Helmholtz solves (u, v, w and every
// concentration field), then the Poisson solve and the projection; after
// both stages the 0.5/0.5 RK average
// (TVD_RK2, definitions.h:148-149) and the divergence check.
#include <cstring>

#include "tb_common.h"

int tb_eno_rhs_async(tb_ctx *c, double dt, bool store_conv);

extern "C" int tb_step(tb_ctx *c, double dt, tb_step_stats *st) {
    if (!c || !(dt > 0.0)) { if (c) c->err = "tb_step: bad argument"; return -1; }
    tb_step_stats loc;
    memset(&loc, 0, sizeof(loc));
    hipEvent_t e0 = c->step_ev[0], e1 = c->step_ev[1];
    TB_CHECK(c, hipEventRecord(e0, c->s));
    int worst = 0;
    if (tb_store_old(c)) return -1;
    const int sel = c->conc_sel;
    for (int stage = 0; stage < 2; stage++) {
        // the convective outflow condition from the stage-start fields
        // (Outflow_impose_convective_boundary; its place in the unpublished
        // driver is reconstructed: each RK stage advances the boundary ODE
        // with an Euler step, and the RK average combines the two)
        if (c->obc.outflow && tb_outflow_async(c, dt)) return -1;
        // c_total of the stage-start fields for the v buoyancy (several fields)
        if (tb_total_concentration(c)) return -1;
        // ENO and the right-hand sides (tb_eno_conv + tb_set_rhs) in one
        // pass; the convective terms stay in registers
        if (tb_eno_rhs_async(c, dt, false)) return -1;
        for (int q = 0; q < 3; q++) {
            int rc = tb_helmholtz_solve(c, q, dt, &loc.helm[stage][q]);
            if (rc < 0) return -1;
            if (rc > worst) worst = rc;
        }
        // every concentration field; the stats report field 0
        for (int n = 0; n < c->nconc; n++) {
            if (c->nconc > 1 && tb_conc_use(c, n)) return -1;
            tb_solve_stats cst;
            int rc = tb_helmholtz_solve(c, TB_Q_C, dt, &cst);
            if (rc < 0) return -1;
            if (rc > worst) worst = rc;
            if (n == 0) loc.helm[stage][TB_Q_C] = cst;
        }
        if (c->nconc > 1 && tb_conc_use(c, sel)) return -1;
        if (tb_poisson_rhs(c, dt)) return -1;
        int rc = tb_poisson_solve(c, &loc.pois[stage]);
        if (rc < 0) return -1;
        if (rc > worst) worst = rc;
        if (tb_project(c, dt)) return -1;
    }
    if (tb_rk_average(c, 0.5, 0.5)) return -1;
    double dm = 0.0;
    if (tb_div_max(c, &dm, nullptr)) return -1;
    TB_CHECK(c, hipEventRecord(e1, c->s));
    TB_CHECK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    TB_CHECK(c, hipEventElapsedTime(&ms, e0, e1));
    loc.div_max = dm;
    loc.ms = ms;
    if (st) *st = loc;
    return worst;
}
