// tb_net.hip -- learned initial guess for the Poisson solve (filled in by
// the network implementation; without a loaded network the solver uses the
// previous solution as its guess).
#include "tb_common.h"

struct NetState {
    int dummy;
};

int tb_net_predict(tb_ctx *c, double *phi_out) {
    (void)phi_out;
    if (!c || !c->net) return -1;
    return -1;
}

void tb_net_free(tb_ctx *c) {
    if (!c) return;
    delete c->net;
    c->net = nullptr;
}
