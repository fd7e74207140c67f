"""The levels between the finest multigrid level and the LDS-resident coarse
cycle are smoothed by one blocked launch per V-cycle side (each box runs its
sweep pairs in LDS on a recomputed halo, the post-smoother with the
prolongation folded in).  It must give the same bits as one launch per half
sweep: two contexts, one per smoother, run the same steps and every field,
every Poisson iteration count and residual agree exactly.  Grids with
several mid levels and every coarsening pattern they use: (0,1,1), (0,1,0)
and (1,1,1) at 256x128x128; ragged boxes at 192x48x40."""
import numpy as np
import pytest

from turbins_amd import Solver, abi, workload as W

pytestmark = pytest.mark.gpu


def _run(c, mode, blocked, steps):
    g = Solver(W.config_for(c, pois_mode=mode, pois_rtol=1e-8))
    g.set_mg_smoother(blocked)
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)]).items():
        g.upload(k, v)
    out, its = [], []
    for _ in range(steps):
        st = g.step(c["dt"])
        assert all(p.converged for p in st.pois)
        its.append([(p.iterations, p.vcycles, p.res) for p in st.pois])
        out.append({n: g.download(n) for n in ("U", "V", "W", "C", "PHI")})
    g.close()
    return out, its


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR])
@pytest.mark.parametrize("name,over", [("learn256", {}), ("bump_small", dict(nx=192, ny=48, nz=40))])
def test_blocked_smoother_matches_half_sweeps(name, over, mode):
    c = W.case(name, **over)
    print("blocked smoother", name, over, "mode", mode, flush=True)
    a, ia = _run(c, mode, True, 2)
    b, ib = _run(c, mode, False, 2)
    print("poisson (iterations, vcycles, res)", ia)
    assert ia == ib
    for s, (x, y) in enumerate(zip(a, b)):
        for n in x:
            assert np.array_equal(x[n], y[n]), f"step {s} field {n}: blocked and half-sweep smoothers differ"


def test_smoother_switch_rejects_bad_values():
    c = W.case("bump_small")
    g = Solver(W.config_for(c))
    assert g.L.tb_set_mg_smoother(g.h, 2) < 0 and g.L.tb_set_mg_smoother(g.h, -1) < 0
    assert g.L.tb_set_mg_smoother(g.h, 0) == 0 and g.L.tb_set_mg_smoother(g.h, 1) == 0
    g.close()
