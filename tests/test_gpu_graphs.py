"""The multigrid V-cycles replay from HIP graphs.  Every replay must give the
same bits as launching the kernels one by one: two contexts run the same
steps (ENO, Helmholtz, Poisson, projection: eager work between the replays),
one with graphs and one without, and every step's fields are compared."""
import numpy as np
import pytest

from turbins_amd import Solver, abi, workload as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [abi.POIS_FP64_PCG, abi.POIS_MIXED_IR, abi.POIS_LEARNED_IR])
def test_graph_replays_match_eager(mode):
    c = W.case("bump_small")
    cfg = W.config_for(c, pois_mode=mode)
    runs = []
    for graphs in (True, False):
        g = Solver(cfg)
        g.set_graphs(graphs)
        g.setup_geometry(W.topography(c))
        for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)]).items():
            g.upload(k, v)
        out = []
        for _ in range(4):
            st = g.step(c["dt"])
            assert all(p.converged for p in st.pois)
            out.append({n: g.download(n) for n in ("U", "V", "W", "C", "PHI")})
        g.close()
        runs.append(out)
    for s, (a, b) in enumerate(zip(*runs)):
        for n in a:
            assert np.array_equal(a[n], b[n]), f"step {s} field {n} differs between graph and eager V-cycles"
