"""The learned mode's Galerkin guess (matrix kept from solve to solve), the
finest-level sweep setting of the V-cycles, and the r . z partial sums that
the last finest-level sweep pair forms on grids whose finest level is smoothed
by fused pairs (>= 4 M cells).  Every solve stops on its true FP64 residual,
so each check compares solutions of the same system reached by different
paths: they agree to the solve tolerance, and the guess must cut the initial
residual and the iterations."""
import numpy as np
import pytest

from turbins_amd import Solver, TBError, abi, workload as W

pytestmark = pytest.mark.gpu


def _rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def _solver(c, **kw):
    g = Solver(W.config_for(c, **kw))
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)]).items():
        g.upload(k, v)
    return g


def test_learned_guess_same_solutions_fewer_iterations():
    c = W.case("bump_small")
    runs = {}
    for mode in (abi.POIS_FP64_PCG, abi.POIS_LEARNED_IR):
        g = _solver(c, pois_rtol=1e-10, helm_rtol=1e-13, pois_mode=mode)
        fields, res0, its = [], [], []
        for _ in range(8):
            st = g.step(c["dt"])
            for p in st.pois:
                assert p.converged and p.res <= 1e-10
                res0.append(p.res0)
                its.append(p.iterations)
            fields.append({n: g.download(n) for n in ("U", "V", "W", "DPDX", "DPDY", "DPDZ")})
        g.close()
        runs[mode] = (fields, res0, its)
    (fp, r0p, itp), (fl, r0l, itl) = runs[abi.POIS_FP64_PCG], runs[abi.POIS_LEARNED_IR]
    for s, (a, b) in enumerate(zip(fl, fp)):
        for n in a:
            assert _rel(a[n], b[n]) <= 1e-7, (s, n, _rel(a[n], b[n]))
    # from the third step on the guess spans several earlier solutions
    print("res0 plain", np.median(r0p[4:]), "learned", np.median(r0l[4:]), "its", np.mean(itp[4:]), np.mean(itl[4:]))
    assert np.median(r0l[4:]) < 0.1 * np.median(r0p[4:])
    assert np.mean(itl[4:]) < np.mean(itp[4:])


def test_guess_history_lengths_and_reset():
    c = W.case("bump_small")
    g = _solver(c, pois_rtol=1e-10, helm_rtol=1e-13, pois_mode=abi.POIS_LEARNED_IR)
    for h in (1, 3, 10):
        g.set_guess_history(h)
        for _ in range(2):
            st = g.step(c["dt"])
            assert all(p.converged and p.res <= 1e-10 for p in st.pois)
    for bad in (0, 11):
        with pytest.raises(TBError):
            g.set_guess_history(bad)
    # uploading phi clears the history: the next solve starts from phi itself
    g.upload("PHI", g.download("PHI"))
    st = g.step(c["dt"])
    assert all(p.converged for p in st.pois)
    g.close()


def test_finest_sweeps_setting():
    c = W.case("bump_small")
    out = []
    for sweeps in ((0, 0, 0, 0), (3, 3, 3, 3), (2, 2, 1, 1)):
        for mode in (abi.POIS_FP64_PCG, abi.POIS_MIXED_IR):
            g = _solver(c, pois_rtol=1e-11, helm_rtol=1e-13, pois_mode=mode)
            g.set_mg_finest_sweeps(*sweeps)
            for _ in range(2):
                st = g.step(c["dt"])
                assert all(p.converged and p.res <= 1e-11 for p in st.pois)
            out.append(g.download("U"))
            for bad in ((-1, 0, 0, 0), (0, 9, 0, 0)):
                with pytest.raises(TBError):
                    g.set_mg_finest_sweeps(*bad)
            g.close()
    for u in out[1:]:
        assert _rel(u, out[0]) <= 1e-9


def test_fused_rz_on_a_paired_finest_level():
    """256 x 128 x 128 cells: the finest level (4.2 M cells) is smoothed by
    fused pairs, whose last one forms r . z for both Krylov methods."""
    c = W.case("learn256")
    sols = []
    for mode in (abi.POIS_FP64_PCG, abi.POIS_MIXED_IR):
        g = _solver(c, pois_rtol=1e-10, helm_rtol=1e-12, pois_mode=mode)
        dt = c["dt"]
        g.eno_conv_rhs(dt)
        for q in range(4):
            assert g.helmholtz_solve(q, dt).converged
        g.poisson_rhs(dt)
        st = g.poisson_solve()
        print("mode", mode, "its", st.iterations, "res", st.res, "fallback", st.fallback)
        assert st.converged and st.res <= 1e-10 and st.fallback == 0
        assert st.iterations <= 25
        sols.append(g.download("PHI"))
        g.close()
    assert _rel(sols[1], sols[0]) <= 1e-7
