"""Fields and restart sets through the reference's file formats on the GPU:
every field survives a save/load bit for bit (the file is exactly the
oracle's encoding of the downloaded field), and a run resumed from the
restart set continues as the uninterrupted run does."""
import numpy as np
import pytest

from oracle import petsc_bin as PB
from turbins_amd import Solver, TBError, abi, workload as W

pytestmark = pytest.mark.gpu


def _solver(c, **kw):
    g = Solver(W.config_for(c, **kw))
    g.setup_geometry(W.topography(c))
    for k, v in W.initial_fields(c, [g.flags(q) for q in range(4)]).items():
        g.upload(k, v)
    return g


def test_field_files_round_trip(tmp_path):
    c = W.case("bump_small")
    g = _solver(c)
    g.step(c["dt"])
    for name in ("U", "V", "W", "C", "DPDX", "DPDY", "DPDZ", "PHI"):
        a = g.download(name)
        p = tmp_path / f"{name}.bin"
        g.save_field(name, p)
        assert p.read_bytes() == PB.encode_vec(a.ravel())
        g.upload("SCRATCH0", np.zeros_like(a))
        g.load_field("SCRATCH0", p)
        assert g.download("SCRATCH0").tobytes() == a.tobytes()
    # a vector of another length is refused, the field left as it was
    short = tmp_path / "short.bin"
    short.write_bytes(PB.encode_vec(np.arange(10.0)))
    before = g.download("U")
    with pytest.raises(TBError, match="length"):
        g.load_field("U", short)
    assert g.download("U").tobytes() == before.tobytes()
    g.close()


def test_resume_continues_the_run(tmp_path):
    """Resume_write_data / Resume_read_data (Resume.c:92-139): three steps,
    write the restart set, two more steps; a fresh context that reads the
    set and takes the same two steps reaches the same fields.  The restart
    set carries no phi (neither does the reference's), so the first Poisson
    solve of the resumed run starts from another guess: the fields agree to
    the solve tolerance, not to the bit."""
    c = W.case("bump_small")
    dt = c["dt"]
    kw = dict(pois_rtol=1e-11, helm_rtol=1e-13, pois_mode=abi.POIS_FP64_PCG)
    a = _solver(c, **kw)
    for _ in range(3):
        a.step(dt)
    a.write_resume(tmp_path, time=3 * dt, dt=dt, dt_old=dt, iteration=3, output_iter=1)
    names = [n for _, n in PB.RESUME_FILES]
    for fname, n in PB.RESUME_FILES:
        assert (tmp_path / fname).read_bytes() == PB.encode_vec(a.download(n).ravel())
    for _ in range(2):
        a.step(dt)
    ref = {n: a.download(n) for n in names}
    a.close()

    b = Solver(W.config_for(c, **kw))
    b.setup_geometry(W.topography(c))
    info = b.read_resume(tmp_path)
    assert info == {"time": 0.003, "dt": 0.001, "dt_old": 0.001, "iteration": 3, "output_iter": 1}
    for _ in range(2):
        b.step(dt)
    for n in names:
        x, y = b.download(n), ref[n]
        err = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300)
        assert err <= 1e-8, (n, err)
    b.close()
