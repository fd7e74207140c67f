"""Restart and field files in the reference's formats (CPU: no context
needed).  The byte layout is pinned against PETSc's published binary-viewer
format, written out by hand below, and against the oracle's numpy
restatement; the library's reader must reject what PETSc would."""
import os

import numpy as np
import pytest

import turbins_amd as T
from oracle import petsc_bin as PB


def test_petsc_vec_bytes_pinned(tmp_path):
    # class id 1211214 = 0x00127B4E, N = 3, then 1.0, -2.5, 3.25 as
    # big-endian IEEE doubles
    want = bytes.fromhex("00127b4e" "00000003" "3ff0000000000000" "c004000000000000" "400a000000000000")
    assert PB.encode_vec([1.0, -2.5, 3.25]) == want
    p = tmp_path / "v.bin"
    T.petsc_write(p, np.array([1.0, -2.5, 3.25]))
    assert p.read_bytes() == want
    assert np.array_equal(T.petsc_read(p), [1.0, -2.5, 3.25])


def test_petsc_vec_golden_fixture(tmp_path):
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "petsc_vec_pinned.bin")
    a = T.petsc_read(golden)
    assert a.tobytes() == np.array([1.0, -2.5, 3.25, -0.0]).tobytes()
    p = tmp_path / "again.bin"
    T.petsc_write(p, a)
    assert p.read_bytes() == open(golden, "rb").read()


@pytest.mark.parametrize("n", [0, 1, 7, 4099, (1 << 20) + 3])
def test_petsc_vec_matches_oracle(tmp_path, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) * 10.0 ** rng.integers(-300, 300, n)
    if n > 4:
        a[:4] = [0.0, -0.0, np.inf, 5e-324]   # signed zero, infinity, subnormal
    p = tmp_path / "v.bin"
    T.petsc_write(p, a)
    assert p.read_bytes() == PB.encode_vec(a)
    b = T.petsc_read(p)
    assert b.tobytes() == a.tobytes()
    q = tmp_path / "w.bin"
    q.write_bytes(PB.encode_vec(a))
    assert T.petsc_read(q).tobytes() == a.tobytes()
    assert PB.decode_vec(q.read_bytes()).tobytes() == a.tobytes()


def test_petsc_vec_rejects(tmp_path):
    good = PB.encode_vec(np.arange(10.0))
    bad_id = tmp_path / "id.bin"
    bad_id.write_bytes(b"\x00\x12\x7b\x4f" + good[4:])
    with pytest.raises(OSError, match="not a PETSc"):
        T.petsc_read(bad_id)
    little = tmp_path / "le.bin"     # a little-endian header is not PETSc's
    little.write_bytes(np.array([PB.VEC_FILE_CLASSID, 10], dtype="<i4").tobytes() + good[8:])
    with pytest.raises(OSError, match="not a PETSc"):
        T.petsc_read(little)
    short = tmp_path / "short.bin"
    short.write_bytes(good[:-8])
    with pytest.raises(OSError, match="short"):
        T.petsc_read(short, capacity=10)
    full = tmp_path / "full.bin"
    full.write_bytes(good)
    with pytest.raises(OSError, match="exceeds"):
        T.petsc_read(full, capacity=9)
    with pytest.raises(OSError, match="open"):
        T.petsc_read(tmp_path / "missing.bin", capacity=1)


def test_resume_primary_record(tmp_path):
    T.write_resume_primary(tmp_path, 12.3456789, 1e-3, 2.5e-4, 4000, 17)
    text = (tmp_path / "Resume_primary.dat").read_text()
    assert text == PB.primary_text(12.3456789, 1e-3, 2.5e-4, 4000, 17)
    assert text.splitlines()[1] == "12.345679"     # the reference's %lf: six decimals
    got = T.read_resume_primary(tmp_path)
    assert got == {"time": 12.345679, "dt": 0.001, "dt_old": 0.00025, "iteration": 4000, "output_iter": 17,
                   "zero_step": False}
    (tmp_path / "Resume_primary.dat").write_text("time=\n1.0\ndt=\n0.1\n")
    with pytest.raises(OSError):
        T.read_resume_primary(tmp_path)
    os.remove(tmp_path / "Resume_primary.dat")
    with pytest.raises(OSError, match="open"):
        T.read_resume_primary(tmp_path)


def test_resume_primary_zero_step_warning(tmp_path):
    """Resume_primary.dat keeps the reference's %lf (six decimals,
    Resume.c:272): a dt of 1e-7 is written as 0.000000 and read back as 0,
    which the reader reports (zero_step) instead of passing silently."""
    T.write_resume_primary(tmp_path, 1.0, 1e-7, 1e-3, 3, 0)
    got = T.read_resume_primary(tmp_path)
    assert got["dt"] == 0.0 and got["dt_old"] == 1e-3 and got["zero_step"]
    T.write_resume_primary(tmp_path, 1.0, 2e-3, 1e-3, 3, 0)
    assert not T.read_resume_primary(tmp_path)["zero_step"]
