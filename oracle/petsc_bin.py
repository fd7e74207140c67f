"""PETSc binary Vec and Resume_primary.dat restatements -- TEST INFRASTRUCTURE ONLY.

The reference writes its restart vectors with PetscViewerBinaryOpen + VecView
and reads them with VecLoadIntoVector (Resume.c:96-139; field dumps
Output.c:1610-1611).  The format belongs to PETSc 3.1 (not vendored in the
reference, README.md:18-20); its published binary-viewer layout for a Vec is

    int32 VEC_FILE_CLASSID = 1211214, int32 N, N doubles   (all big-endian)

with a DA global vector in the DA's natural order (x fastest).  The restart
set and its file names are those of Resume_set_data (Resume.c:172-244); the
primary record is the fprintf / fscanf pair of Resume_write_primary_data /
Resume_read_primary_data (Resume.c:247-319).  Only tests/ may import this
module, as the checker of the library's tb_petsc_* / tb_*_resume functions.
"""
import numpy as np

VEC_FILE_CLASSID = 1211214

# Resume_set_data (Resume.c:186-238), one concentration field
RESUME_FILES = [("Resume_u.bin", "U"), ("Resume_v.bin", "V"), ("Resume_w.bin", "W"),
                ("Resume_dpdx.bin", "DPDX"), ("Resume_dpdy.bin", "DPDY"), ("Resume_dpdz.bin", "DPDZ"),
                ("Resume_conc0.bin", "C")]


def encode_vec(a):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    return np.array([VEC_FILE_CLASSID, a.size], dtype=">i4").tobytes() + a.astype(">f8").tobytes()


def decode_vec(raw):
    head = np.frombuffer(raw[:8], dtype=">i4")
    if head.size != 2 or int(head[0]) != VEC_FILE_CLASSID:
        raise ValueError("not a PETSc binary Vec")
    n = int(head[1])
    body = np.frombuffer(raw[8:8 + 8 * n], dtype=">f8")
    if body.size != n:
        raise ValueError("short vector")
    return body.astype(np.float64)


def primary_text(time, dt, dt_old, iteration, output_iter):
    # C's %lf prints six decimals, as Python's %f does
    return "time=\n%f\ndt=\n%f\ndt_old=\n%f\niteration=\n%d\noutput_iter=\n%d\n" % (
        time, dt, dt_old, iteration, output_iter)
