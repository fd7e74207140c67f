"""Writes petsc_vec_pinned.bin: a PETSc binary Vec spelled out byte by byte
from PETSc's published binary-viewer layout (the format of the reference's
Resume_*.bin files, Resume.c:102-103): big-endian int32 class id 1211214,
int32 N = 4, then 1.0, -2.5, 3.25 and -0.0 as big-endian IEEE doubles."""
import os

RAW = bytes.fromhex(
    "00127b4e"            # VEC_FILE_CLASSID
    "00000004"            # N
    "3ff0000000000000"    # 1.0
    "c004000000000000"    # -2.5
    "400a000000000000"    # 3.25
    "8000000000000000"    # -0.0
)

if __name__ == "__main__":
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "petsc_vec_pinned.bin"), "wb") as f:
        f.write(RAW)
