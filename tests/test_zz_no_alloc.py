"""tb_div_max_noalloc (the allocation-free divergence check) against
tb_div_max on the same fields: the per-cell |div u| bit for bit, the maximum
and its cell index equal, on two calls in a row -- the second on a smaller
box, so the workspace holds the first call's partials when it starts -- and
a NaN found with its cell.  The caller's device buffers are allocated by the
test through the HIP runtime the library links (plumbing only; the entry
point itself allocates nothing); the call goes through the C-ABI."""
import ctypes as C

import numpy as np
import pytest

from turbins_amd import abi, workload as W

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    from turbins_amd import lib
    lib()                                   # the library's runtime is loaded by now
    h = C.CDLL("libamdhip64.so.7")          # same soname: the already loaded instance
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class _Dev:
    """A device copy of a host array (and back)."""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data_as(C.c_void_p), self.a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _state(name):
    from turbins_amd import Solver
    c = W.case(name)
    cfg = W.config_for(c)
    g = Solver(cfg)
    g.setup_geometry(W.topography(c))
    flags = [g.flags(q) for q in range(4)]
    fields = W.initial_fields(c, flags)
    for k, v in fields.items():
        g.upload(k, v)
    return c, g, fields, flags[3]


def _noalloc(c, g, fields, flags_c):
    from turbins_amd import div_max_noalloc
    hip = _hip()
    bufs = {k: _Dev(hip, fields[k]) for k in ("U", "V", "W", "C")}
    bufs["flags"] = _Dev(hip, flags_c)
    for k in ("mxc", "myc", "mzc"):
        bufs[k] = _Dev(hip, g.grid(k))
    bufs["div"] = _Dev(hip, np.full(g.NT, -7.0))
    bufs["max"] = _Dev(hip, np.full(1, -7.0))
    bufs["idx"] = _Dev(hip, np.full(1, -7, dtype=np.int64))
    box = abi.Box(c["nx"], c["ny"], c["nz"], c["nx"] + 1, 1.0 / c["nx"], 1.0 / c["ny"], 1.0 / c["nz"])
    p = {k: b.p.value for k, b in bufs.items()}
    div_max_noalloc(box, p["mxc"], p["myc"], p["mzc"], p["U"], p["V"], p["W"], p["C"], p["flags"], p["div"],
                    p["max"], p["idx"], None)
    assert hip.hipDeviceSynchronize() == 0
    out = bufs["div"].get(), float(bufs["max"].get()[0]), int(bufs["idx"].get()[0])
    for b in bufs.values():
        b.free()
    return out


def _ijk(c, idx):
    NX, NY = c["nx"] + 1, c["ny"] + 1
    return (idx % NX, (idx // NX) % NY, idx // (NX * NY))


def test_div_max_noalloc_matches_tb_div_max():
    for name in ("bump_small", "lock64"):      # the second box is the smaller one
        c, g, fields, flags_c = _state(name)
        dm, ijk = g.div_max()
        div_ref = g.download("DIV")
        div, mx, idx = _noalloc(c, g, fields, flags_c)
        print(name, "div max", dm, mx, "cell", ijk, _ijk(c, idx), flush=True)
        assert dm > 0.0
        assert np.array_equal(div, div_ref)
        assert mx == dm and _ijk(c, idx) == ijk
        g.close()


def test_div_max_noalloc_nan_and_bad_arguments():
    from turbins_amd import TBError, div_max_noalloc
    c, g, fields, flags_c = _state("bump_small")
    NX, NY = c["nx"] + 1, c["ny"] + 1
    u = fields["U"].copy()
    u[(7 * NY + 25) * NX + 20] = np.nan
    fields = dict(fields, U=u)
    div, mx, idx = _noalloc(c, g, fields, flags_c)
    assert np.isinf(mx) and _ijk(c, idx) == (20, 25, 7)
    # checked on the host before anything is launched
    big = abi.Box(2048, 2048, 1024, 2049, 1.0 / 2048, 1.0 / 2048, 1.0 / 1024)
    with pytest.raises(TBError, match="2\\^31"):
        div_max_noalloc(big, 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, None)
    with pytest.raises(TBError, match="bad argument"):
        div_max_noalloc(abi.Box(8, 8, 8, 4, 0.125, 0.125, 0.125), 1, 1, 1, 1, 1, 1, 1, 1, 0, 1, 1, None)
    g.close()
