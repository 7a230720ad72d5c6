"""mac3d_div_norms_noalloc: the divergence norms of caller-owned device
arrays, with no context and no allocation (partial sums in the library's
static device scratch), against mac3d_div_norms of a context holding the
same velocity.

Two calls in a row share the static scratch, the second on a smaller grid:
each must match its context's norms (max bit for bit: the same partials and
fold order; RMS within 2 ulp: the same sum, square root on the device against
the host).  Device memory for the inputs and output comes from the same HIP
runtime the library uses, through its C API."""
import ctypes as C
import os

import numpy as np
import pytest

from mac3d import MacSolver, scenarios
from mac3d import lib as L
from mac3d.lib import config_for

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    for p in ("/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        if os.path.exists(p):
            h = C.CDLL(p)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            return h
    pytest.skip("HIP runtime library not found")


class DevArray:
    def __init__(self, rt, host):
        self.rt, self.n = rt, host.nbytes
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), self.n) == 0
        assert rt.hipMemcpy(self.p, host.ctypes.data, self.n, H2D) == 0

    def read(self, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, self.p, self.n, D2H) == 0
        return out

    def free(self):
        self.rt.hipFree(self.p)


def noalloc(rt, s, dtype):
    lib = L.load_library()
    f = [np.ascontiguousarray(s.download(n, dtype=dtype)) for n in ("U", "V", "W")]
    dev = [DevArray(rt, a) for a in f]
    out = DevArray(rt, np.full(2, -1.0))
    nz, ny, nx = s.shape
    rc = lib.mac3d_div_norms_noalloc(dev[0].p, dev[1].p, dev[2].p,
                                     L.F32 if dtype == np.float32 else L.F64, nx, ny, nz,
                                     out.p, None)
    assert rc == 0, lib.mac3d_last_error(None)
    assert rt.hipDeviceSynchronize() == 0
    got = out.read(np.zeros(2))
    for d in dev + [out]:
        d.free()
    return got


def check(rt, name, steps, dtype):
    w = scenarios.workload(name, 1)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.step_nostats(w.dt, steps)
        mx, l2 = s.div_norms()
        got = noalloc(rt, s, dtype)
    print(f"[noalloc] {name} {w.nx}x{w.ny}x{w.nz} {np.dtype(dtype).name}: {got} vs "
          f"({mx}, {l2})", flush=True)
    assert got[0] == mx
    assert abs(got[1] - l2) <= 2 * np.spacing(l2)
    assert mx > 0 and l2 > 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_matches_div_norms_twice_second_smaller(dtype):
    rt = hip()
    check(rt, "adjoint", 3, dtype)   # 64x96x64 (1536 partial sums)
    check(rt, "tiny", 2, dtype)      # 32x48x32, the same scratch again


def test_noalloc_rejects_bad_arguments():
    lib = L.load_library()
    rt = hip()
    a = DevArray(rt, np.zeros(8, dtype=np.float32))
    out = DevArray(rt, np.zeros(2))
    try:
        # 4 x 128 x 257 = 131584 partial sums > 131072: refused before any launch
        rc = lib.mac3d_div_norms_noalloc(a.p, a.p, a.p, L.F32, 256, 512, 257, out.p, None)
        assert rc == 1 and b"static scratch" in lib.mac3d_last_error(None)
        assert lib.mac3d_div_norms_noalloc(None, a.p, a.p, L.F32, 2, 2, 2, out.p, None) == 1
        assert lib.mac3d_div_norms_noalloc(a.p, a.p, a.p, 7, 2, 2, 2, out.p, None) == 1
        assert lib.mac3d_div_norms_noalloc(a.p, a.p, a.p, L.F32, 0, 2, 2, out.p, None) == 1
    finally:
        a.free()
        out.free()
