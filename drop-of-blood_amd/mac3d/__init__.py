"""Python host of the MI355X MAC step (ctypes over include/mac3d.h).

Mirrors the TypeScript solver API of js/MacSolver.js: a solver object with
step(dt, n) and field buffers u, v, w, p, dye.  The only compute path is the
HIP library drop-of-blood_amd/lib/libmac3d.so; importing this package on a
tree where it is not built raises (there is no CPU fallback).
"""
from .lib import (  # noqa: F401
    Config, Blob, Stats, MacSolver, Mac3dError, load_library, LIB_PATH,
    FIELD, STAGE, MG, KC, PREC, F32, F64,
)
