"""ctypes binding of include/mac3d.h and the MacSolver host class."""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG, "lib", "libmac3d.so")

F32, F64 = 0, 1
PREC = {"fp32": 0, "mixed": 1, "fp64": 2}
FIELD = dict(U=0, V=1, W=2, P=3, DYE=4, CU=5, CV=6, CW=7, RHS=8, MGX=9, MGB=10, MGR=11,
             DYE_PRED=12, WAVE=13, PTOP=14, CURL=15, CONC=16, CONC_PRED=17)
STAGE = dict(FACE_TO_CENTRE=0, ADVECT=1, CASSON=2, DIVERGENCE=3, VCYCLE=4, PROJECT=5,
             DYE_PREDICT=6, DYE_RESOLVE=7, WAVE=8, TOP_PRESSURE=9, PARCELS=10, CONC_PREDICT=11,
             CONC_RESOLVE=12)
MG = dict(SMOOTH=0, RESIDUAL=1, RESTRICT=2, PROLONG=3, COARSE=4, ZERO=5, SMOOTH_ZERO=6,
          SMOOTH2_MARCH=7)
KC = dict(FACE_TO_CENTRE=0, ADVECT=1, CASSON=2, DIVERGENCE=3, SMOOTH_L0=4, RESTRICT_L0=5,
          PROLONG_L0=6, MG_COARSE_LEVELS=7, COARSE_SOLVE=8, PROJECT=9, DYE=10, REFINE64=11,
          SMOOTH2_L0=12, WAVE=13, PARCELS=14)
STATUS = {0: "OK", 1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ESTATE", 5: "ENONFINITE"}


class Config(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nx", "ny", "nz", "levels", "nu1", "nu2", "cycles",
                                          "precision", "ir_max_iters", "device",
                                          "extrapolate")] + \
               [(n, C.c_double) for n in ("dx", "rho_w", "mu_w", "rho_b", "tau_y", "mu_inf", "g",
                                           "gamma_reg", "shear_eps2", "alpha_max", "vel_damp",
                                           "maccormack", "ir_tol", "sor_omega")]


class Blob(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cx", "cy", "cz", "radius", "w0")]


class WaveParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("scene", C.c_int32)] + \
               [(n, C.c_double) for n in ("half_extent", "half_xz", "depth", "surface_y",
                                           "marangoni", "vort_conf", "sim_time", "schlieren",
                                           "hemolysis")] + \
               [("physics", C.c_int32), ("reserved", C.c_int32), ("casson_scale", C.c_double),
                ("buoyancy", C.c_double)]


class Drop(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "radius", "vel_y", "osc", "drain",
                                           "seed")]


MAX_DROPS = 8


class Stats(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("steps", "vcycles", "ir_iters", "converged")] + \
               [(n, C.c_double) for n in ("max_div", "l2_div", "max_div_rate", "rel_res", "ms")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


MAX_BLOBS = 16


class Grad(C.Structure):
    _fields_ = [("loss", C.c_double), ("d_rho_b", C.c_double), ("nblobs", C.c_int32),
                ("steps", C.c_int32), ("d_blob", (C.c_double * 5) * MAX_BLOBS)]

    def as_dict(self):
        return {"loss": self.loss, "d_rho_b": self.d_rho_b, "steps": self.steps,
                "d_blob": [[self.d_blob[q][p] for p in range(5)] for q in range(self.nblobs)]}


RELEASE_PARAMS = ("x", "z", "radius", "start_y", "vel_y")
BLOB_PARAMS = ("cx", "cy", "cz", "radius", "w0")


class Release(C.Structure):
    """MultiDropSystem.spawnAerialDrop's parameters (dropKinematics.ts:354-391)"""
    _fields_ = [(n, C.c_double) for n in RELEASE_PARAMS]


class ReleaseMap(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("origin_x", "origin_y", "origin_z", "cell", "surface_y",
                                           "gravity", "vel_scale", "submerge", "max_offset",
                                           "r_min", "r_max")]


class ReleaseGrad(C.Structure):
    _fields_ = [("loss", C.c_double), ("d_rho_b", C.c_double), ("n", C.c_int32),
                ("steps", C.c_int32), ("d_release", (C.c_double * 5) * MAX_BLOBS),
                ("d_blob", (C.c_double * 5) * MAX_BLOBS)]

    def as_dict(self):
        return {"loss": self.loss, "d_rho_b": self.d_rho_b, "steps": self.steps,
                "d_release": [[self.d_release[q][p] for p in range(5)] for q in range(self.n)],
                "d_blob": [[self.d_blob[q][p] for p in range(5)] for q in range(self.n)]}


class FitOpts(C.Structure):
    _fields_ = [("max_iters", C.c_int32), ("checkpoint_every", C.c_int32), ("mask", C.c_uint64),
                ("tol", C.c_double), ("gtol", C.c_double)]


class FitReport(C.Structure):
    _fields_ = [("iters", C.c_int32), ("evals", C.c_int32), ("converged", C.c_int32),
                ("reserved", C.c_int32), ("loss0", C.c_double), ("loss", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved"}


def _releases(rel: Sequence):
    """Release structs from objects or dicts with the five fields"""
    def get(r, n):
        return r[n] if isinstance(r, dict) else getattr(r, n)
    return (Release * len(rel))(*[Release(*[float(get(r, n)) for n in RELEASE_PARAMS])
                                  for r in rel])


def release_default_map(cfg: "Config") -> ReleaseMap:
    """the reference tank's width on nx cells, top face at the surface
    (mac3d_release_default_map)"""
    L = load_library()
    m = ReleaseMap()
    rc = L.mac3d_release_default_map(C.byref(cfg), C.byref(m))
    if rc:
        raise Mac3dError(rc, L.mac3d_last_error(None).decode())
    return m


def release_blobs(rmap: ReleaseMap, rel: Sequence):
    """(blobs, jacobians): the blobs of the releases and, per release, the
    5x5 d(cx, cy, cz, radius, w0)/d(x, z, radius, start_y, vel_y)
    (mac3d_release_blobs; host arithmetic, no device)"""
    L = load_library()
    arr = _releases(rel)
    n = len(rel)
    blobs = (Blob * n)()
    jac = np.zeros((n, 5, 5))
    rc = L.mac3d_release_blobs(C.byref(rmap), arr, n, blobs, jac.ctypes.data)
    if rc:
        raise Mac3dError(rc, L.mac3d_last_error(None).decode())
    return [blobs[q] for q in range(n)], jac


class Mac3dError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mac3d {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None
EXPORTS = ("mac3d_default_config", "mac3d_create", "mac3d_destroy", "mac3d_upload",
           "mac3d_download", "mac3d_init_blobs", "mac3d_step", "mac3d_stage", "mac3d_mg_op",
           "mac3d_div_norms", "mac3d_profile", "mac3d_profile_read", "mac3d_adjoint",
           "mac3d_last_error", "mac3d_version", "mac3d_wave_default_params",
           "mac3d_wave_enable", "mac3d_set_drops", "mac3d_set_uniforms", "mac3d_parcels_upload",
           "mac3d_parcels_download", "mac3d_splat_download", "mac3d_checkpoint_size",
           "mac3d_checkpoint_save", "mac3d_checkpoint_load", "mac3d_div_norms_noalloc",
           "mac3d_conc_enable", "mac3d_parcels_emit", "mac3d_release_default_map",
           "mac3d_release_blobs", "mac3d_adjoint_release", "mac3d_fit_default_opts",
           "mac3d_fit_release")


def load_library(path: Optional[str] = None):
    """Load libmac3d.so (raises if it is not built: no fallback path)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"{p} is not built; run __graft_entry__.build()")
    L = C.CDLL(p)
    vp = C.c_void_p
    L.mac3d_default_config.argtypes = [C.POINTER(Config)]
    L.mac3d_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.mac3d_destroy.argtypes = [vp]
    L.mac3d_upload.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_size_t]
    L.mac3d_download.argtypes = [vp, C.c_int, C.c_int, vp, C.c_int, C.c_size_t]
    L.mac3d_init_blobs.argtypes = [vp, C.POINTER(Blob), C.c_int]
    L.mac3d_step.argtypes = [vp, C.c_double, C.c_int, C.POINTER(Stats)]
    L.mac3d_stage.argtypes = [vp, C.c_int, C.c_double]
    L.mac3d_mg_op.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    L.mac3d_div_norms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mac3d_profile.argtypes = [vp, C.c_int]
    L.mac3d_profile_read.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
    L.mac3d_adjoint.argtypes = [vp, C.POINTER(Blob), C.c_int, C.c_double, C.c_int, C.c_int, vp,
                                C.c_int, C.POINTER(Grad)]
    L.mac3d_wave_default_params.argtypes = [C.POINTER(WaveParams)]
    L.mac3d_wave_enable.argtypes = [vp, C.POINTER(WaveParams)]
    L.mac3d_set_drops.argtypes = [vp, C.POINTER(Drop), C.c_int]
    L.mac3d_set_uniforms.argtypes = [vp, vp, C.c_size_t]
    L.mac3d_parcels_upload.argtypes = [vp, vp, C.c_int32]
    L.mac3d_parcels_download.argtypes = [vp, vp, C.c_int32]
    L.mac3d_splat_download.argtypes = [vp, C.c_int, vp, C.c_size_t]
    L.mac3d_checkpoint_size.argtypes = [vp, C.POINTER(C.c_size_t)]
    L.mac3d_checkpoint_save.argtypes = [vp, vp, C.c_size_t]
    L.mac3d_checkpoint_load.argtypes = [vp, vp, C.c_size_t]
    L.mac3d_conc_enable.argtypes = [vp, C.c_int]
    L.mac3d_parcels_emit.argtypes = [vp, C.c_int32, C.c_int32]
    L.mac3d_div_norms_noalloc.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
    L.mac3d_release_default_map.argtypes = [C.POINTER(Config), C.POINTER(ReleaseMap)]
    L.mac3d_release_blobs.argtypes = [C.POINTER(ReleaseMap), C.POINTER(Release), C.c_int,
                                      C.POINTER(Blob), vp]
    L.mac3d_adjoint_release.argtypes = [vp, C.POINTER(ReleaseMap), C.POINTER(Release), C.c_int,
                                        C.c_double, C.c_int, C.c_int, vp, C.c_int,
                                        C.POINTER(ReleaseGrad)]
    L.mac3d_fit_default_opts.argtypes = [C.POINTER(FitOpts)]
    L.mac3d_fit_release.argtypes = [vp, C.POINTER(ReleaseMap), C.POINTER(Release), C.c_int,
                                    C.c_double, C.c_int, vp, C.c_int, C.POINTER(FitOpts),
                                    C.POINTER(FitReport)]
    L.mac3d_last_error.argtypes = [vp]
    L.mac3d_last_error.restype = C.c_char_p
    L.mac3d_version.restype = C.c_char_p
    for n in EXPORTS:
        if n not in ("mac3d_last_error", "mac3d_version"):
            getattr(L, n).restype = C.c_int
    if path is None:
        _lib = L
    return L


def default_config(**kw) -> Config:
    L = load_library()
    c = Config()
    L.mac3d_default_config(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


class MacSolver:
    """One simulation on one GPU: step(dt, n) and field buffers.

    Mirrors the TypeScript MacSolver (js/MacSolver.js); arrays are numpy
    (nz, ny, nx) with x fastest, as the reference's textures are laid out.
    """

    n_parcels = 0

    def __init__(self, config: Optional[Config] = None, **kw):
        self.L = load_library()
        self.cfg = config if config is not None else default_config(**kw)
        h = C.c_void_p()
        rc = self.L.mac3d_create(C.byref(self.cfg), C.byref(h))
        if rc:
            raise Mac3dError(rc, self.L.mac3d_last_error(None).decode())
        self.h = h
        self.shape = (self.cfg.nz, self.cfg.ny, self.cfg.nx)
        self.conc_refine = 0  # the concentration grid per axis, 0 while it is off

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.L.mac3d_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise Mac3dError(rc, self.L.mac3d_last_error(self.h).decode())

    # -- fields
    def level_shape(self, level: int):
        nz, ny, nx = self.shape
        return (nz >> level, ny >> level, nx >> level)

    def field_shape(self, fid: int, level: int = 0):
        if fid == FIELD["WAVE"]:
            n = self.wave_n
            return (n, n, 4)
        if fid == FIELD["PTOP"]:
            return (self.cfg.nz, self.cfg.nx)
        if fid == FIELD["CURL"]:
            return self.level_shape(0) + (4,)
        if fid == FIELD["CONC"]:
            return self.conc_shape() + (4,)
        if fid == FIELD["CONC_PRED"]:
            return self.conc_shape()
        return self.level_shape(level)

    def conc_shape(self):
        """(nz, ny, nx) of the concentration grid (the splat grid while it is
        on; else the MAC grid)"""
        r = max(self.conc_refine, 1)
        return tuple(r * a for a in self.shape)

    def download(self, field, level: int = 0, dtype=np.float64) -> np.ndarray:
        fid = FIELD[field] if isinstance(field, str) else field
        out = np.empty(self.field_shape(fid, level), dtype=dtype)
        dt = F64 if out.dtype == np.float64 else F32
        self._check(self.L.mac3d_download(self.h, fid, level, out.ctypes.data, dt, out.size))
        return out

    def download_into(self, field, out: np.ndarray, level: int = 0):
        fid = FIELD[field] if isinstance(field, str) else field
        dt = F64 if out.dtype == np.float64 else F32
        self._check(self.L.mac3d_download(self.h, fid, level, out.ctypes.data, dt, out.size))
        return out

    def upload(self, field, arr, level: int = 0):
        fid = FIELD[field] if isinstance(field, str) else field
        a = np.ascontiguousarray(arr)
        if a.dtype not in (np.float32, np.float64):
            a = a.astype(np.float64)
        dt = F64 if a.dtype == np.float64 else F32
        self._check(self.L.mac3d_upload(self.h, fid, level, a.ctypes.data, dt, a.size))

    @property
    def u(self): return self.download("U")

    @property
    def v(self): return self.download("V")

    @property
    def w(self): return self.download("W")

    @property
    def p(self): return self.download("P")

    @property
    def dye(self): return self.download("DYE")

    # -- free-surface coupling
    wave_n = 0

    def wave_enable(self, enable: bool = True, **kw):
        """Enable the free-surface wave coupling (flat surface, no drops) with
        mac3d_wave_params fields as keywords (n, scene, half_extent, half_xz,
        depth, surface_y, marangoni, vort_conf, sim_time, schlieren), or
        disable it."""
        if not enable:
            self._check(self.L.mac3d_wave_enable(self.h, None))
            self.wave_n = 0
            return
        p = WaveParams()
        self._check(self.L.mac3d_wave_default_params(C.byref(p)))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        self._check(self.L.mac3d_wave_enable(self.h, C.byref(p)))
        self.wave_n = p.n

    def set_drops(self, drops: Sequence):
        """drops: up to 8 sequences (x, y, z, radius, vel_y, osc, drain, seed)"""
        arr = (Drop * max(1, len(drops)))(*[Drop(*map(float, d)) for d in drops])
        self._check(self.L.mac3d_set_drops(self.h, arr, len(drops)))

    def set_uniforms(self, block):
        """the reference's packed 128-float SimUniforms block (surface height,
        Marangoni, domain bounds and the 8 drop slots are read)"""
        b = np.ascontiguousarray(block, dtype=np.float32)
        self._check(self.L.mac3d_set_uniforms(self.h, b.ctypes.data, b.size))

    def parcels_upload(self, parcels):
        """Lagrangian parcels as an (n, 8) array: x, y, z, mass, stretch x, y,
        z, age ticks + lysis (float32 on the device)"""
        p = np.ascontiguousarray(parcels, dtype=np.float32).reshape(-1, 8)
        self._check(self.L.mac3d_parcels_upload(self.h, p.ctypes.data, p.shape[0]))
        self.n_parcels = p.shape[0]

    def parcels_emit(self, cursor: int, count: int):
        """the next parcel stage first emits `count` parcels into the ring from
        slot `cursor` (PARTICLE_EMIT)"""
        self._check(self.L.mac3d_parcels_emit(self.h, int(cursor), int(count)))

    def conc_enable(self, enable: bool = True, refine: int = 1):
        """the reference's 4-channel concentration (field CONC; pass B in the
        step: parcels, CONC_PREDICT, CONC_RESOLVE) on the MAC grid (refine 1)
        or a grid twice as fine per axis (refine 2, as the reference holds
        it); a change of grid starts from zero"""
        e = int(refine) if enable else 0
        self._check(self.L.mac3d_conc_enable(self.h, e))
        self.conc_refine = e

    def parcels_download(self, n=None) -> np.ndarray:
        n = self.n_parcels if n is None else n
        out = np.empty((n, 8), dtype=np.float32)
        self._check(self.L.mac3d_parcels_download(self.h, out.ctypes.data, n))
        return out

    def splat_download(self, which) -> np.ndarray:
        """the splat counters conc_shape() uint32: which 0 or "INTACT", 1 or
        "LYSED"."""
        w = {"INTACT": 0, "LYSED": 1}.get(which, which)
        out = np.empty(self.conc_shape(), dtype=np.uint32)
        self._check(self.L.mac3d_splat_download(self.h, int(w), out.ctypes.data, out.size))
        return out

    # -- stepping
    def init_blobs(self, blobs: Sequence):
        arr = (Blob * len(blobs))(*[Blob(b.cx, b.cy, b.cz, b.radius, b.w0) for b in blobs])
        self._check(self.L.mac3d_init_blobs(self.h, arr, len(blobs)))

    def step(self, dt: float, n: int = 1) -> dict:
        st = Stats()
        self._check(self.L.mac3d_step(self.h, dt, n, C.byref(st)))
        return st.as_dict()

    def step_nostats(self, dt: float, n: int = 1):
        self._check(self.L.mac3d_step(self.h, dt, n, None))

    def adjoint(self, blobs: Sequence, dt: float, steps: int, target, checkpoint_every: int = 10):
        """Gradient of 1/2 ||dye_N - target||^2 with respect to the blob
        parameters (cx, cy, cz, radius, w0 per blob) and rho_b, by the
        hand-written adjoint with checkpoints every `checkpoint_every` steps."""
        arr = (Blob * len(blobs))(*[Blob(b.cx, b.cy, b.cz, b.radius, b.w0) for b in blobs])
        t = np.ascontiguousarray(target, dtype=np.float64)
        if t.size != self.cfg.nx * self.cfg.ny * self.cfg.nz:
            raise ValueError("target must have one value per cell")
        g = Grad()
        self._check(self.L.mac3d_adjoint(self.h, arr, len(blobs), dt, steps, checkpoint_every,
                                         t.ctypes.data, F64, C.byref(g)))
        return g.as_dict()

    # -- drop releases
    def release_map(self) -> ReleaseMap:
        """mac3d_release_default_map for this solver's grid"""
        return release_default_map(self.cfg)

    def _target(self, target):
        t = np.ascontiguousarray(target, dtype=np.float64)
        if t.size != self.cfg.nx * self.cfg.ny * self.cfg.nz:
            raise ValueError("target must have one value per cell")
        return t

    def adjoint_release(self, rmap: ReleaseMap, releases: Sequence, dt: float, steps: int,
                        target, checkpoint_every: int = 10) -> dict:
        """Gradient of 1/2 ||dye_N - target||^2 with respect to the release
        parameters (x, z, radius, start_y, vel_y per release), from the
        blobs the release map makes of them (mac3d_adjoint_release)."""
        t = self._target(target)
        g = ReleaseGrad()
        self._check(self.L.mac3d_adjoint_release(self.h, C.byref(rmap), _releases(releases),
                                                 len(releases), dt, steps, checkpoint_every,
                                                 t.ctypes.data, F64, C.byref(g)))
        return g.as_dict()

    def fit_release(self, rmap: ReleaseMap, releases: Sequence, dt: float, steps: int, target,
                    max_iters: int = 20, checkpoint_every: int = 10, mask: int = 0,
                    tol: float = 1e-6, gtol: float = 1e-9):
        """(fitted releases as dicts, report): quasi-Newton fit of the
        releases to a target dye field from the given guess
        (mac3d_fit_release)."""
        t = self._target(target)
        arr = _releases(releases)
        o = FitOpts(max_iters, checkpoint_every, mask, tol, gtol)
        rep = FitReport()
        self._check(self.L.mac3d_fit_release(self.h, C.byref(rmap), arr, len(releases), dt, steps,
                                             t.ctypes.data, F64, C.byref(o), C.byref(rep)))
        out = [{n: getattr(arr[q], n) for n in RELEASE_PARAMS} for q in range(len(releases))]
        return out, rep.as_dict()

    # -- checkpoints
    def checkpoint_save(self) -> np.ndarray:
        """The evolving state (fields, pressure history, wave, parcels) as a
        byte array (mac3d_checkpoint_save)."""
        n = C.c_size_t()
        self._check(self.L.mac3d_checkpoint_size(self.h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        self._check(self.L.mac3d_checkpoint_save(self.h, buf.ctypes.data, buf.size))
        return buf

    def checkpoint_load(self, buf) -> None:
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        self._check(self.L.mac3d_checkpoint_load(self.h, b.ctypes.data, b.size))
        # the parcel count is the header's int32 after magic (8 bytes) and 8
        # int32 fields (ABI, nx, ny, nz, levels, precision, element size, wave n)
        self.n_parcels = int(np.frombuffer(b[40:44].tobytes(), dtype="<i4")[0])

    def stage(self, stage, dt: float = 0.0):
        sid = STAGE[stage] if isinstance(stage, str) else stage
        self._check(self.L.mac3d_stage(self.h, sid, dt))

    def mg(self, op, level: int, count: int = 1):
        oid = MG[op] if isinstance(op, str) else op
        self._check(self.L.mac3d_mg_op(self.h, oid, level, count))

    def div_norms(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.L.mac3d_div_norms(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def profile(self, enable: bool):
        self._check(self.L.mac3d_profile(self.h, int(enable)))

    def profile_read(self, kclass) -> tuple:
        kid = KC[kclass] if isinstance(kclass, str) else kclass
        ms, n = C.c_double(), C.c_int64()
        self._check(self.L.mac3d_profile_read(self.h, kid, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def config_for(workload, **kw) -> Config:
    """Config of a scenarios.Workload."""
    return default_config(nx=workload.nx, ny=workload.ny, nz=workload.nz, dx=workload.dx,
                          levels=workload.levels, nu1=workload.nu1, nu2=workload.nu2,
                          cycles=workload.cycles, precision=workload.precision,
                          ir_tol=workload.ir_tol, **kw)
