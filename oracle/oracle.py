"""ctypes front end of the FP64 oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py as the checker of the HIP path; the product never
imports it.  Field and stage ids are those of include/mac3d.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MAC3D_ORACLE_LIB: another build of the same oracle (the sanitizer build of
# tests/test_sanitizers.py); by default the one beside this file
LIB = os.environ.get("MAC3D_ORACLE_LIB") or os.path.join(HERE, "liborc.so")

(F_U, F_V, F_W, F_P, F_DYE, F_CU, F_CV, F_CW, F_RHS, F_MGX, F_MGB, F_MGR, F_DYE_PRED, F_WAVE,
 F_PTOP, F_CURL, F_CONC, F_CONC_PRED) = range(18)
(ST_F2C, ST_ADVECT, ST_CASSON, ST_DIV, ST_VCYCLE, ST_PROJECT, ST_DYE_PRED, ST_DYE_RES, ST_WAVE,
 ST_TOP_PRESSURE, ST_PARCELS, ST_CONC_PRED, ST_CONC_RES) = range(13)
MG_SMOOTH, MG_RESIDUAL, MG_RESTRICT, MG_PROLONG, MG_COARSE, MG_ZERO = range(6)


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("nx", "ny", "nz", "levels", "nu1", "nu2", "cycles",
                                        "solve_to_tol", "max_cycles")] + \
               [(n, C.c_double) for n in ("dx", "rho_w", "mu_w", "rho_b", "tau_y", "mu_inf", "g",
                                           "gamma_reg", "shear_eps2", "alpha_max", "vel_damp",
                                           "maccormack", "tol", "omega")]


class OrcBlob(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("cx", "cy", "cz", "radius", "w0")]


class OrcWaveParams(C.Structure):
    _fields_ = [("n", C.c_int), ("scene", C.c_int)] + [(n, C.c_double) for n in (
        "half_extent", "half_xz", "depth", "surface_y", "marangoni", "vort_conf", "sim_time",
        "schlieren", "hemolysis")] + [("physics", C.c_int), ("reserved", C.c_int),
                                      ("casson_scale", C.c_double), ("buoyancy", C.c_double)]


class OrcDrop(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "radius", "vel_y", "osc", "drain", "seed")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_field.restype = C.POINTER(C.c_double)
        L.orc_field.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_field_len.restype = C.c_long
        L.orc_field_len.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_init_blobs.argtypes = [C.c_void_p, C.POINTER(OrcBlob), C.c_int]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_double)]
        L.orc_stage.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_mg_op.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
        L.orc_div_norms.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_wave_enable.argtypes = [C.c_void_p, C.POINTER(OrcWaveParams)]
        L.orc_parcels_set.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int]
        L.orc_parcels.argtypes = [C.c_void_p]
        L.orc_parcels.restype = C.POINTER(C.c_double)
        L.orc_splat.argtypes = [C.c_void_p, C.c_int]
        L.orc_splat.restype = C.POINTER(C.c_uint32)
        L.orc_splat_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_set_drops.argtypes = [C.c_void_p, C.POINTER(OrcDrop), C.c_int]
        L.orc_parcels_emit.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.orc_conc_enable.argtypes = [C.c_void_p, C.c_int]
        L.orc_fft.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int]
        L.orc_set_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_sample_cell.restype = C.c_double
        L.orc_sample_cell.argtypes = [C.POINTER(C.c_double), C.c_int, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_double]
        L.orc_smoothstep.restype = C.c_double
        L.orc_smoothstep.argtypes = [C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


DEFAULTS = dict(rho_w=1000.0, mu_w=1e-3, rho_b=1060.0, tau_y=0.005, mu_inf=0.0035, g=9.81,
                gamma_reg=0.40, shear_eps2=1e-6, alpha_max=1.0 / 6.0, vel_damp=1.0,
                maccormack=0.45 * 0.995, tol=1e-8, omega=1.0)


class Oracle:
    """One FP64 simulation on the CPU."""

    n_parcels = 0

    def __init__(self, nx, ny, nz, dx, levels=3, nu1=2, nu2=2, cycles=8, solve_to_tol=False,
                 max_cycles=50, **kw):
        p = dict(DEFAULTS)
        p.update(kw)
        self.params = OrcParams(nx=nx, ny=ny, nz=nz, levels=levels, nu1=nu1, nu2=nu2,
                                cycles=cycles, solve_to_tol=int(solve_to_tol),
                                max_cycles=max_cycles, dx=dx, **p)
        self.shape = (nz, ny, nx)
        self.levels = levels
        self.h = lib().orc_create(C.byref(self.params))
        if not self.h:
            raise ValueError("oracle: bad parameters")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def level_shape(self, level):
        nz, ny, nx = self.shape
        return (nz >> level, ny >> level, nx >> level)

    def view(self, field, level=0):
        """numpy view (nz, ny, nx) of an oracle field; writes go through."""
        n = lib().orc_field_len(self.h, field, level)
        ptr = lib().orc_field(self.h, field, level)
        if not ptr:
            raise ValueError("oracle: field not available")
        if field in (F_MGX, F_MGB, F_MGR):
            shape = self.level_shape(level)
        elif field == F_WAVE:
            w = int(round((n // 4) ** 0.5))
            shape = (w, w, 4)
        elif field == F_PTOP:
            shape = (self.shape[0], self.shape[2])
        elif field == F_CURL:
            shape = tuple(self.shape) + (4,)
        elif field == F_CONC:
            shape = (4,) + self.splat_shape()  # planes: intact, lysed, oxygenation, bed
        elif field == F_CONC_PRED:
            shape = self.splat_shape()  # the lysed channel's predictor
        else:
            shape = self.shape
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(shape)

    def get(self, field, level=0):
        return self.view(field, level).copy()

    def set(self, field, arr, level=0):
        v = self.view(field, level)
        v[...] = np.asarray(arr, dtype=np.float64).reshape(v.shape)

    def init_blobs(self, blobs):
        arr = (OrcBlob * len(blobs))(*[OrcBlob(b.cx, b.cy, b.cz, b.radius, b.w0) for b in blobs])
        lib().orc_init_blobs(self.h, arr, len(blobs))

    def step(self, dt, nsteps=1):
        rr = C.c_double(0.0)
        cycles = lib().orc_step(self.h, dt, nsteps, C.byref(rr))
        return cycles, rr.value

    def stage(self, stage, dt=0.0):
        lib().orc_stage(self.h, stage, dt)

    def mg(self, op, level, count=1):
        lib().orc_mg_op(self.h, op, level, count)

    def div_norms(self):
        a, b = C.c_double(), C.c_double()
        lib().orc_div_norms(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def wave_enable(self, n=512, half_extent=1.60, half_xz=0.0, depth=0.0, surface_y=0.30,
                    marangoni=0.0, scene=0, vort_conf=1.02, sim_time=0.0, schlieren=0.032,
                    hemolysis=0.088, physics=0, casson_scale=1.0, buoyancy=0.48, off=False):
        """enable (flat surface, no drops) or, off=True, disable the wave coupling"""
        if off:
            lib().orc_wave_enable(self.h, None)
            return
        self.wave_params = OrcWaveParams(n=n, scene=int(scene), half_extent=half_extent,
                                         half_xz=half_xz, depth=depth, surface_y=surface_y,
                                         marangoni=marangoni, vort_conf=vort_conf,
                                         sim_time=sim_time, schlieren=schlieren,
                                         hemolysis=hemolysis, physics=int(physics),
                                         casson_scale=casson_scale, buoyancy=buoyancy)
        lib().orc_wave_enable(self.h, C.byref(self.wave_params))

    def parcels_set(self, parcels):
        """Lagrangian parcels, an (n, 8) array (x, y, z, mass, stretch x, y, z,
        age ticks + lysis)"""
        p = np.ascontiguousarray(parcels, dtype=np.float64).reshape(-1, 8)
        lib().orc_parcels_set(self.h, p.ctypes.data_as(C.POINTER(C.c_double)), p.shape[0])
        self.n_parcels = p.shape[0]

    def parcels(self):
        ptr = lib().orc_parcels(self.h)
        return np.ctypeslib.as_array(ptr, shape=(self.n_parcels * 8,)).reshape(-1, 8).copy()

    def splat_shape(self):
        """(nz, ny, nx) of the splat grid: the concentration grid while it
        is on, else the MAC grid"""
        out = (C.c_int * 3)()
        lib().orc_splat_dims(self.h, out)
        return (out[2], out[1], out[0])

    def splat_view(self, which):
        """numpy view (writes go through) of the intact (0) or lysed (1)
        splat counters, splat_shape() uint32"""
        ptr = lib().orc_splat(self.h, int(which))
        shape = self.splat_shape()
        return np.ctypeslib.as_array(ptr, shape=(int(np.prod(shape)),)).reshape(shape)

    def splat(self, which):
        """the intact (0) or lysed (1) splat counters, splat_shape() uint32"""
        return self.splat_view(which).copy()

    def conc_enable(self, on=True, refine=1):
        """the reference's 4-channel concentration (pass B in the step), on
        the MAC grid (refine 1) or a grid twice as fine per axis (refine 2,
        the reference's CONC_GRID over its VEL_GRID)"""
        lib().orc_conc_enable(self.h, int(refine) if on else 0)

    def parcels_emit(self, cursor, count):
        """the next parcel stage first emits `count` parcels from ring slot
        `cursor` (PARTICLE_EMIT)"""
        lib().orc_parcels_emit(self.h, int(cursor), int(count))

    def set_drops(self, drops):
        """drops: sequences (x, y, z, radius, vel_y, osc, drain, seed)"""
        arr = (OrcDrop * max(1, len(drops)))(*[OrcDrop(*map(float, d)) for d in drops])
        lib().orc_set_drops(self.h, arr, len(drops))


def fft(z, sign=-1):
    """the oracle's unitary radix-2 FFT of a complex vector (length a power of two)"""
    z = np.asarray(z, dtype=np.complex128)
    buf = np.ascontiguousarray(np.stack([z.real, z.imag], axis=-1).ravel())
    lib().orc_fft(buf.ctypes.data_as(C.POINTER(C.c_double)), z.size, 1, int(sign))
    return buf[0::2] + 1j * buf[1::2]


def sample_cell(f, ux, uy, uz):
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, ny, nx = f.shape
    return lib().orc_sample_cell(f.ctypes.data_as(C.POINTER(C.c_double)), nx, ny, nz, ux, uy, uz)


def smoothstep(e0, e1, x):
    return lib().orc_smoothstep(e0, e1, x)
