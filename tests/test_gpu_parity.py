"""HIP path vs the FP64 oracle, through the C ABI (needs an MI355X).

Tolerances: FP32 kernels against FP64 arithmetic, relative L2 <= 1e-5 per
field (the workload's parity bar); the FP64 device mode against the oracle
<= 1e-11.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from gpu_helpers import pair, copy_state

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-11


def _cmp(s, o, fields, tol, level=0):
    for f in fields:
        fid = getattr(oracle, "F_" + f)
        a = s.download(f, level)
        b = o.get(fid, level)
        e = rel_l2(a, b)
        assert e <= tol, f"{f} level {level}: rel L2 {e:.3e} > {tol:.1e}"


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, TOL64)])
@pytest.mark.parametrize("stage,fields", [
    ("FACE_TO_CENTRE", ("CU", "CV", "CW")),
    ("ADVECT", ("U", "V", "W")),
    ("CASSON", ("U", "V", "W")),
    ("DIVERGENCE", ("RHS",)),
    ("PROJECT", ("U", "V", "W")),
])
def test_stage(stage, fields, prec, tol):
    w, s, o = pair(precision=prec)
    with s:
        # centred velocity feeds advection; set it in both
        o.stage(oracle.ST_F2C)
        s.stage("FACE_TO_CENTRE")
        if stage in ("DIVERGENCE", "PROJECT"):
            # a projected field has D(u) ~ 0: advect and force it first so the
            # divergence and the projection act on an O(1) right-hand side
            o.stage(oracle.ST_ADVECT, w.dt)
            o.stage(oracle.ST_CASSON, w.dt)
            copy_state(o, s, ("U", "V", "W"))
        if stage == "PROJECT":
            o.stage(oracle.ST_DIV)
            o.stage(oracle.ST_VCYCLE)
            copy_state(o, s, ("P",))
        o.stage(getattr(oracle, {"FACE_TO_CENTRE": "ST_F2C", "ADVECT": "ST_ADVECT",
                                  "CASSON": "ST_CASSON", "DIVERGENCE": "ST_DIV",
                                  "PROJECT": "ST_PROJECT"}[stage]), w.dt)
        s.stage(stage, w.dt)
        _cmp(s, o, fields, tol)


@pytest.mark.parametrize("scale", [1.0, 6.0])
def test_advect_stage_large_steps(scale):
    """FP32 advection samples from each block's staged tile and gathers from
    memory where a sample's corners leave it (displacements over a cell):
    the advection stage at the 192x256x192 workload's step and at 6 times
    it (traces of up to ~2.5 cells), against the oracle."""
    w, s, o = pair(name="tiny", precision=0)
    with s:
        dt = w.dt * scale
        o.stage(oracle.ST_F2C)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_ADVECT, dt)
        s.stage("ADVECT", dt)
        _cmp(s, o, ("U", "V", "W"), TOL32)


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, TOL64)])
def test_dye_stages(prec, tol):
    w, s, o = pair(precision=prec)
    with s:
        o.stage(oracle.ST_F2C)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_DYE_PRED, w.dt)
        s.stage("DYE_PREDICT", w.dt)
        _cmp(s, o, ("DYE_PRED",), tol)
        o.stage(oracle.ST_DYE_RES, w.dt)
        s.stage("DYE_RESOLVE", w.dt)
        _cmp(s, o, ("DYE",), tol)


@pytest.mark.parametrize("scale", [1.0, 6.0])
def test_dye_stages_large_steps(scale):
    """The FP32 dye predictor samples from each block's staged tile and
    gathers from memory where a sample's corners leave it: the dye stages at
    the workload's step and at 6 times it, against the oracle."""
    w, s, o = pair(name="tiny", precision=0)
    with s:
        dt = w.dt * scale
        o.stage(oracle.ST_F2C)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_DYE_PRED, dt)
        s.stage("DYE_PREDICT", dt)
        _cmp(s, o, ("DYE_PRED",), TOL32)
        o.stage(oracle.ST_DYE_RES, dt)
        s.stage("DYE_RESOLVE", dt)
        _cmp(s, o, ("DYE",), TOL32)


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, TOL64)])
def test_dye_resolve_after_dye_change(prec, tol):
    """The resolve stage after the dye changed since the predict (an upload
    in between): the library drops the limiter bounds the predict formed and
    traces back again, as the oracle's resolve does on the current dye."""
    w, s, o = pair(precision=prec)
    with s:
        o.stage(oracle.ST_F2C)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_DYE_PRED, w.dt)
        s.stage("DYE_PREDICT", w.dt)
        dye = o.get(oracle.F_DYE)
        dye = dye * 0.5 + np.roll(dye, 3, axis=2) * 0.5
        o.set(oracle.F_DYE, dye)
        s.upload("DYE", dye)
        o.stage(oracle.ST_DYE_RES, w.dt)
        s.stage("DYE_RESOLVE", w.dt)
        _cmp(s, o, ("DYE",), tol)


@pytest.mark.parametrize("prec,tol", [(0, 2e-6), (2, TOL64)])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_mg_smooth_residual_restrict_prolong(level, prec, tol):
    w, s, o = pair(precision=prec, warm_steps=0)
    rng = np.random.default_rng(level)
    with s:
        shp = s.level_shape(level)
        x = rng.standard_normal(shp)
        b = rng.standard_normal(shp)
        for dev_f, orc_f, arr in (("MGX", oracle.F_MGX, x), ("MGB", oracle.F_MGB, b)):
            s.upload(dev_f, arr, level)
            o.set(orc_f, arr, level)
        s.mg("SMOOTH", level, 2)
        o.mg(oracle.MG_SMOOTH, level, 2)
        _cmp(s, o, ("MGX",), tol, level)
        s.mg("RESIDUAL", level)
        o.mg(oracle.MG_RESIDUAL, level)
        _cmp(s, o, ("MGR",), tol * 10, level)
        s.mg("RESTRICT", level)
        o.mg(oracle.MG_RESTRICT, level)
        _cmp(s, o, ("MGB",), tol * 10, level + 1)
        xc = rng.standard_normal(s.level_shape(level + 1))
        s.upload("MGX", xc, level + 1)
        o.set(oracle.F_MGX, xc, level + 1)
        s.mg("PROLONG", level)
        o.mg(oracle.MG_PROLONG, level)
        _cmp(s, o, ("MGX",), tol, level)


@pytest.mark.parametrize("prec,tol", [(0, 1e-5), (2, 1e-11)])
@pytest.mark.parametrize("name", ["tiny", "adjoint"])
def test_coarse_solve_exact(name, prec, tol):
    w, s, o = pair(name, precision=prec, warm_steps=0)
    L = w.levels
    rng = np.random.default_rng(7)
    with s:
        b = rng.standard_normal(s.level_shape(L))
        s.upload("MGB", b, L)
        o.set(oracle.F_MGB, b, L)
        s.mg("COARSE", L)
        o.mg(oracle.MG_COARSE, L)
        _cmp(s, o, ("MGX",), tol, L)


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, 1e-10)])
def test_vcycle(prec, tol):
    w, s, o = pair(precision=prec)
    with s:
        o.stage(oracle.ST_F2C); o.stage(oracle.ST_ADVECT, w.dt); o.stage(oracle.ST_CASSON, w.dt)
        o.stage(oracle.ST_DIV)
        copy_state(o, s, ("U", "V", "W", "P"))
        s.stage("DIVERGENCE")
        for _ in range(3):
            o.stage(oracle.ST_VCYCLE)
            s.stage("VCYCLE")
        _cmp(s, o, ("P",), tol)


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, 1e-10)])
@pytest.mark.parametrize("nu1,nu2,levels", [(1, 1, 3), (3, 2, 3), (2, 3, 2), (2, 2, 4)])
def test_vcycle_schedules(nu1, nu2, levels, prec, tol):
    """V-cycles with odd sweep counts (a two-sweep pass and a single sweep on
    one level) and with 2 and 4 levels, against the oracle's cycle with the
    same schedule (32x48x32)."""
    w, s, o = pair(precision=prec, nu1=nu1, nu2=nu2, levels=levels)
    with s:
        o.stage(oracle.ST_F2C); o.stage(oracle.ST_ADVECT, w.dt); o.stage(oracle.ST_CASSON, w.dt)
        o.stage(oracle.ST_DIV)
        copy_state(o, s, ("U", "V", "W", "P"))
        s.stage("DIVERGENCE")
        for _ in range(2):
            o.stage(oracle.ST_VCYCLE)
            s.stage("VCYCLE")
        _cmp(s, o, ("P",), tol)
    o.close()


def _full_size_rhs(precision):
    w, s, o = pair("ref", precision=precision, warm_steps=0)
    o.stage(oracle.ST_F2C); o.stage(oracle.ST_ADVECT, w.dt); o.stage(oracle.ST_CASSON, w.dt)
    o.stage(oracle.ST_DIV)
    copy_state(o, s, ("U", "V", "W", "P"))
    s.stage("DIVERGENCE")
    return w, s, o


def test_vcycle_full_size():
    """FP32 V-cycles at 192x256x192 against the oracle's: level 0 smooths in
    z-marching two-sweep passes, levels 1 (96x128x96) and 2 (48x64x48) in the
    tile pass, and the residual-restriction of levels 0 and 1 marches in z.

    The pressure is held to the FP32 bar after every one of the workload's
    8 cycles from a cold start, and the projected velocities after the last.
    (The level-0 residual is formed in FP64 arithmetic: with an FP32 one the
    iterate drifted by ~1.4e-4 of rms(p) in the smoothest mode over these
    cycles, the FP32 rounding of a residual that cancels to ~N^-2 of its
    terms; measured now 2.7e-7 to 4.7e-7.)"""
    w, s, o = _full_size_rhs(0)
    with s:
        for c in range(w.cycles):
            o.stage(oracle.ST_VCYCLE)
            s.stage("VCYCLE")
            _cmp(s, o, ("P",), TOL32)
        o.stage(oracle.ST_PROJECT)
        s.stage("PROJECT")
        _cmp(s, o, ("U", "V", "W"), TOL32)
    o.close()


def test_vcycle_full_size_fp64():
    """Two FP64 V-cycles at 192x256x192 (the z-marching residual-restriction
    of levels 0 and 1 and the single-sweep passes in FP64) against the
    oracle's, to 1e-10."""
    w, s, o = _full_size_rhs(2)
    with s:
        for _ in range(2):
            o.stage(oracle.ST_VCYCLE)
            s.stage("VCYCLE")
        _cmp(s, o, ("P",), 1e-10)
    o.close()


@pytest.mark.parametrize("prec,tol", [(0, TOL32), (2, 1e-10)])
def test_full_step_tiny(prec, tol):
    """The tiny workload: 32x48x32, 20 steps, dt 0.5 ms, from the seeded blob."""
    w, s, o = pair(precision=prec, warm_steps=0)
    with s:
        st = s.step(w.dt, w.steps)
        o.step(w.dt, w.steps)
        _cmp(s, o, ("U", "V", "W", "P", "DYE"), tol)
        assert st["max_div"] <= 1e-6, st
        omax, _ = o.div_norms()
        assert omax <= 1e-9


def test_mixed_tiny():
    """Mixed precision (FP16 level>=2 storage, FP64 refinement to 1e-8)
    against the FP64 oracle solving to the same tolerance."""
    w, s, o = pair(precision=1, warm_steps=0)
    with s:
        st = s.step(w.dt, w.steps)
        o.step(w.dt, w.steps)
        assert st["converged"] == 1 and st["rel_res"] <= 1e-8, st
        assert st["max_div"] <= 1e-6, st
        _cmp(s, o, ("U", "V", "W", "DYE"), TOL32)
        _cmp(s, o, ("P",), 1e-5)


def test_step_rejects_bad_dt():
    w, s, o = pair(warm_steps=0)
    with s:
        before = s.download("V")
        from mac3d import Mac3dError
        for bad in (float("nan"), -1e-3, 0.0, float("inf")):
            with pytest.raises(Mac3dError):
                s.step(bad, 1)
        assert np.array_equal(before, s.download("V"))


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replays_match_eager(prec):
    """Steps replayed from captured graphs (the default path) against the
    same steps launched eagerly (profiling switches graphs off), one step per
    call with downloads in between: every replay is bit-identical."""
    w, s, o = pair(precision=prec, warm_steps=0)
    from mac3d import MacSolver
    from mac3d.lib import config_for
    with s, MacSolver(config_for(w)) as e:
        e.init_blobs(w.blobs)
        e.profile(True)
        for step in range(6):
            s.step(w.dt, 1)
            e.step(w.dt, 1)
            for f in ("U", "V", "W", "P", "DYE"):
                a, b = s.download(f), e.download(f)
                assert np.array_equal(a, b), (step, f)


@pytest.mark.parametrize("level", [0, 1])
def test_over_relaxed_smoother_fp64(level):
    """The over-relaxed red-black sweep (omega 1.2; the mixed mode smooths
    with 1.25) against the oracle's, in the FP64 device mode."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=2)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, omega=1.2)
    rng = np.random.default_rng(11 + level)
    with MacSolver(config_for(w, sor_omega=1.2)) as s:
        shp = s.level_shape(level)
        x, b = rng.standard_normal(shp), rng.standard_normal(shp)
        s.upload("MGX", x, level)
        s.upload("MGB", b, level)
        o.set(oracle.F_MGX, x, level)
        o.set(oracle.F_MGB, b, level)
        s.mg("SMOOTH", level, 2)
        o.mg(oracle.MG_SMOOTH, level, 2)
        assert rel_l2(s.download("MGX", level), o.get(oracle.F_MGX, level)) <= 1e-11


@pytest.mark.parametrize("op", ["SMOOTH", "SMOOTH_ZERO"])
def test_two_sweep_pass_fp16_level_against_oracle(op):
    """Mixed mode, 192x256x192: level 2 (48x64x48, FP16 storage) smooths a
    pair of sweeps in one two-sweep pass, FP32 in between and rounded to FP16
    once on the store.  Against the oracle's two over-relaxed sweeps (omega
    1.25, the mode's default) from the same FP16 values: within 1e-3 relative
    (FP16 rounds to 2^-11)."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.scaled(scenarios.workload("ref", 1), precision=1)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, omega=1.25)
    rng = np.random.default_rng(5)
    level = 2
    with MacSolver(config_for(w, sor_omega=1.25)) as s:
        shp = s.level_shape(level)
        x = rng.standard_normal(shp).astype(np.float16).astype(np.float64)
        b = rng.standard_normal(shp).astype(np.float16).astype(np.float64)
        s.upload("MGX", x, level)
        s.upload("MGB", b, level)
        o.set(oracle.F_MGX, x, level)
        o.set(oracle.F_MGB, b, level)
        s.mg(op, level, 2)
        if op == "SMOOTH_ZERO":
            o.set(oracle.F_MGX, np.zeros(shp), level)
        o.mg(oracle.MG_SMOOTH, level, 2)
        r = rel_l2(s.download("MGX", level), o.get(oracle.F_MGX, level))
        assert r <= 1e-3, r
    o.close()


@pytest.mark.parametrize("dims,levels", [((32, 48, 32), 3), ((96, 40, 20), 2), ((64, 96, 64), 3)])
@pytest.mark.parametrize("omega", [1.0, 1.2])
def test_two_sweep_pass_equals_two_sweeps(dims, levels, omega):
    """Level 0, FP32: a smoothing call of two sweeps runs as one temporally
    blocked pass; it must equal two calls of one sweep bit for bit, from a
    given guess and from a zero guess (grids with partial 64 x 16 tiles and
    partial z-chunks included)."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.scaled(scenarios.workload("tiny", 1), nx=dims[0], ny=dims[1], nz=dims[2],
                         levels=levels, precision=0)
    rng = np.random.default_rng(int(omega * 10) + dims[2])
    with MacSolver(config_for(w, sor_omega=omega)) as s:
        shp = s.level_shape(0)
        x = rng.standard_normal(shp).astype(np.float32)
        b = rng.standard_normal(shp).astype(np.float32)
        res = {}
        for name, ops in [("pass", [("SMOOTH", 2)]), ("sweeps", [("SMOOTH", 1), ("SMOOTH", 1)]),
                          ("zpass", [("SMOOTH_ZERO", 2)]),
                          ("zsweeps", [("SMOOTH_ZERO", 1), ("SMOOTH", 1)])]:
            s.upload("MGX", x, 0)
            s.upload("MGB", b, 0)
            for op, n in ops:
                s.mg(op, 0, n)
            res[name] = s.download("MGX", 0, dtype=np.float32)
        assert np.array_equal(res["pass"], res["sweeps"])
        assert np.array_equal(res["zpass"], res["zsweeps"])
        assert not np.array_equal(res["pass"], res["zpass"])


def test_two_sweep_pass_against_oracle():
    """The FP32 two-sweep pass with over-relaxation (omega 1.2) against the
    oracle's two FP64 sweeps on the tiny grid."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=0)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, omega=1.2)
    rng = np.random.default_rng(5)
    with MacSolver(config_for(w, sor_omega=1.2)) as s:
        shp = s.level_shape(0)
        x, b = rng.standard_normal(shp), rng.standard_normal(shp)
        s.upload("MGX", x, 0)
        s.upload("MGB", b, 0)
        o.set(oracle.F_MGX, x, 0)
        o.set(oracle.F_MGB, b, 0)
        s.mg("SMOOTH", 0, 2)
        o.mg(oracle.MG_SMOOTH, 0, 2)
        assert rel_l2(s.download("MGX", 0), o.get(oracle.F_MGX, 0)) <= TOL32


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("omega", [1.0, 1.2])
def test_two_sweep_pass_coarse_levels_against_oracle(level, omega):
    """Levels 1 and 2 of the 192x256x192 grid (96x128x96 and 48x64x48, at or
    the tile pass's sizes) smooth pairs of sweeps in the tile pass: against the
    oracle's two FP64 sweeps (h^2 = 4^l, lid coefficient 2 2^l / (2^l + 1))."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.workload("ref", 1)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, omega=omega)
    rng = np.random.default_rng(17)
    with MacSolver(config_for(w, sor_omega=omega)) as s:
        shp = s.level_shape(level)
        assert np.prod(shp) >= 1 << 17
        x, b = rng.standard_normal(shp), rng.standard_normal(shp)
        s.upload("MGX", x, level)
        s.upload("MGB", b, level)
        o.set(oracle.F_MGX, x, level)
        o.set(oracle.F_MGB, b, level)
        s.mg("SMOOTH", level, 2)
        o.mg(oracle.MG_SMOOTH, level, 2)
        assert rel_l2(s.download("MGX", level), o.get(oracle.F_MGX, level)) <= TOL32
    o.close()
