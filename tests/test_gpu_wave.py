"""GPU parity of the free-surface coupling against the FP64 oracle.

The wave passes (src/shaders/wave.ts:56-388) compute in FP32 on the device,
as the reference's spectra do; the oracle restates them in FP64, so every
comparison here is at FP32 tolerance, written in each test.  The set-up
puts two active drops (one with the capillary contact window open), a
surfactant film with Marangoni pull, a non-flat surface and an upward MAC
top layer under the surface, so every forcing term of the force pass and
both terms of the lid pressure (fluid.ts:16-78) are exercised.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from gpu_helpers import pair
from mac3d.lib import Mac3dError

pytestmark = pytest.mark.gpu

TOLW = 2e-5  # FP32 FFTs and transcendental functions against FP64
GEOM = dict(half_extent=0.06, half_xz=0.016, depth=0.048, surface_y=0.0, marangoni=0.8)
DROPS = [(0.004, -0.003, -0.002, 0.020, -0.4, 0.0, 0.55, 1.3),
         (-0.010, 0.010, 0.006, 0.018, -0.2, 0.0, 0.20, 2.1)]


def wave_state(n, seed=5):
    """a smooth random surface (eta, d eta/dt, Laplacian slot, film)"""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n)
    filt = np.exp(-(k[:, None] ** 2 + k[None, :] ** 2) * 400.0)
    W = np.empty((n, n, 4))
    for c, a in enumerate((0.01, 0.05, 0.002, 0.3)):
        f = np.fft.ifft2(np.fft.fft2(rng.standard_normal((n, n))) * filt).real
        W[..., c] = a * f / np.abs(f).max()
    W[..., 3] = np.abs(W[..., 3])
    return W


def coupled(n=64, prec=0, warm_steps=2, scene=0):
    w, s, o = pair(precision=prec, warm_steps=warm_steps)
    s.wave_enable(n=n, scene=scene, **GEOM)
    o.wave_enable(n=n, scene=scene, **GEOM)
    s.set_drops(DROPS)
    o.set_drops(DROPS)
    W = wave_state(n)
    s.upload("WAVE", W)
    o.view(oracle.F_WAVE)[...] = W
    # an upward top layer under the surface feeds the force pass
    V = o.get(oracle.F_V)
    V[:, -3:, :] += 0.2
    o.set(oracle.F_V, V)
    s.upload("V", V)
    return w, s, o


def _cmp_wave(s, o, tol):
    """eta, d eta/dt and film to rel L2 <= tol; the 9-point Laplacian is a
    difference of near-equal FP32 values, so its error is bounded by the
    stencil's weight (2.4 (n/256)^2) times tol times the size of eta"""
    a, b = s.download("WAVE"), o.get(oracle.F_WAVE)
    n = a.shape[0]
    for c in (0, 1, 3):
        e = rel_l2(a[..., c], b[..., c])
        assert e <= tol, f"wave channel {c}: rel L2 {e:.3e} > {tol:.1e}"
    bound = 2.4 * (n / 256.0) ** 2 * tol * np.linalg.norm(b[..., 0])
    e = np.linalg.norm(a[..., 2] - b[..., 2])
    assert e <= bound, f"wave Laplacian: L2 error {e:.3e} > {bound:.3e}"


@pytest.mark.parametrize("n", [64, 512])
def test_wave_step_against_oracle(n):
    """Three wave steps (force, FFT rows/columns, propagation, inverse FFT,
    finalize) from the same surface, drops and MAC top layer."""
    w, s, o = coupled(n)
    with s:
        for _ in range(3):
            s.stage("WAVE", w.dt)
            o.stage(oracle.ST_WAVE, w.dt)
        _cmp_wave(s, o, TOLW)
    o.close()


def test_top_pressure_against_oracle():
    """evalTopDirichletPressure on the lid: the wave's hydrostatic and
    capillary terms and both drops' neck overpressure."""
    w, s, o = coupled(64)
    with s:
        s.stage("WAVE", w.dt)
        o.stage(oracle.ST_WAVE, w.dt)
        s.stage("TOP_PRESSURE", w.dt)
        o.stage(oracle.ST_TOP_PRESSURE, w.dt)
        a, b = s.download("PTOP"), o.get(oracle.F_PTOP)
        assert np.abs(b).max() > 0
        assert rel_l2(a, b) <= TOLW
    o.close()


@pytest.mark.parametrize("prec,tol", [(0, 2e-5), (2, 2e-5)])
def test_coupled_steps_against_oracle(prec, tol):
    """Five full steps with the coupling on: the wave step and lid pressure
    first, then the MAC step with p_top in its right-hand side and top-face
    projection.  FP64 mode too: its wave is FP32 (bound TOLW-like)."""
    w, s, o = coupled(64, prec=prec)
    with s:
        s.step(w.dt, 5)
        o.step(w.dt, 5)
        for f in ("U", "V", "W", "P", "DYE"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            assert e <= tol, f"{f}: rel L2 {e:.3e} > {tol:.1e}"
        _cmp_wave(s, o, TOLW)
        assert rel_l2(s.download("PTOP"), o.get(oracle.F_PTOP)) <= TOLW
    o.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_flat_calm_surface_is_the_flat_lid(prec):
    """With a flat surface and no drops the lid pressure is exactly 0 and the
    coupled step equals the flat-lid step bit for bit."""
    w, s1, o = pair(precision=prec, warm_steps=2)
    o.close()
    w, s2, o = pair(precision=prec, warm_steps=2)
    with s1, s2:
        s2.wave_enable(n=32, **GEOM)
        s1.step(w.dt, 3)
        s2.step(w.dt, 3)
        assert np.all(s2.download("PTOP") == 0.0)
        for f in ("U", "V", "W", "P", "DYE"):
            assert np.array_equal(s1.download(f), s2.download(f)), f
    o.close()


def test_wave_calls_need_the_coupling_and_valid_arguments():
    """ESTATE before mac3d_wave_enable; EINVAL for a grid that is not a
    power of two in [16, 1024] or more than 8 drops; disabling restores the
    flat lid (the wave fields are gone)."""
    w, s, o = pair(warm_steps=0)
    o.close()
    with s:
        with pytest.raises(Mac3dError, match="ESTATE"):
            s.stage("WAVE", w.dt)
        with pytest.raises(Mac3dError, match="ESTATE"):
            s.set_drops(DROPS)
        for n in (8, 100, 2048):
            with pytest.raises(Mac3dError, match="EINVAL"):
                s.wave_enable(n=n)
        s.wave_enable(n=16)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.set_drops(DROPS * 5)
        s.wave_enable(False)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.download("PTOP")


def test_packed_uniform_block_sets_the_coupling():
    """mac3d_set_uniforms reads surfaceParams, domainBounds and the 8 drop
    slots from the reference's packed 128-float block (common.ts:26-43,
    uniformLayout.ts:56-190) and gives the same wave step as the same values
    set one by one."""
    w, s1, o = pair(warm_steps=0)
    o.close()
    w, s2, o = pair(warm_steps=0)
    o.close()
    blk = np.zeros(128, dtype=np.float32)
    blk[40], blk[41] = 0.0, 0.8                       # surfaceY, marangoni
    blk[52], blk[53], blk[54], blk[55] = 0.016, -0.048, 0.0, 1.1  # domainBounds (top = surface)
    for q, d in enumerate(DROPS):
        blk[64 + 8 * q: 72 + 8 * q] = d
    depth = float(np.float32(blk[54]) - np.float32(blk[53]))
    with s1, s2:
        s1.wave_enable(n=64, half_extent=0.06)
        s1.set_uniforms(blk)
        s2.wave_enable(n=64, half_extent=0.06, half_xz=float(blk[52]), depth=depth,
                       surface_y=0.0, marangoni=float(blk[41]))
        s2.set_drops([tuple(float(x) for x in np.float32(d)) for d in DROPS])
        W = wave_state(64)
        s1.upload("WAVE", W)
        s2.upload("WAVE", W)
        for s in (s1, s2):
            s.stage("WAVE", w.dt)
            s.stage("TOP_PRESSURE", w.dt)
        assert np.array_equal(s1.download("WAVE"), s2.download("WAVE"))
        assert np.array_equal(s1.download("PTOP"), s2.download("PTOP"))
        with pytest.raises(Mac3dError, match="EINVAL"):
            s1.set_uniforms(blk[:64])
        # the tank top (domainBounds.z) must be the surface height
        bad = blk.copy()
        bad[54] = 0.004
        with pytest.raises(Mac3dError, match="EINVAL"):
            s1.set_uniforms(bad)


def test_scene_advection_against_oracle():
    """The advection with the scene terms (drop necks, Marangoni pull and
    kinematic top from the wave, sponges; fluid.ts:198-321) after one wave
    step, against the oracle's."""
    w, s, o = coupled(64, scene=1)
    with s:
        s.stage("WAVE", w.dt)
        o.stage(oracle.ST_WAVE, w.dt)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_F2C)
        before = s.download("V")
        s.stage("ADVECT", w.dt)
        o.stage(oracle.ST_ADVECT, w.dt)
        assert not np.allclose(s.download("V"), before)
        for f in ("U", "V", "W"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            assert e <= TOLW, f"{f}: rel L2 {e:.3e}"
    o.close()


def test_scene_coupled_steps_against_oracle():
    """Five full FP32 steps with the wave, the lid pressure and the
    advection's scene terms."""
    w, s, o = coupled(64, scene=1)
    with s:
        s.step(w.dt, 5)
        o.step(w.dt, 5)
        for f in ("U", "V", "W", "P", "DYE"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            assert e <= 2e-5, f"{f}: rel L2 {e:.3e}"
        _cmp_wave(s, o, TOLW)
    o.close()


DEEP = dict(half_extent=0.5, half_xz=0.2, depth=0.8, surface_y=0.0, marangoni=0.3,
            vort_conf=1.4, sim_time=0.9, schlieren=0.05)


def deep_scene(scene=1, prec=0):
    """A deep tank (the unfurling gate opens 0.11 m under the surface) with a
    swirling velocity whose curl is 0.25-1.26 in grid units on 98% of the
    cells (so the confinement acts nearly everywhere) and random dye."""
    w, s, o = pair(precision=prec, warm_steps=0)
    nz, ny, nx = o.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    fields = {"U": 0.9 * np.sin(2 * np.pi * j / 9.0) * np.cos(2 * np.pi * k / 11.0),
              "V": 0.9 * np.sin(2 * np.pi * k / 7.0 + 0.3) * np.cos(2 * np.pi * i / 10.0),
              "W": 0.9 * np.sin(2 * np.pi * i / 8.0 + 0.5) * np.cos(2 * np.pi * j / 12.0),
              "DYE": np.clip(np.random.default_rng(3).random(o.shape) * 1.2 - 0.1, 0, 1.2)}
    for f, a in fields.items():
        o.set(getattr(oracle, "F_" + f), a)
        s.upload(f, a)
    s.wave_enable(n=64, scene=scene, **DEEP)
    o.wave_enable(n=64, scene=scene, **DEEP)
    return w, s, o


def test_scene_casson_against_oracle():
    """The Casson pass with the scene terms (fluid.ts:331-509): the curl of
    the advected centre velocity, then the vorticity confinement, unfurling
    forcing and clamp after the viscous update.  The scene's own increment
    (scene on minus scene off, about 1% of the velocity) is compared by
    itself, so FP32 rounding of the velocity does not hide it."""
    dt = 0.02
    got = {}
    for scene in (1, 0):
        w, s, o = deep_scene(scene)
        with s:
            s.stage("CASSON", dt)
            o.stage(oracle.ST_CASSON, dt)
            got[scene] = {f: (s.download(f), o.get(getattr(oracle, "F_" + f)))
                          for f in ("U", "V", "W")}
            if scene:
                c_dev, c_orc = s.download("CURL"), o.get(oracle.F_CURL)
                assert rel_l2(c_dev, c_orc) <= 2e-6, "curl"
        o.close()
    for f in ("U", "V", "W"):
        (d1, o1), (d0, o0) = got[1][f], got[0][f]
        assert rel_l2(d1, o1) <= TOLW, f
        inc_d, inc_o = d1 - d0, o1 - o0
        assert np.linalg.norm(inc_o) > 3e-3 * np.linalg.norm(o0), f
        e = rel_l2(inc_d, inc_o)
        assert e <= 1e-4, f"{f} scene increment: rel L2 {e:.3e}"


def test_scene_deep_steps_against_oracle():
    """Three full FP32 steps of the deep swirling set-up with every scene
    term on (wave, lid pressure, advection and Casson scene terms)."""
    w, s, o = deep_scene(1)
    with s:
        s.step(w.dt, 3)
        o.step(w.dt, 3)
        for f in ("U", "V", "W", "P", "DYE"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            assert e <= 2e-5, f"{f}: rel L2 {e:.3e}"
    o.close()


def test_scene_clock_from_the_uniform_block():
    """simTime [35], fluidParams.x [44] and particleParams.x [60] of the
    packed block set the Casson scene terms as the same values given to
    mac3d_wave_enable do (device memory: a captured step sees them)."""
    w, s1, o = deep_scene(1)
    o.close()
    w, s2, o = deep_scene(1)
    o.close()
    blk = np.zeros(128, dtype=np.float32)
    blk[35], blk[44], blk[60] = 2.5, 0.7, 0.11
    blk[40], blk[41] = DEEP["surface_y"], DEEP["marangoni"]
    blk[52], blk[53], blk[54] = DEEP["half_xz"], -DEEP["depth"], 0.0
    with s1, s2:
        s1.set_uniforms(blk)
        s2.wave_enable(n=64, scene=1, **dict(DEEP, sim_time=float(blk[35]),
                                             vort_conf=float(blk[44]),
                                             schlieren=float(blk[60])))
        for s in (s1, s2):
            s.stage("CASSON", 0.02)
        for f in ("U", "V", "W"):
            assert np.array_equal(s1.download(f), s2.download(f)), f


def test_scene_face_to_centre_against_oracle():
    """The centre velocity with the scene's top-row stagnation gate
    (fluid.ts:944-959) on the deep swirling set-up with the two drops: equal
    to the oracle's, held back only on downward top-row centres."""
    w, s, o = deep_scene(1)
    with s:
        s.set_drops(DROPS)
        o.set_drops(DROPS)
        V = o.get(oracle.F_V)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_F2C)
        for f in ("CU", "CV", "CW"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            assert e <= 1e-6, f"{f}: rel L2 {e:.3e}"
        cv = s.download("CV")
        plain = 0.5 * (V[:, -2] + V[:, -1])
        down = plain < 0
        assert np.all(np.abs(cv[:, -1][~down] - plain[~down]) <= 1e-6)
        assert np.all(np.abs(cv[:, -1][down]) < np.abs(plain[down]) + 1e-6)
        assert np.abs(cv[:, -1] - plain).max() > 0.05
    o.close()


def test_scene_graph_replays_see_the_uniform_block():
    """The scene step replayed from its captured graph against the same
    steps launched eagerly (profiling switches graphs off), with the packed
    block changing simTime, vorticity confinement, hemolysis and the drop
    slots between steps: these live in device memory, so every replay sees
    them and matches the eager step bit for bit."""
    w, s, o = deep_scene(1)
    o.close()
    w, e, o = deep_scene(1)
    o.close()
    blk = np.zeros(128, dtype=np.float32)
    blk[40], blk[41] = DEEP["surface_y"], DEEP["marangoni"]
    blk[52], blk[53], blk[54] = DEEP["half_xz"], -DEEP["depth"], 0.0
    with s, e:
        e.profile(True)
        for step in range(4):
            blk[35], blk[44], blk[59] = 0.3 * step, 1.0 + 0.2 * step, 0.05 * step
            d = DROPS[step % 2]
            blk[64:72] = (d[0], d[1], d[2], d[3], d[4], d[5], d[6] * (1 + step), d[7])
            for x in (s, e):
                x.set_uniforms(blk)
                x.step(w.dt, 1)
            for f in ("U", "V", "W", "P", "DYE"):
                assert np.array_equal(s.download(f), e.download(f)), (step, f)
