"""GPU parity of the reference's pass B on the MAC grid against the FP64
oracle: parcel emission (PARTICLE_EMIT, src/shaders/particles.ts:53-186),
the 4-channel concentration resolve (CONC_RESOLVE,
src/shaders/advect.ts:81-298) with its lysed-channel predictor, and full
steps with the concentration on (BloodDropEngine.ts:2073-2108).

The device computes in FP32 as the reference does.  The emission's R3
sequence and hash are formed in single precision without fused
multiply-adds on both sides, so every new parcel's angle and radius match to
FP32 rounding of the trigonometry; the overwritten parcels' transfers into
the lysed counter are integer and equal.  The resolve's cut-offs (lysed
below 2.5e-6, bed below 1e-6, the early-outs) can fall the other way where
FP32 and FP64 straddle a threshold; the bars written in each test hold
anyway."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from gpu_helpers import pair
from mac3d import MacSolver, scenarios
from mac3d.lib import config_for

pytestmark = pytest.mark.gpu

# a 0.2 m wide, 0.2 m deep tank (6.25 mm x 4.2 mm cells on 32 x 48 x 32), so
# the necks' parcels (up to ~35 mm from a drop's axis) stay inside
GEOM = dict(half_extent=0.25, half_xz=0.1, depth=0.2, surface_y=0.0, marangoni=0.8)
DROPS = [(0.004, -0.003, -0.002, 0.020, -0.4, 0.0, 0.85, 1.3),
         (-0.045, 0.010, 0.030, 0.018, -0.2, 0.0, 0.30, 2.1)]


def conc_pair(prec=0, physics=1, scene=1, ring=4096, warm=0, refine=1):
    """the tiny workload's solver and oracle with the coupling, two drops
    and the 4-channel concentration on its refine-times grid"""
    w, s, o = pair(precision=prec, warm_steps=warm)
    for x in (s, o):
        x.wave_enable(n=64, scene=scene, physics=physics, **GEOM)
        x.set_drops(DROPS)
        x.conc_enable(True, refine=refine)
    P = np.zeros((ring, 8))
    s.parcels_upload(P)
    o.parcels_set(P)
    return w, s, o


def _parcels_close(a, b, pos_tol=2e-6):
    assert np.array_equal(a[:, 3] > 0, b[:, 3] > 0)
    live = b[:, 3] > 0
    dp = np.abs(a[live, :3] - b[live, :3]).max()
    dm = np.abs(a[live, 3] - b[live, 3]).max()
    print(f"[conc] parcels: {live.sum()} live, max |dx| {dp:.2e} m, max |dm| {dm:.2e}", flush=True)
    assert dp <= pos_tol and dm <= 1e-6


@pytest.mark.parametrize("refine", [1, 2])
def test_emission_against_oracle(refine):
    """600 emissions into a 4096-slot ring from slot 3900 (wrapping), two
    drops sharing by drain rate, over a ring whose written slots hold 50 live
    parcels: the new parcels' positions within 2 um of the oracle's, masses,
    stretch and lysis to FP32 rounding, the transfers into the lysed counter
    (on the concentration grid, the MAC grid's or twice as fine) equal, and
    the untouched slots unchanged."""
    w, s, o = conc_pair(refine=refine)
    rng = np.random.default_rng(4)
    P = np.zeros((4096, 8))
    live = (3900 + rng.choice(600, 50, replace=False)) % 4096
    P[live, 0] = rng.uniform(-0.05, 0.05, 50)
    P[live, 2] = rng.uniform(-0.05, 0.05, 50)
    P[live, 1] = rng.uniform(-0.19, -0.01, 50)
    P[live, 3] = rng.uniform(0.01, 0.134, 50)
    P[live, 4] = 0.002
    P = P.astype(np.float32).astype(np.float64)
    with s:
        s.parcels_upload(P)
        o.parcels_set(P)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_F2C)
        s.parcels_emit(3900, 600)
        o.parcels_emit(3900, 600)
        s.stage("PARCELS", 0.0)
        o.stage(oracle.ST_PARCELS, 0.0)
        a, b = s.parcels_download().astype(np.float64), o.parcels()
        slots = (3900 + np.arange(600)) % 4096
        assert np.all(b[slots, 3] == 0.134)
        _parcels_close(a, b)
        assert np.abs(a[slots, 4:8] - b[slots, 4:8]).max() <= 1e-6
        rest = np.ones(4096, dtype=bool)
        rest[slots] = False
        assert np.array_equal(a[rest], P[rest].astype(np.float32))
        ls, lo = s.splat_download(1), o.splat(1)
        assert lo.sum() > 0 and np.array_equal(ls, lo)
        # the request is consumed: a second stage emits nothing new
        s.stage("PARCELS", 0.0)
        assert np.array_equal(s.parcels_download(), a.astype(np.float32))
    o.close()


@pytest.mark.parametrize("refine", [1, 2])
def test_resolve_against_oracle(refine):
    """Predict + resolve of a smooth random 4-channel state in a swirling
    flow, with the counters of 3000 splatted parcels (intact and lysed, floor
    bed near the bottom): intact, lysed and bed within 1e-5 relative L2; the
    oxygenation/shear channel within 1e-4 -- it is a chain of smoothsteps of
    the convective shear |v0 - v_mid| / (dt/2), which divides an FP32
    difference of two velocity samples by half a step (measured 3.6e-5).
    On the grid twice as fine the lysed channel's predictor is compared as
    well (its own field), at 1e-5."""
    w, s, o = conc_pair(refine=refine)
    rng = np.random.default_rng(7)
    nz, ny, nx = o.splat_shape()
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i, j, k = i / refine, j / refine, k / refine  # the same fields in MAC-cell units
    C = np.stack([0.2 * (1 + np.sin(i / 3.0) * np.cos(j / 5.0)) * (k > 2) * (k < nz / refine - 3),
                  0.05 * (1 + np.cos(i / 4.0 + j / 6.0)),
                  np.zeros(i.shape),
                  0.3 * np.exp(-0.5 * j) * (1 + np.sin(k / 4.0))], axis=-1)
    nz, ny, nx = o.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    P = np.zeros((3000, 8))
    P[:, 0], P[:, 2] = rng.uniform(-0.05, 0.05, (2, 3000))
    P[:, 1] = rng.uniform(-0.195, -0.02, 3000)
    P[:, 3] = 0.1
    P[:, 4:7] = rng.standard_normal((3000, 3)) * 0.002
    P[:, 7] = rng.integers(0, 3000, 3000)
    P = P.astype(np.float32).astype(np.float64)
    U = {f: 0.1 * np.sin(0.3 * i + 0.5 * q) * np.cos(0.2 * j + q) for q, f in enumerate("UVW")}
    with s:
        for f, a in U.items():
            s.upload(f, a)
            o.set(getattr(oracle, "F_" + f), a)
        s.upload("CONC", C)
        o.view(oracle.F_CONC)[...] = np.moveaxis(C, -1, 0)
        s.parcels_upload(P)
        o.parcels_set(P)
        s.stage("FACE_TO_CENTRE")
        o.stage(oracle.ST_F2C)
        s.stage("PARCELS", 1e-3)
        o.stage(oracle.ST_PARCELS, 1e-3)
        s.stage("CONC_PREDICT", 1e-3)
        o.stage(oracle.ST_CONC_PRED, 1e-3)
        e = rel_l2(s.download("CONC_PRED"), o.get(oracle.F_CONC_PRED))
        print(f"[conc] predictor: rel L2 {e:.3e}", flush=True)
        assert e <= 1e-5
        s.stage("CONC_RESOLVE", 1e-3)
        o.stage(oracle.ST_CONC_RES, 1e-3)
        got = s.download("CONC")
        want = np.moveaxis(o.get(oracle.F_CONC), 0, -1)
        for q, name in enumerate(("intact", "lysed", "oxygenation", "bed")):
            e = rel_l2(got[..., q], want[..., q])
            print(f"[conc] {name}: rel L2 {e:.3e}, |want| {np.abs(want[..., q]).max():.3e}",
                  flush=True)
            assert np.abs(want[..., q]).max() > 0, name
            assert e <= (1e-4 if name == "oxygenation" else 1e-5), name
    o.close()


@pytest.mark.parametrize("refine", [1, 2])
def test_steps_with_the_concentration_against_oracle(refine):
    """Three full FP32 steps of the reference scene with the 4-channel
    concentration on the MAC grid or twice as fine: wave, drops, scene terms
    and reference physics, 400 emissions per step into a 4096-slot ring,
    parcels, predict and resolve; the Casson pass, buoyancy and hemolysis
    read the concentration."""
    w, s, o = conc_pair(0, warm=0, refine=refine)
    with s:
        for step in range(3):
            for x in (s, o):
                x.parcels_emit(400 * step, 400)
            s.step(w.dt, 1)
            o.step(w.dt, 1)
        for f in ("U", "V", "W", "P"):
            e = rel_l2(s.download(f), o.get(getattr(oracle, "F_" + f)))
            print(f"[conc] {f}: rel L2 {e:.3e}", flush=True)
            assert e <= 2e-5, f
        got = s.download("CONC")
        want = np.moveaxis(o.get(oracle.F_CONC), 0, -1)
        for q, name in enumerate(("intact", "lysed", "oxygenation", "bed")):
            e = rel_l2(got[..., q], want[..., q])
            print(f"[conc] {name}: rel L2 {e:.3e}", flush=True)
            assert e <= 1e-4, name
        assert np.abs(want[..., 0]).max() > 0
        _parcels_close(s.parcels_download().astype(np.float64), o.parcels(), pos_tol=1e-5)
    o.close()


def test_conc_calls_and_checkpoint():
    """ESTATE without the coupling; EINVAL in FP64 and for an emission above
    8192; the concentration is part of a checkpoint and must match the
    context's setting on load."""
    w = scenarios.workload("tiny", 1)
    from mac3d.lib import Mac3dError
    with MacSolver(config_for(w)) as s:
        with pytest.raises(Mac3dError, match="ESTATE"):
            s.conc_enable(True)
        s.wave_enable(n=64, **GEOM)
        s.conc_enable(True)
        s.parcels_upload(np.zeros((16, 8)))
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.parcels_emit(0, 8193)
        C = np.random.default_rng(1).random(s.shape + (4,)) * 0.1
        s.upload("CONC", C)
        assert np.allclose(s.download("CONC"), C.astype(np.float32))
        ck = s.checkpoint_save()
        s.conc_enable(False)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.checkpoint_load(ck)
        s.conc_enable(True)
        s.upload("CONC", C * 0)
        s.checkpoint_load(ck)
        assert np.allclose(s.download("CONC"), C.astype(np.float32))
    with MacSolver(config_for(w)) as s:
        # on the grid twice as fine: its own field sizes, counters and
        # checkpoint section; a checkpoint of one grid does not load into the
        # other; a change of grid starts from zero
        s.wave_enable(n=64, **GEOM)
        s.conc_enable(True, refine=2)
        assert s.conc_shape() == tuple(2 * a for a in s.shape)
        s.parcels_upload(np.zeros((16, 8)))
        assert s.splat_download(0).shape == s.conc_shape()
        C = np.random.default_rng(2).random(s.conc_shape() + (4,)) * 0.1
        s.upload("CONC", C)
        ck = s.checkpoint_save()
        assert ck.nbytes > 8 * 4 * int(np.prod(s.shape)) * 4
        s.conc_enable(True, refine=1)
        assert np.all(s.download("CONC") == 0)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.checkpoint_load(ck)
        s.conc_enable(True, refine=2)
        s.checkpoint_load(ck)
        assert np.allclose(s.download("CONC"), C.astype(np.float32))
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.conc_enable(True, refine=3)
    w64 = scenarios.scaled(w, precision=2)
    with MacSolver(config_for(w64)) as s:
        s.wave_enable(n=64, **GEOM)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.conc_enable(True)
