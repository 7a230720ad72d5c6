"""The reference's own scene at the reference's own shape, every part of it
at once, on the GPU against the FP64 oracle:
- the MAC grid VEL_GRID 96x128x96 with 10 mm cells over the 0.96 x 1.28 x
  0.96 m box under the surface at y = 0.30 (common.ts:4-6,
  uniformLayout.ts:92-101), 3 coarse levels (MG_L1..L3, BloodDropEngine.ts:
  88-98), dt FIXED_SIM_DT = 1/60 s (BloodDropEngine.ts:82), velDamp 0.9993;
- the 512-texel free-surface wave (WAVE_GRID_SIZE, common.ts:1);
- the scene forcing with the reference's physics (fluid.ts:198-511);
- the 393,216-parcel ring (PARTICLE_COUNT, common.ts:18) with emission from
  two draining drops;
- the 4-channel concentration on CONC_GRID 192x256x192 (common.ts:12-15),
  twice the MAC grid per axis.

The device steps in FP32 as the reference does, the oracle in FP64, from a
quiescent start; three steps.  The bars are the ones the tiny-grid
concentration steps hold (faces and pressure 2e-5 relative L2, the
concentration channels 1e-4, parcel positions 1e-5 m), except the lysed
channel: it holds only a few hundred counts of the lysed splat counter
(1/65536 each) and is held to two counts per cell absolute, the counters'
quantum (see the test)."""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from gpu_helpers import pair

pytestmark = pytest.mark.gpu

REF_SHAPE = dict(nx=96, ny=128, nz=96, dx=0.01, dt=1.0 / 60.0, levels=3, blobs=[])
GEOM = dict(n=512, half_extent=1.60, half_xz=0.48, depth=1.28, surface_y=0.30,
            marangoni=1.0, scene=1, physics=1, casson_scale=1.0, buoyancy=0.48)
# two drop slots (x, y, z, radius, velY, l2 oscillation, drain rate, seed):
# one draining through the surface, one just entering it
DROPS = [(0.03, 0.285, -0.02, 0.024, -0.35, 0.0, 0.62, 1.3),
         (-0.11, 0.31, 0.07, 0.018, -0.9, 0.0, 0.21, 2.4)]
RING = 393216


def test_reference_scene_at_its_shape():
    w, s, o = pair(warm_steps=0, vel_damp=0.9993, **REF_SHAPE)
    assert (w.nx, w.ny, w.nz) == (96, 128, 96)
    for x in (s, o):
        x.wave_enable(**GEOM)
        x.set_drops(DROPS)
        x.conc_enable(True, refine=2)
    assert o.splat_shape() == (192, 256, 192) and s.conc_shape() == (192, 256, 192)
    P = np.zeros((RING, 8))
    s.parcels_upload(P)
    o.parcels_set(P)
    with s:
        for step in range(3):
            for x in (s, o):
                x.parcels_emit(2000 * step, 2000)
            s.step(w.dt, 1)
            o.step(w.dt, 1)
            print(f"[ref-scene] step {step} done", flush=True)
        for f in ("U", "V", "W", "P"):
            want = o.get(getattr(oracle, "F_" + f))
            e = rel_l2(s.download(f), want)
            print(f"[ref-scene] {f}: rel L2 {e:.3e}, max |want| {np.abs(want).max():.3e}",
                  flush=True)
            assert np.abs(want).max() > 0, f
            assert e <= 2e-5, f
        # the last step's splat counters: integer sums of round(65536 m w)
        # per weight, where FP32 and FP64 products can round a weight either
        # way, so a count may differ by one per weight landing in the cell
        for which, name, l1 in ((0, "intact", 1e-4), (1, "lysed", 1e-2)):
            a_, b_ = s.splat_download(which).astype(np.int64), o.splat(which).astype(np.int64)
            dc = np.abs(a_ - b_)
            print(f"[ref-scene] {name} counter: total {b_.sum()}, cells {int((b_ > 0).sum())}, "
                  f"differing {int((dc > 0).sum())}, max |d| {dc.max()}", flush=True)
            assert b_.sum() > 0 and dc.sum() <= l1 * b_.sum(), name
        got = s.download("CONC")
        want = np.moveaxis(o.get(oracle.F_CONC), 0, -1)
        for q, name in enumerate(("intact", "lysed", "oxygenation", "bed")):
            e = rel_l2(got[..., q], want[..., q])
            d = np.abs(got[..., q] - want[..., q]).max()
            print(f"[ref-scene] {name}: rel L2 {e:.3e}, max {np.abs(want[..., q]).max():.3e}, "
                  f"max |d| {d:.3e} ({d * 65536:.2f} counts)", flush=True)
            if q < 3:
                assert np.abs(want[..., q]).max() > 0, name
            if name == "lysed":
                # at most ~1e-3 per cell (some 70 counts), built from the
                # counters of the three steps: a one-count rounding difference
                # in an earlier step's counter stays in the channel (blended
                # and carried), so the bar is the counter's quantum: two counts
                assert d <= 2.0 / 65536, name
            else:
                assert e <= 1e-4, name
        a, b = s.parcels_download().astype(np.float64), o.parcels()
        live = b[:, 3] > 0
        assert np.array_equal(a[:, 3] > 0, live) and live.sum() >= 5000
        dp = np.abs(a[live, :3] - b[live, :3]).max()
        print(f"[ref-scene] parcels: {live.sum()} live, max |dx| {dp:.2e} m", flush=True)
        assert dp <= 1e-5
    o.close()
