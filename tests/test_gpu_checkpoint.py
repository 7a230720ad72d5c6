"""mac3d_checkpoint_save / mac3d_checkpoint_load through the C ABI: the state
a step evolves goes to a caller-owned buffer and comes back bit for bit.

The reference resumes by replay (resetSimulation + seekToSimTime,
BloodDropEngine.ts:1682-1763, 2142-2179): a checkpoint must resume to the
same numbers the uninterrupted run produces, in every precision mode, with
the wave coupling, the scene terms and the parcels on, in the same context
and in a fresh one."""
import numpy as np
import pytest

from mac3d import Mac3dError, MacSolver, scenarios
from mac3d.lib import config_for

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "P", "DYE")
DROPS = [(0.004, -0.003, -0.002, 0.02, -0.4, 0.0, 0.55, 1.3),
         (-0.01, 0.01, 0.006, 0.018, -0.2, 0.0, 0.2, 2.1)]
BOX = dict(n=64, half_extent=0.06, half_xz=0.016, depth=0.048, surface_y=0.0, marangoni=0.8)


def fields(s, wave=False, parcels=False):
    out = {f: s.download(f) for f in FIELDS}
    if wave:
        out["WAVE"] = s.download("WAVE")
    if parcels:
        out["parcels"] = s.parcels_download()
        out["splat0"] = s.splat_download(0)
        out["splat1"] = s.splat_download(1)
    return out


def same(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_round_trip_is_bit_exact(precision):
    """save after 3 steps; 3 more steps; load; the same 3 steps again: equal
    bit for bit (mixed mode: the FP64 pressure history and the refinement
    log come back too, so the warm start and the iteration count repeat);
    a fresh context of the same configuration loading the same bytes gives
    the same numbers."""
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=precision)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.step(w.dt, 3)
        ck = s.checkpoint_save()
        st_a = s.step(w.dt, 3)
        a = fields(s)
        s.checkpoint_load(ck)
        st_b = s.step(w.dt, 3)
        same(a, fields(s))
        assert st_a["vcycles"] == st_b["vcycles"]
    with MacSolver(config_for(w)) as t:
        t.checkpoint_load(ck)
        t.step(w.dt, 3)
        same(a, fields(t))


def test_round_trip_with_the_scene_wave_and_parcels():
    """The coupled scene (wave, drops, scene terms) with 4000 parcels: the
    wave state and the parcels are in the checkpoint; after a load the next
    step and parcel stage repeat bit for bit, splat counters included."""
    w = scenarios.workload("tiny", 1)
    rng = np.random.default_rng(2)
    P = np.zeros((4000, 8), dtype=np.float32)
    P[:, 0], P[:, 2] = rng.uniform(-0.008, 0.008, (2, 4000))
    P[:, 1] = rng.uniform(-0.04, -0.004, 4000)
    P[:, 3] = 0.1
    P[:, 4] = 0.002
    P[:, 7] = rng.integers(0, 2000, 4000)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.wave_enable(scene=1, **BOX)
        s.set_drops(DROPS)
        s.parcels_upload(P)
        s.step(w.dt, 2)
        s.stage("PARCELS", w.dt)
        ck = s.checkpoint_save()

        def advance(x):
            x.step(w.dt, 1)
            x.stage("PARCELS", w.dt)
            return fields(x, wave=True, parcels=True)
        a = advance(s)
        s.checkpoint_load(ck)
        same(a, advance(s))
    with MacSolver(config_for(w)) as t:
        t.wave_enable(scene=1, **BOX)
        t.set_drops(DROPS)
        t.checkpoint_load(ck)
        same(a, advance(t))


def test_load_rejects_mismatches():
    """EINVAL, state untouched, for: another grid, another precision, a wave
    grid that differs from the context's coupling, a truncated buffer, a
    corrupted magic."""
    w = scenarios.workload("tiny", 1)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.step(w.dt, 1)
        ck = s.checkpoint_save()
        before = fields(s)
        bad = ck.copy()
        bad[0] ^= 0xFF
        for buf in (ck[:100], ck[:-1], bad):
            with pytest.raises(Mac3dError, match="EINVAL"):
                s.checkpoint_load(buf)
        s.wave_enable(**BOX)
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.checkpoint_load(ck)  # saved without the coupling
        s.wave_enable(False)
        same(before, fields(s))
        s.checkpoint_load(ck)
    for other in (dict(nx=64, levels=3), dict(precision=1)):
        w2 = scenarios.scaled(w, **other)
        with MacSolver(config_for(w2)) as t:
            with pytest.raises(Mac3dError, match="EINVAL"):
                t.checkpoint_load(ck)
