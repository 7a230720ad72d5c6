"""GPU parity of the Lagrangian parcel step (PARTICLE_ADVECT_SPLAT,
src/shaders/particles.ts:189-463) against the FP64 oracle.

The device computes in FP32, as the reference's parcels do; positions and
masses are compared at FP32 tolerance.  The splat counters are integer sums
of round(65536 m w): a weight can round the other way where the FP32 and
FP64 products straddle a half, and a parcel near one of the step's branch
thresholds could take the other branch, so the counters are compared by their
totals and their L1 difference, with the bounds written in each test; the
parcels themselves are bounded one by one."""
import numpy as np
import pytest

import oracle
from gpu_helpers import pair
from mac3d.lib import Mac3dError

pytestmark = pytest.mark.gpu

BOX = dict(half_extent=0.05, half_xz=0.016, depth=0.048, surface_y=0.0)


def cloud(n, seed=7, hx=0.016, ymin=-0.048, ymax=0.0):
    """parcels inside 0.6 hx of the axis: 40% stretched (10 mm, a 4-point
    ribbon), the rest short (2 mm), random ages and lysis"""
    rng = np.random.default_rng(seed)
    P = np.zeros((n, 8))
    r, th = rng.uniform(0, 0.6 * hx, n), rng.uniform(0, 2 * np.pi, n)
    P[:, 0], P[:, 2] = r * np.cos(th), r * np.sin(th)
    P[:, 1] = rng.uniform(ymin + 0.003, ymax - 0.003, n)
    P[:, 3] = rng.uniform(0.05, 0.134, n)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    P[:, 4:7] = d * np.where(rng.random(n) < 0.4, 0.010, 0.002)[:, None]
    P[:, 7] = rng.integers(0, 4000, n) + rng.uniform(0, 0.6, n)
    return P.astype(np.float32).astype(np.float64)


def run_both(n, dt, hemolysis=1.0):
    w, s, o = pair(warm_steps=2)
    P = cloud(n)
    for x in (s, o):
        x.wave_enable(n=16, hemolysis=hemolysis, **BOX)
    s.stage("FACE_TO_CENTRE")
    o.stage(oracle.ST_F2C)
    s.parcels_upload(P)
    o.parcels_set(P)
    s.stage("PARCELS", dt)
    o.stage(oracle.ST_PARCELS, dt)
    return w, s, o, P


def _cmp_splat(a, b, l1_rel, tot_rel):
    a, b = a.astype(np.int64), b.astype(np.int64)
    assert b.sum() > 0
    assert abs(a.sum() - b.sum()) <= tot_rel * b.sum(), (a.sum(), b.sum())
    e = np.abs(a - b).sum() / b.sum()
    assert e <= l1_rel, f"splat L1 {e:.3e} > {l1_rel:.1e}"


def test_splat_only_against_oracle():
    """dt = 0: no motion, only the splat of 20000 parcels (twice: each call
    clears the counters first); each weight is at most one count off (FP32
    rounding), so the L1 difference is below 2e-5 of the total."""
    w, s, o, P = run_both(20000, 0.0)
    with s:
        s.stage("PARCELS", 0.0)  # again: the counters are cleared first, not added to
        assert np.array_equal(s.parcels_download().astype(np.float64), P)
        _cmp_splat(s.splat_download(0), o.splat(0), 2e-5, 2e-6)
        assert s.splat_download(1).sum() == 0 and o.splat(1).sum() == 0
    o.close()


def test_parcel_step_against_oracle():
    """One step (dt 0.5 ms) of 20000 parcels through the workload's flow:
    RK3/RK2 advection with settling and unfurling, stretch, floor bed,
    hemolysis and both splats.  Every parcel: position within 1e-6 m, mass
    within 1e-5 relative, stretch within 1e-4 relative (measured maxima over
    the 20000: 1.9e-7 m, 5.3e-7, 1.4e-5; none of these parcels lies close
    enough to a branch threshold of the step -- RK3 above speed^2 0.0025, the
    stretch update, the ribbon splat above stretch 3.8 mm, the hemolysis
    headroom -- for FP32 rounding to change its branch); age and active flag
    equal; counters within 1e-3 (L1) and 1e-4 (totals)."""
    w, s, o, P = run_both(20000, 0.0005)
    with s:
        a, b = s.parcels_download().astype(np.float64), o.parcels()
        assert np.array_equal(a[:, 3] > 0, b[:, 3] > 0)
        moved = np.linalg.norm(b[:, :3] - P[:, :3], axis=1)
        assert np.median(moved) > 1e-6
        dpos = np.linalg.norm(a[:, :3] - b[:, :3], axis=1)
        dm = np.abs(a[:, 3] - b[:, 3]) / np.maximum(b[:, 3], 1e-6)
        dst = np.linalg.norm(a[:, 4:7] - b[:, 4:7], axis=1) / np.maximum(
            np.linalg.norm(b[:, 4:7], axis=1), 1e-6)
        for name, e, tol in (("position", dpos, 1e-6), ("mass", dm, 1e-5), ("stretch", dst, 1e-4)):
            worst = int(np.argmax(e))
            assert e[worst] <= tol, f"{name}: parcel {worst} off by {e[worst]:.3e} > {tol}"
        assert np.all(np.abs(np.floor(a[:, 7]) - np.floor(b[:, 7])) == 0)
        _cmp_splat(s.splat_download(0), o.splat(0), 1e-3, 1e-4)
        _cmp_splat(s.splat_download(1), o.splat(1), 1e-3, 1e-4)
    o.close()


def test_parcel_calls_need_the_coupling():
    """ESTATE for parcels before mac3d_wave_enable and for the stage before
    any upload; EINVAL for a download beyond the uploaded parcels."""
    w, s, o = pair(warm_steps=0)
    o.close()
    with s:
        with pytest.raises(Mac3dError, match="ESTATE"):
            s.parcels_upload(np.zeros((4, 8)))
        s.wave_enable(n=16, **BOX)
        with pytest.raises(Mac3dError, match="ESTATE"):
            s.stage("PARCELS", 0.001)
        s.parcels_upload(np.zeros((4, 8)))
        s.stage("PARCELS", 0.001)  # inactive parcels: counters cleared, nothing splatted
        assert s.splat_download(0).sum() == 0
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.parcels_download(5)


def test_parcels_after_a_velocity_upload():
    """The parcel stage samples the centre velocity of the current faces:
    after an upload of U the library forms it first, so the stage equals an
    explicit face->centre pass followed by the stage, bit for bit."""
    res = []
    for explicit in (False, True):
        w, s, o = pair(warm_steps=2)
        o.close()
        with s:
            s.wave_enable(n=16, **BOX)
            s.stage("FACE_TO_CENTRE")
            s.parcels_upload(cloud(5000, seed=3))
            U = s.download("U")
            s.upload("U", U * 1.5 + 0.01)
            if explicit:
                s.stage("FACE_TO_CENTRE")
            s.stage("PARCELS", 0.0005)
            res.append((s.parcels_download(), s.splat_download(0), s.splat_download(1)))
    for x, y in zip(*res):
        assert np.array_equal(x, y)


def test_parcel_upload_rejects_counter_overflow():
    """The splat counters hold 65536 per unit mass in 32 bits, so parcels
    whose total mass reaches 65536 units are refused (EINVAL), as are
    negative or non-finite masses."""
    w, s, o = pair(warm_steps=0)
    o.close()
    with s:
        s.wave_enable(n=16, **BOX)
        P = cloud(4)
        P[:, 3] = 16384.0
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.parcels_upload(P)
        P[:, 3] = 16383.0
        s.parcels_upload(P)
        P[0, 3] = -1.0
        with pytest.raises(Mac3dError, match="EINVAL"):
            s.parcels_upload(P)
