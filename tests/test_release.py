"""Drop releases as blobs (mac3d_release_blobs, host arithmetic of the
library; no device).

The reference lets a drop go with MultiDropSystem.spawnAerialDrop
(src/sim/dropKinematics.ts:354-391) and moves it with the aerial branch of
MultiDropSystem.step (:651-707): per substep, free fall under G_SLOW, and on
the substep where the lowest point passes WATER_SURFACE_Y the contact time
solved from the quadratic, the centre set to surface + R and the velocity to
v0 + G t.  The library takes the contact state in closed form.  These tests
step the reference's loop itself (restated below, substep by substep, as the
engine calls it at 1/240 s) and hold the closed form to it, check the clamps
of the offset disk and the radius, and hold the chain-rule Jacobian to
central differences of the map."""
import ctypes as C
import math

import numpy as np
import pytest

from mac3d import lib

G_SLOW, SURFACE, MAX_OFF, R_MIN, R_MAX = -0.155, 0.30, 0.26, 0.018, 0.095


def aerial_contact(x, z, radius, start_y, vel_y, dt=1.0 / 240.0, max_steps=100000):
    """spawnAerialDrop then MultiDropSystem.step's aerial branch, substep by
    substep, until the drop changes to 'coalescing': returns the centre and
    the vertical velocity at contact (dropKinematics.ts:49-58, 354-391,
    651-707)."""
    r = math.hypot(x, z)
    if not (r <= MAX_OFF or r < 1e-8):
        x, z = x * MAX_OFF / r, z * MAX_OFF / r
    R = min(max(radius, R_MIN), R_MAX)
    y, v = start_y, vel_y
    for _ in range(max_steps):
        y_next = y + v * dt + 0.5 * G_SLOW * dt * dt
        if y_next - R <= SURFACE:
            dt_air = 0.0
            if y - R > SURFACE:
                qa, qb, qc = 0.5 * G_SLOW, v, y - R - SURFACE
                disc = max(0.0, qb * qb - 4.0 * qa * qc)
                dt_air = min(max((-qb - math.sqrt(disc)) / (2.0 * qa), 0.0), dt)
            return (x, SURFACE + R, z), v + G_SLOW * dt_air, R
        y, v = y_next, v + G_SLOW * dt
    raise AssertionError("no contact")


RELEASES = [
    dict(x=0.05, z=-0.03, radius=0.058, start_y=0.495, vel_y=-0.075),  # the hero drop
    dict(x=-0.12, z=0.2, radius=0.03, start_y=0.55, vel_y=0.0),        # from the pipette tip
    dict(x=0.0, z=0.0, radius=0.07, start_y=0.45, vel_y=0.2),          # thrown upward first
    dict(x=0.3, z=0.2, radius=0.05, start_y=0.5, vel_y=-0.1),          # outside the disk
    dict(x=0.01, z=0.02, radius=0.12, start_y=0.6, vel_y=-0.05),       # radius capped
    dict(x=0.01, z=0.02, radius=0.01, start_y=0.6, vel_y=-0.05),       # radius floored
    dict(x=0.02, z=0.0, radius=0.04, start_y=0.32, vel_y=-0.3),        # already touching
]


def grid_map(nx=48, ny=64, nz=48):
    return lib.release_default_map(lib.default_config(nx=nx, ny=ny, nz=nz))


def test_default_map():
    m = grid_map()
    assert m.cell == pytest.approx(0.96 / 48, rel=1e-15)
    assert (m.origin_x, m.origin_z) == pytest.approx((-0.48, -0.48), rel=1e-15)
    assert m.origin_y + 64 * m.cell == pytest.approx(SURFACE, abs=1e-15)  # top face at the surface
    assert m.origin_y == pytest.approx(-0.98, abs=1e-14)  # DOMAIN_Y_MIN at the reference aspect
    assert (m.gravity, m.max_offset, m.r_min, m.r_max) == (G_SLOW, MAX_OFF, R_MIN, R_MAX)
    assert (m.vel_scale, m.submerge) == (1.0, 1.0)


@pytest.mark.parametrize("rel", RELEASES)
def test_blob_is_the_reference_contact_state(rel):
    m = grid_map()
    m.vel_scale, m.submerge = 2.5, 1.3
    (xc, yc, zc), vc, R = aerial_contact(**rel)
    (b,), _ = lib.release_blobs(m, [rel])
    assert b.cx == pytest.approx((xc - m.origin_x) / m.cell, abs=1e-12)
    assert b.cz == pytest.approx((zc - m.origin_z) / m.cell, abs=1e-12)
    assert b.radius == pytest.approx(R / m.cell, rel=1e-14)
    # the contact centre is one radius above the surface; the blob's sits
    # `submerge` radii under it
    assert yc == pytest.approx(SURFACE + R, rel=1e-15)
    assert b.cy == pytest.approx((SURFACE - 1.3 * R - m.origin_y) / m.cell, rel=1e-14)
    assert b.w0 == pytest.approx(-2.5 * vc, rel=1e-10, abs=1e-14)


def test_contact_speed_does_not_depend_on_the_substep():
    rel = RELEASES[0]
    speeds = [aerial_contact(**rel, dt=dt)[1] for dt in (1 / 30, 1 / 240, 1 / 1000, 0.37)]
    assert max(speeds) - min(speeds) <= 1e-12


def _blob_vec(m, rel):
    (b,), _ = lib.release_blobs(m, [rel])
    return np.array([getattr(b, n) for n in lib.BLOB_PARAMS])


@pytest.mark.parametrize("rel", RELEASES)
def test_jacobian_against_central_differences(rel):
    m = grid_map()
    m.vel_scale, m.submerge = 2.5, 1.3
    _, J = lib.release_blobs(m, [rel])
    h = 1e-7
    for b, name in enumerate(lib.RELEASE_PARAMS):
        lo, hi = dict(rel), dict(rel)
        lo[name] -= h
        hi[name] += h
        fd = (_blob_vec(m, hi) - _blob_vec(m, lo)) / (2 * h)
        assert np.allclose(J[0][:, b], fd, rtol=1e-6, atol=1e-6 * np.abs(J[0]).max()), (name, J[0][:, b], fd)


def test_clamped_parameters_pass_no_gradient():
    m = grid_map()
    _, J = lib.release_blobs(m, [RELEASES[4], RELEASES[5]])
    for q in range(2):  # the radius clamp binds: no blob entry moves with the radius
        assert np.all(J[q][:, 2] == 0.0)
    _, J = lib.release_blobs(m, [RELEASES[6]])
    # at or below the surface at release: the contact speed is the release
    # velocity, so only vel_y reaches w0
    assert J[0][4, 3] == 0.0 and J[0][4, 2] == 0.0 and J[0][4, 4] == -1.0
    # outside the disk the offset moves only along the rim: the radial
    # direction is in the Jacobian's null space
    x, z = RELEASES[3]["x"], RELEASES[3]["z"]
    _, J = lib.release_blobs(m, [RELEASES[3]])
    radial = np.array([x, z]) / math.hypot(x, z)
    assert np.abs(J[0][[0, 2]][:, :2] @ radial).max() <= 1e-12


def test_bad_arguments():
    m = grid_map()
    L = lib.load_library()
    with pytest.raises(lib.Mac3dError):
        lib.release_blobs(m, [dict(RELEASES[0], x=float("nan"))])
    bad = grid_map()
    bad.cell = 0.0
    with pytest.raises(lib.Mac3dError):
        lib.release_blobs(bad, [RELEASES[0]])
    bad = grid_map()
    bad.gravity = 0.1
    with pytest.raises(lib.Mac3dError):
        lib.release_blobs(bad, [RELEASES[0]])
    blobs = (lib.Blob * 1)()
    assert L.mac3d_release_blobs(C.byref(m), lib._releases([RELEASES[0]]), 0, blobs, None) == 1
    assert L.mac3d_release_blobs(C.byref(m), lib._releases([RELEASES[0]]), 1, blobs, None) == 0
