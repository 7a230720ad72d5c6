"""Pins of the oracle's Lagrangian parcel step (PARTICLE_ADVECT_SPLAT,
src/shaders/particles.ts:189-463): the fixed-point splat against the textbook
trilinear split, the 4-point ribbon's mass and centroid, settling in still
water by the RK2 midpoint rule, and mass conservation between the parcels
and the lysed counter."""
import math

import numpy as np
import pytest

import oracle

S = 65536.0  # SPLAT_FP_SCALE


def tank(nx=16, ny=24, nz=16, hx=0.008, depth=0.024, surface=0.0, **kw):
    """a still tank (zero velocity and dye) with the coupling on"""
    o = oracle.Oracle(nx, ny, nz, 2 * hx / nx, levels=2)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_DYE):
        o.set(f, np.zeros(o.shape))
    o.wave_enable(n=16, half_extent=0.05, half_xz=hx, depth=depth, surface_y=surface, **kw)
    o.stage(oracle.ST_F2C)
    return o


def uvw_of(o, p, hx=0.008, depth=0.024, surface=0.0):
    return np.array([(p[0] + hx) / (2 * hx), (p[1] - (surface - depth)) / depth, (p[2] + hx) / (2 * hx)])


def trilinear(shape, u, mass):
    """round(mass 65536 w) on the eight cells around normalised position u"""
    nz, ny, nx = shape
    out = np.zeros(shape, dtype=np.int64)
    c = np.array([u[0] * nx, u[1] * ny, u[2] * nz]) - 0.5
    b = np.floor(c).astype(int)
    f = c - b
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                out[b[2] + dz, b[1] + dy, b[0] + dx] += int(np.rint(mass * S * w))
    return out


def one(p, stretch=(0.002, 0.0, 0.0), mass=0.1, ticks=0.0):
    return [p[0], p[1], p[2], mass, *stretch, ticks]


def test_splat_of_a_parcel_at_a_cell_centre():
    """dt = 0 leaves the parcel where it is; a short parcel (stretch below
    3.8 mm) at a cell centre puts round(65536 m) into that cell alone."""
    o = tank()
    h = 0.001
    p = (-0.008 + 5.5 * h, -0.024 + 9.5 * h, -0.008 + 7.5 * h)
    o.parcels_set([one(p, mass=0.1)])
    o.stage(oracle.ST_PARCELS, 0.0)
    sp = o.splat(0)
    assert sp[7, 9, 5] == round(0.1 * S) and sp.sum() == sp[7, 9, 5]
    assert o.splat(1).sum() == 0
    assert np.array_equal(o.parcels()[0], np.array(one(p, mass=0.1)))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_splat_is_the_trilinear_split(seed):
    """dt = 0, a short parcel anywhere inside: the eight counters are the
    textbook trilinear weights of its position, each rounded."""
    rng = np.random.default_rng(seed)
    o = tank()
    p = (rng.uniform(-0.004, 0.004), rng.uniform(-0.02, -0.005), rng.uniform(-0.004, 0.004))
    m = rng.uniform(0.05, 0.13)
    o.parcels_set([one(p, mass=m)])
    o.stage(oracle.ST_PARCELS, 0.0)
    want = trilinear(o.shape, uvw_of(o, p), m)
    assert np.array_equal(o.splat(0).astype(np.int64), want)
    assert abs(int(want.sum()) - m * S) <= 4


@pytest.mark.parametrize("direction", [(1, 0, 0), (0.3, -0.9, 0.2), (0, 1, 0)])
def test_ribbon_splat_keeps_mass_and_centroid(direction):
    """A stretched parcel (12 mm) splats a 4-point ribbon with weights
    0.21, 0.29, 0.29, 0.21 at -0.42 s, -0.14 s + v, 0.14 s - v, 0.42 s
    (particles.ts:440-461): its counters hold the whole mass (to rounding)
    and their centroid is the parcel's position (trilinear splats keep the
    first moment away from the walls)."""
    o = tank(nx=32, ny=48, nz=32, hx=0.016, depth=0.048)
    d = np.array(direction, float)
    st = 0.012 * d / np.linalg.norm(d)
    p = (0.0011, -0.0213, -0.0007)
    m = 0.12
    o.parcels_set([one(p, stretch=st, mass=m)])
    o.stage(oracle.ST_PARCELS, 0.0)
    sp = o.splat(0).astype(float)
    assert abs(sp.sum() - m * S) <= 16
    nz, ny, nx = o.shape
    h = 0.032 / nx
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cx = (-0.016 + (i + 0.5) * h * 1.0)
    cy = (-0.048 + (j + 0.5) * 0.048 / ny)
    cz = (-0.016 + (k + 0.5) * h)
    cen = np.array([(sp * cx).sum(), (sp * cy).sum(), (sp * cz).sum()]) / sp.sum()
    assert np.allclose(cen, p, atol=2e-6)
    assert np.count_nonzero(sp) > 8  # more than one trilinear splat


def test_settling_in_still_water_is_the_rk2_midpoint():
    """No flow, 2 cm under the surface (no unfurling, no floor contact), a
    young parcel whose meniscus hold is 0 (tail rank below 0.36): it falls by
    dt settle(y - dt settle(y) / 2) with settle(y) =
    mix(0.020, 0.046, smoothstep(0.04, 0.42, b)) mix(0.38, 1, smoothstep(0.002,
    0.055, b)), b the depth below the surface (particles.ts:201-215, 330-352);
    x and z stay."""
    ss = oracle.smoothstep

    def settle(y):
        b = 0.0 - y
        return (0.020 + 0.026 * ss(0.04, 0.42, b)) * (0.38 + 0.62 * ss(0.002, 0.055, b))
    o = tank(depth=0.1)  # floor contact starts 4.2 cm above the floor
    g = 0  # tail rank fract(0 * 0.7548776662) = 0: no meniscus hold
    p = (0.0012, -0.020, -0.0009)
    o.parcels_set([one(p, mass=0.1)])
    dt = 0.01
    o.stage(oracle.ST_PARCELS, dt)
    q = o.parcels()[g]
    y1 = p[1] - dt * settle(p[1] - 0.5 * dt * settle(p[1]))
    assert q[1] == pytest.approx(y1, rel=0, abs=1e-15)
    assert q[0] == p[0] and q[2] == p[2]
    assert math.floor(q[7]) == max(1, round(dt * 240))  # age ticks + 2 (+ lysis fraction)


def test_hemolysis_and_bed_conserve_mass():
    """Over one step of many parcels in still water (no wall sponge inside
    0.68 hx), each parcel's lost mass is in the lysed counter: hemolysis
    moves dm, the floor bed 0.68 dm of its share (particles.ts:386-425), so
    stored mass + lysed / 65536 equals the starting mass less 0.32 of the bed
    share; the intact counter holds the stored mass."""
    rng = np.random.default_rng(4)
    o = tank(hemolysis=3.0)
    n = 400
    r = rng.uniform(0, 0.6 * 0.008, n)
    th = rng.uniform(0, 2 * np.pi, n)
    P = np.zeros((n, 8))
    P[:, 0], P[:, 2] = r * np.cos(th), r * np.sin(th)
    P[:, 1] = rng.uniform(-0.0235, -0.003, n)
    P[:, 3] = rng.uniform(0.08, 0.134, n)
    P[:, 4] = 0.002
    P[:, 7] = rng.integers(0, 3000, n) + rng.uniform(0, 0.6, n)
    o.parcels_set(P)
    dt = 0.02
    o.stage(oracle.ST_PARCELS, dt)
    Q = o.parcels()
    lysed, intact = o.splat(1).sum() / S, o.splat(0).sum() / S
    assert lysed > 0.01 * P[:, 3].sum()
    # bed share: the parcels that ended within 2.6 cm of the floor lose
    # 0.32 of their bed transfer; bound the identity by that transfer
    near = Q[:, 1] < -0.024 + 0.026
    moved = P[:, 3].sum() - Q[:, 3].sum()
    assert moved > 0
    assert lysed <= moved + 8 * n / S
    bed_max = (P[near, 3] * 0.095).sum()
    assert moved - lysed <= 0.32 * bed_max + 8 * n / S
    assert abs(intact - Q[:, 3].sum()) <= 8 * n / S


def _settle_young(y, surface=0.0):
    """settling speed of a young parcel with no meniscus hold at height y
    (particles.ts:201-215): mono(b) gate(b), b the depth below the surface"""
    ss = oracle.smoothstep
    b = surface - y
    return (0.020 + 0.026 * ss(0.04, 0.42, b)) * (0.38 + 0.62 * ss(0.002, 0.055, b))


@pytest.mark.parametrize("fast", [True, False])
def test_advection_in_a_linear_flow_is_ralston_rk3(fast):
    """A parcel in an affine centre-velocity field c(p) = A + B p (trilinear
    sampling reproduces it exactly inside the tank), 2 cm under the surface
    (no unfurling, no floor or wall terms), young and with no meniscus hold:
    its velocity is v(p) = c(p) - settle(y) e_y.  Above speed^2 0.0025 the
    step is Ralston's RK3, p + dt (2 k1 + 3 k2 + 4 k3) / 9 with k2 at
    p + dt k1 / 2 and k3 at p + 3 dt k2 / 4; below it the RK2 midpoint
    p + dt k2 (particles.ts:310-336).  Both against the closed form to
    rounding, and each far from the other rule."""
    o = tank(nx=16, ny=48, nz=16, depth=0.1)
    nz, ny, nx = o.shape
    hx, depth = 0.008, 0.1
    x = -hx + (np.arange(nx) + 0.5) * (2 * hx / nx)
    y = -depth + (np.arange(ny) + 0.5) * (depth / ny)
    z = -hx + (np.arange(nz) + 0.5) * (2 * hx / nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    A = np.array([0.07, 0.02, 0.05]) if fast else np.array([0.012, 0.004, -0.008])
    B = np.array([[0.0, 0.8, 0.5], [1.5, 0.0, -0.7], [0.4, 0.9, -1.0]])
    P3 = np.stack([X, Y, Z], axis=-1)
    c = A + P3 @ B.T
    for a, f in enumerate((oracle.F_CU, oracle.F_CV, oracle.F_CW)):
        o.set(f, c[..., a])

    def v(p):
        w = A + B @ p
        w[1] -= _settle_young(p[1])
        return w
    p0 = np.array([0.0011, -0.020, -0.0007])
    o.parcels_set([one(p0, mass=0.1)])
    dt = 0.01
    o.stage(oracle.ST_PARCELS, dt)
    got = o.parcels()[0][:3]
    k1 = v(p0)
    k2 = v(p0 + 0.5 * dt * k1)
    k3 = v(p0 + 0.75 * dt * k2)
    rk3 = p0 + dt * (2 * k1 + 3 * k2 + 4 * k3) / 9.0
    rk2 = p0 + dt * k2
    assert (k1 @ k1 > 0.0025) == fast
    want, other = (rk3, rk2) if fast else (rk2, rk3)
    assert np.abs(got - want).max() < 1e-15
    assert np.abs(other - want).max() > 1e-9
