"""Closed-form pins of the oracle's semi-Lagrangian pieces (CPU only).

Trilinear sampling reproduces a field that is linear in the texel
coordinates exactly (between the first and last texel), so under a centre
velocity that is itself linear in position the RK2 backtrace has a closed
form:

    s(t)   = (A + B t) dt / h                  (cells per step)
    t_mid  = t0 - s(t0) / 2                    (fluid.ts:169-174, advect.ts:60-69)
    t_back = t0 - s(t_mid)

and the traced sample of a linear field F(t) = P + Q.t is F(t_back), with the
staggered components sampled half a texel back along their own axis
(sampleMacU/V/W, fluid.ts:154-167).  A non-diagonal B makes the midpoint
differ from the start point by B s(t0)/2, so an Euler step (v0 for the whole
trace), a dropped or sign-flipped half-texel shift, or a transposed axis fails
by far more than rounding.

The MacCormack core of CONC_RESOLVE (advect.ts:196-222) is pinned on a
half-cell uniform shift, where every sample falls on a texel or midway
between two: pred = (f[-1] + f[0])/2, est = (f[-1] + 2 f[0] + f[1])/4,
corr = pred + mc (f[0] - est), clamped to the min/max of the 2x2x2 source
texels at floor(t_back).
"""
import numpy as np
import pytest

import oracle


def _texel_grid(shape):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([i, j, k], axis=-1).astype(np.float64)  # (nz, ny, nx, 3): t = (i, j, k)


def _linear(t, P, Q):
    return P + t @ Q


def _inside(t, shape):
    nz, ny, nx = shape
    hi = np.array([nx - 1, ny - 1, nz - 1], dtype=np.float64)
    return np.all((t >= 0.0) & (t <= hi), axis=-1)


def _set_linear_cvel(o, shape, A, B):
    t = _texel_grid(shape)
    c = A + t @ B.T  # c[..., a] = A[a] + sum_b B[a, b] t[b]
    o.set(oracle.F_CU, c[..., 0])
    o.set(oracle.F_CV, c[..., 1])
    o.set(oracle.F_CW, c[..., 2])


def _trace(t0, A, B, dt, dx):
    s = lambda t: (A + t @ B.T) * (dt / dx)  # noqa: E731
    tm = t0 - 0.5 * s(t0)
    return tm, t0 - s(tm)


# a linear centre velocity (m/s) with a full gradient, about 0.1-0.3 cells per
# step of 1 ms on 1 mm cells, and a displacement gradient of ~1e-2 per cell
A_VEL = np.array([0.21, -0.17, 0.13])
B_VEL = np.array([[0.012, -0.007, 0.009],
                  [0.006, 0.015, -0.011],
                  [-0.008, 0.010, 0.013]])


@pytest.mark.parametrize("axis,field", [(0, oracle.F_U), (1, oracle.F_V), (2, oracle.F_W)])
def test_advect_rk2_midpoint_in_a_linear_flow(axis, field):
    """Each staggered component after FLUID_ADVECT_MAC (fluid.ts:176-192)
    against the closed-form RK2 trace from its face centre (texel t0 = cell
    + 1/2 along its own axis) and the sample half a texel back along that
    axis, on a 16x20x12 grid; faces whose trace leaves the texel range are
    skipped (sampling clamps there)."""
    nx, ny, nz, dx, dt = 16, 20, 12, 1e-3, 1e-3
    o = oracle.Oracle(nx, ny, nz, dx, levels=1)
    shape = o.level_shape(0)
    t = _texel_grid(shape)
    rng = np.random.default_rng(31 + axis)
    fields = {}
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        P, Q = rng.standard_normal(), rng.standard_normal(3)
        fields[f] = (P, Q)
        o.set(f, _linear(t, P, Q))
    _set_linear_cvel(o, shape, A_VEL, B_VEL)
    o.stage(oracle.ST_ADVECT, dt)
    got = o.get(field)

    e = np.zeros(3)
    e[axis] = 0.5
    t0 = t + e
    tm, tb = _trace(t0, A_VEL, B_VEL, dt, dx)
    ts = tb - e
    P, Q = fields[field]
    want = _linear(ts, P, Q)
    ok = _inside(t0, shape) & _inside(tm, shape) & _inside(tb, shape) & _inside(ts, shape)
    assert ok.sum() > 0.5 * ok.size
    err = np.abs(got - want)[ok].max()
    assert err < 1e-12, err
    # the same trace with the start velocity for the whole step (Euler) is
    # far from it: the pin sees the midpoint
    tb_euler = t0 - (A_VEL + t0 @ B_VEL.T) * (dt / dx)
    euler = _linear(tb_euler - e, P, Q)
    assert np.abs(euler - want)[ok].max() > 1e-4


def test_dye_predict_linear_dye_in_a_linear_flow():
    """CONC_PREDICT (advect.ts:48-71) from each cell centre: the predictor of
    a linear dye is the dye at the closed-form RK2 back point.  Under a
    uniform flow (B = 0) that is the dye shifted by the displacement."""
    nx, ny, nz, dx, dt = 16, 20, 12, 1e-3, 1e-3
    for B in (np.zeros((3, 3)), B_VEL):
        o = oracle.Oracle(nx, ny, nz, dx, levels=1)
        shape = o.level_shape(0)
        t = _texel_grid(shape)
        P, Q = 0.4, np.array([0.031, -0.027, 0.019])
        o.set(oracle.F_DYE, _linear(t, P, Q))
        _set_linear_cvel(o, shape, A_VEL, B)
        o.stage(oracle.ST_DYE_PRED, dt)
        got = o.get(oracle.F_DYE_PRED)
        tm, tb = _trace(t, A_VEL, B, dt, dx)
        want = _linear(tb, P, Q)
        ok = _inside(tm, shape) & _inside(tb, shape)
        assert ok.sum() > 0.5 * ok.size
        assert np.abs(got - want)[ok].max() < 1e-12
        if not B.any():
            assert np.abs(want - _linear(t - A_VEL * dt / dx, P, Q)).max() < 1e-15


@pytest.mark.parametrize("axis,sign", [(0, 1.0), (1, -1.0), (2, 1.0), (0, -1.0)])
def test_maccormack_half_cell_shift_closed_form(axis, sign):
    """Predict + resolve of a random dye under a uniform centre velocity of
    half a cell per step along one axis (h = dt = 2^-10, so every texel
    coordinate is exact): interior cells against the closed form above; the
    limiter clamps a few percent of them."""
    n = (12, 16, 8)  # nx, ny, nz
    h = 2.0 ** -10
    o = oracle.Oracle(*n, h, levels=1)
    shape = o.level_shape(0)
    rng = np.random.default_rng(17 + axis)
    f = rng.uniform(0.0, 1.0, shape)
    o.set(oracle.F_DYE, f)
    c = np.zeros(3)
    c[axis] = sign * 0.5
    _set_linear_cvel(o, shape, c, np.zeros((3, 3)))
    o.stage(oracle.ST_DYE_PRED, h)
    o.stage(oracle.ST_DYE_RES, h)
    got = o.get(oracle.F_DYE)

    ax = 2 - axis  # numpy axis of the shift: (z, y, x) storage
    up = lambda a, s: np.roll(a, s, axis=ax)  # noqa: E731  up(a, 1)[i] = a[i - 1]
    s = int(sign)
    f_m, f_p = up(f, s), up(f, -s)  # upstream and downstream neighbour
    pred = 0.5 * (f_m + f)
    est = 0.25 * (f_m + 2.0 * f + f_p)
    mc = oracle.DEFAULTS["maccormack"]
    corr = pred + mc * (f - est)
    # limiter texels: floor(t_back) and floor + 1 on every axis; the back
    # point is i - s/2 along the shift axis and exactly the texel elsewhere
    lo = np.full(shape, np.inf)
    hi = np.full(shape, -np.inf)
    for dz in (0, 1):
        for dy in (0, 1):
            for dxx in (0, 1):
                off = [dxx, dy, dz]
                if s > 0:
                    off[axis] -= 1  # texels i - 1 and i
                g = f
                for a_, o_ in enumerate(off):
                    g = np.roll(g, -o_, axis=2 - a_)
                lo = np.minimum(lo, g)
                hi = np.maximum(hi, g)
    want = np.clip(corr, lo, hi)
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1, 1:-1] = True
    assert np.abs(got - want)[inner].max() < 1e-14
    clipped = (np.abs(want - corr) > 0) & inner
    assert clipped.sum() >= 10 and (inner & ~clipped).sum() >= 10  # both branches
    # a sign-flipped correction or an estimate sampled at the back point
    # differs from the pinned value at most cells
    flipped = np.clip(pred - mc * (f - est), lo, hi)
    back_est = np.clip(pred + mc * (f - 0.5 * (up(pred, s) + pred)), lo, hi)
    assert (np.abs(flipped - want)[inner] > 1e-6).mean() > 0.3
    assert (np.abs(back_est - want)[inner] > 1e-6).mean() > 0.3
