"""Pins of the oracle's reference-physics mode (the coupling's physics flag):
the reference's own lattice Casson law, tanh-saturated surface-gated
buoyancy, clamp-to-edge viscous neighbours (FLUID_CASSON_FORCES,
src/shaders/fluid.ts:377-511), the projection's +-2 clamp (fluid.ts:908)
and the dye predictor's speed branches and wall zeroing (CONC_PREDICT,
src/shaders/advect.ts:28-72; CONC_RESOLVE's zeroing, advect.ts:100-103).
Each against a closed form a plausible slip (a constant, a sign, a branch
threshold, an edge rule) would fail."""
import math

import numpy as np
import pytest

import oracle


def ref_oracle(nx=8, ny=12, nz=8, dx=1e-3, scene=0, **kw):
    o = oracle.Oracle(nx, ny, nz, dx, levels=1)
    o.wave_enable(n=16, half_extent=0.05, surface_y=0.0, scene=scene, physics=1, **kw)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_DYE):
        o.set(f, np.zeros(o.shape))
    return o


def lattice_alpha(gam, c, scale=1.0):
    c = min(max(c, 0.0), 1.5)
    sq = math.sqrt(0.00010 * (1.0 + 1.35 * c))
    y = math.sqrt(0.0036 * scale * c * c / max(gam + 0.40, 0.40))
    return min(max((sq + y) ** 2, 0.00010), 0.0044)


@pytest.mark.parametrize("c,scale", [(1.0, 1.0), (0.6, 2.5), (2.0, 1.0)])
def test_lattice_casson_quadratic_profile(c, scale):
    """U = a j^2, blood at concentration c everywhere (2.0 is clamped to
    1.5): interior Laplacian 2a and strain 50 |U(j+1) - U(j-1)| = 200 a j, so
    U gains 2a alpha(200 a j, c) with the lattice law (no dt: the reference's
    alpha is per step)."""
    o = ref_oracle(casson_scale=scale)
    nz, ny, nx = o.shape
    a = 2e-4
    j = np.arange(ny, dtype=float)
    U = np.broadcast_to((a * j * j)[None, :, None], o.shape).copy()
    U[:, :, -1] = 0.0
    o.set(oracle.F_U, U)
    o.set(oracle.F_DYE, np.full(o.shape, c))
    o.stage(oracle.ST_CASSON, 5e-4)
    got = o.get(oracle.F_U)
    for jj in range(1, ny - 1):
        gam = 50.0 * math.sqrt((4 * a * jj) ** 2 + 1e-6)
        want = a * jj * jj + 2 * a * lattice_alpha(gam, c, scale)
        assert abs(got[3, jj, 3] - want) <= 1e-15 + 1e-13 * abs(want), jj


def test_clamp_to_edge_neighbours():
    """No blood (alpha = clamp(0.01^2, 1e-4, 0.0044) = 1e-4): a field linear in
    z changes only at the edges, where the clamped neighbour repeats the edge
    value: lap = U(1) - U(0) at k = 0 (the no-slip ghost would give
    U(1) - 3 U(0)) and U(nz-2) - U(nz-1) at the top plane."""
    o = ref_oracle()
    nz, ny, nx = o.shape
    k = np.arange(nz, dtype=float) + 1.0
    U = np.broadcast_to(k[:, None, None], o.shape).copy()
    U[:, :, -1] = 0.0
    o.set(oracle.F_U, U)
    o.stage(oracle.ST_CASSON, 5e-4)
    got = o.get(oracle.F_U)
    assert got[3, 5, 3] == U[3, 5, 3]
    assert got[0, 5, 3] == pytest.approx(1.0 + 1e-4 * (2.0 - 1.0), abs=1e-15)
    assert got[-1, 5, 3] == pytest.approx(nz + 1e-4 * ((nz - 1) - nz), abs=1e-14)
    assert np.all(got[:, :, -1] == 0.0)


def test_tanh_surface_gated_buoyancy():
    """Still fluid, random blood: each y-face gains
    -b 0.345 1.24 tanh(rho/1.24) smoothstep(0.002, 0.024, -y_face) dt with
    rho = 1.24 (c_j + c_j+1)/2 (the top row's own value), nothing where
    rho <= 0.0008; on a 12-row tank of 2 mm cells under a surface at 0 the
    gate runs from 0 (top face) to 1."""
    dx, dt, b = 2e-3, 5e-4, 0.48
    o = ref_oracle(dx=dx, buoyancy=b)
    nz, ny, nx = o.shape
    rng = np.random.default_rng(4)
    c = rng.uniform(0.0, 1.3, o.shape)
    c[:, 3, :] = 1e-4  # rows 3-4 faces: rho below the threshold where both are tiny
    c[:, 4, :] = 1e-4
    o.set(oracle.F_DYE, c)
    o.stage(oracle.ST_CASSON, dt)
    got = o.get(oracle.F_V)
    up = np.concatenate([c[:, 1:, :], c[:, -1:, :]], axis=1)
    rho = 1.24 * 0.5 * (c + up)
    depth = ny * dx
    yf = -depth + (np.arange(ny) + 1.0) * dx
    ss = np.vectorize(oracle.smoothstep)
    gate = ss(0.002, 0.024, np.maximum(0.0, -yf))[None, :, None]
    want = np.where(rho > 0.0008, -b * 0.345 * 1.24 * np.tanh(rho / 1.24) * gate * dt, 0.0)
    assert np.abs(got - want).max() < 1e-15
    assert np.all(got[:, 3, :] == 0.0) and np.all(got[:, -1, :] == 0.0)
    assert np.abs(got).max() > 1e-5
    # the linear workload law would be -g (rho_b - rho_w)/rho_w c_face dt:
    assert not np.allclose(got, -9.81 * 0.06 * 0.5 * (c + up) * dt)


def test_projection_clamps_to_two():
    """Faces beyond +-2 after the pressure step are clipped (fluid.ts:908)."""
    o = ref_oracle()
    nz, ny, nx = o.shape
    rng = np.random.default_rng(5)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        o.set(f, rng.uniform(-3.0, 3.0, o.shape))
    u0 = o.get(oracle.F_U)
    p = rng.uniform(-0.5, 0.5, o.shape)
    o.set(oracle.F_P, p)
    o.stage(oracle.ST_PROJECT)
    got = o.get(oracle.F_U)
    want = u0[:, :, :-1] - (p[:, :, 1:] - p[:, :, :-1])
    assert np.abs(got[:, :, :-1] - np.clip(want, -2.0, 2.0)).max() < 1e-15
    assert np.abs(want).max() > 2.0 and np.abs(got).max() == 2.0


def _lin(shape):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return np.stack([i, j, k], axis=-1).astype(float)


@pytest.mark.parametrize("speed,branch", [(0.003, "static"), (0.05, "euler"), (0.15, "rk2")])
def test_dye_predictor_branches_and_walls(speed, branch):
    """A linear dye in a linear centre flow (trilinear sampling is exact):
    below speed^2 2.5e-5 the predictor copies the cell's own dye; below 0.012
    it traces one stage, t - v(t) dt/h; above, the RK2 midpoint trace; the
    lateral walls and the top row predict 0, and the resolve writes 0 there."""
    nx, ny, nz, dx, dt = 16, 20, 12, 1e-3, 1e-3
    o = ref_oracle(nx, ny, nz, dx)
    t = _lin(o.shape)
    P, Q = 0.4, np.array([0.031, -0.027, 0.019])
    o.set(oracle.F_DYE, P + t @ Q)
    A = np.array([0.6, -0.48, 0.64]) * speed
    B = np.array([[0.01, -0.004, 0.006], [0.003, 0.01, -0.005], [-0.004, 0.006, 0.012]]) * speed
    cv = A + t @ B.T
    for a, f in enumerate((oracle.F_CU, oracle.F_CV, oracle.F_CW)):
        o.set(f, cv[..., a])
    o.stage(oracle.ST_DYE_PRED, dt)
    got = o.get(oracle.F_DYE_PRED)
    s = lambda x: (A + x @ B.T) * (dt / dx)  # noqa: E731
    if branch == "static":
        want = P + t @ Q
    elif branch == "euler":
        want = P + (t - s(t)) @ Q
    else:
        want = P + (t - s(t - 0.5 * s(t))) @ Q
    sp2 = (cv ** 2).sum(-1)
    inner = np.zeros(o.shape, dtype=bool)
    inner[1:-1, :-1, 1:-1] = True
    lim = {"static": (0.0, 2.5e-5), "euler": (2.5e-5, 0.012), "rk2": (0.012, 1e9)}[branch]
    ok = inner & (sp2 >= lim[0]) & (sp2 < lim[1])
    assert ok.sum() > 0.5 * inner.sum()
    assert np.abs(got - want)[ok].max() < 1e-12
    wall = ~inner
    assert np.all(got[wall] == 0.0)
    o.stage(oracle.ST_DYE_RES, dt)
    assert np.all(o.get(oracle.F_DYE)[wall] == 0.0)
