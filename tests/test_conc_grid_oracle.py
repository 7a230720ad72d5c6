"""Pins of the oracle's concentration on its own grid, twice the MAC grid per
axis as the reference holds it (CONC_GRID 192x256x192 over VEL_GRID
96x128x96, src/shaders/common.ts:4-15).

What the finer grid changes is where things are read and written:
- the Casson pass, its buoyancy and the scene's unfurling gate sample the
  concentration at each MAC cell centre and top y-face (fluid.ts:446-447,
  477-478, 491);
- the parcels splat into, and read their hemolysis exposure from, the
  concentration grid (particles.ts:238-301, 403-404);
- the emission moves an overwritten parcel's mass into the lysed counter at
  floor(uvw N) of that grid (particles.ts:102-117);
- CONC_PREDICT and CONC_RESOLVE run per concentration cell, sampling the MAC
  grid's centre velocity (advect.ts:28-298).

The pins:
- trilinear sampling reproduces a linear field, so on linear fields the
  fine grid's reads equal the MAC grid's to rounding;
- a splat's counters keep the parcel's mass and centre it on the parcel's
  texel position;
- the predictor shifts a linear field by exactly v dt in each of its three
  branches;
- the resolve's dilution and counter split hold per fine cell."""
import math

import numpy as np
import pytest

import oracle

S = 65536.0


def tank(refine, nx=16, ny=24, nz=16, hx=0.008, depth=0.024, physics=0, scene=0):
    """still water, the coupling and the 4-channel concentration on its
    refine-times grid"""
    o = oracle.Oracle(nx, ny, nz, 2 * hx / nx, levels=2)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_DYE):
        o.set(f, np.zeros(o.shape))
    o.wave_enable(n=16, half_extent=0.05, half_xz=hx, depth=depth, surface_y=0.0,
                  physics=physics, scene=scene, casson_scale=1.7)
    o.conc_enable(True, refine=refine)
    o.stage(oracle.ST_F2C)
    return o


def centres(shape, hx, depth):
    """world coordinates (x, y, z) of the cell centres of a (nz, ny, nx) grid
    over [-hx, hx] x [-depth, 0] x [-hx, hx]"""
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (-hx + (i + 0.5) / nx * 2 * hx, -depth + (j + 0.5) / ny * depth,
            -hx + (k + 0.5) / nz * 2 * hx)


def test_refine_sets_the_grid():
    o = tank(2)
    assert o.splat_shape() == (32, 48, 32)
    assert o.view(oracle.F_CONC).shape == (4, 32, 48, 32)
    o.conc_enable(True, refine=1)
    assert o.splat_shape() == (16, 24, 16)
    assert np.all(o.get(oracle.F_CONC) == 0)  # a change of grid starts from zero
    o.conc_enable(False)
    assert o.splat_shape() == (16, 24, 16)


@pytest.mark.parametrize("scene", [0, 1])
def test_casson_reads_equal_on_linear_fields(scene):
    """The reference-physics Casson pass (lattice law from the intact
    channel at the cell centre, buoyancy from both channels at the top
    y-face) and, with the scene, the unfurling gate from the cell centre:
    with linear concentration fields the pass gives the same faces whether
    the fields live on the MAC grid or on the grid twice as fine, to
    rounding.  (The top row's face sample, clamped at uvw 0.999, lands
    differently on the two grids, but the buoyancy's surface gate is 0
    there.)  The same linear field laid on the fine grid half a fine cell
    off, which is what reading it at a wrong offset amounts to, moves the
    faces by far more than that."""
    hx, depth = 0.008, 0.024
    lo, hi = (0.1, 1.2) if not scene else (0.002, 0.013)

    def run(refine, off=0.0):
        o = tank(refine, physics=1, scene=scene, hx=hx, depth=depth)
        nz, ny, nx = o.shape
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        # a velocity with curvature, so the Casson law's alpha acts
        for f, ph in ((oracle.F_U, 0.3), (oracle.F_V, 1.1), (oracle.F_W, 2.0)):
            o.set(f, 0.2 * np.sin(0.4 * i + ph) * np.cos(0.3 * j) * np.sin(0.5 * k + ph))
        x, y, z = centres(o.splat_shape(), hx, depth)
        y = y + off
        c = o.view(oracle.F_CONC)
        c[0] = lo + (hi - lo) * (0.3 * (x + hx) / (2 * hx) + 0.5 * (y + depth) / depth +
                                 0.2 * (z + hx) / (2 * hx))
        c[1] = 0.1 * c[0] + 0.02 * (x + hx) / (2 * hx)
        o.stage(oracle.ST_CASSON, 2e-3)
        out = {f: o.get(f) for f in (oracle.F_U, oracle.F_V, oracle.F_W)}
        o.close()
        return out

    a, b = run(1), run(2)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        assert np.abs(a[f] - b[f]).max() <= 1e-12 * max(1.0, np.abs(a[f]).max())
    c = run(2, off=0.5 * depth / 48)
    assert np.abs(c[oracle.F_V] - b[oracle.F_V]).max() > 1e-8


def test_splat_keeps_mass_and_centre_on_the_fine_grid():
    """dt = 0, one short parcel at a time: its intact counts sum to
    round(65536 m) within the eight roundings, and their centroid in the
    fine grid's texel coordinates is the parcel's uvw N - 1/2 (no clamping
    inside the grid)."""
    hx, depth = 0.008, 0.024
    o = tank(2, hx=hx, depth=depth)
    nz, ny, nx = o.splat_shape()
    rng = np.random.default_rng(5)
    for _ in range(6):
        p = np.array([rng.uniform(-0.5, 0.5) * hx, rng.uniform(-0.8, -0.2) * depth,
                      rng.uniform(-0.5, 0.5) * hx])
        P = np.zeros((1, 8))
        P[0, :3] = p
        P[0, 3] = 0.1
        P[0, 4] = 0.001  # stretch 1 mm: one trilinear splat
        o.parcels_set(P)
        o.stage(oracle.ST_PARCELS, 0.0)
        cnt = o.splat(0).astype(np.float64)
        assert abs(cnt.sum() - 0.1 * S) <= 4
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        want = [(p[0] + hx) / (2 * hx) * nx - 0.5, (p[1] + depth) / depth * ny - 0.5,
                (p[2] + hx) / (2 * hx) * nz - 0.5]
        got = [(cnt * i).sum() / cnt.sum(), (cnt * j).sum() / cnt.sum(), (cnt * k).sum() / cnt.sum()]
        assert np.allclose(got, want, atol=1e-3)
        assert o.splat(1).sum() == 0
    o.close()


def test_emission_transfer_lands_in_the_fine_cell():
    """An overwritten live parcel's round(65536 m) goes to the lysed counter
    at floor(uvw N) of the concentration grid."""
    hx, depth = 0.1, 0.5
    o = tank(2, nx=16, ny=32, nz=16, hx=hx, depth=depth)
    nz, ny, nx = o.splat_shape()
    P = np.zeros((8, 8))
    P[3] = [0.0023, -0.0107, -0.0041, 0.0917, 0.002, 0, 0, 3.0]
    o.parcels_set(P)
    o.set_drops([(0.001, -0.02, -0.0005, 0.02, -0.3, 0.0, 0.9, 1.3)])
    o.parcels_emit(3, 1)
    o.stage(oracle.ST_PARCELS, 0.0)
    lys = o.splat(1)
    ci = int(math.floor((0.0023 + hx) / (2 * hx) * nx))
    cj = int(math.floor((-0.0107 + depth) / depth * ny))
    ck = int(math.floor((-0.0041 + hx) / (2 * hx) * nz))
    assert lys[ck, cj, ci] == round(0.0917 * S)
    assert lys.sum() == lys[ck, cj, ci]
    o.close()


@pytest.mark.parametrize("vel", [(0.08, -0.06, 0.05), (0.04, -0.03, 0.05), (0.002, 0.003, -0.001)])
def test_predict_shifts_a_linear_field(vel):
    """A uniform centre velocity and a linear lysed channel on the fine
    grid: away from the walls the predictor holds the field at x - v dt,
    whether it takes the RK2 trace (speed^2 0.0125), the one-stage trace
    (0.0050) or the static copy (speed^2 1.4e-5: the field at x itself);
    the walls and the top row hold 0."""
    hx, depth, dt = 0.008, 0.024, 0.005  # shifts of up to 0.8 fine cells
    o = tank(2, hx=hx, depth=depth)
    for f, v in zip((oracle.F_CU, oracle.F_CV, oracle.F_CW), vel):
        o.set(f, np.full(o.shape, v))
    x, y, z = centres(o.splat_shape(), hx, depth)
    L = lambda x, y, z: 0.3 + 2.0 * x - 1.5 * y + 3.0 * z
    o.view(oracle.F_CONC)[1] = L(x, y, z)
    o.stage(oracle.ST_CONC_PRED, dt)
    got = o.get(oracle.F_CONC_PRED)
    sp2 = sum(v * v for v in vel)
    sh = (0.0, 0.0, 0.0) if sp2 < 2.5e-5 else tuple(v * dt for v in vel)
    want = L(x - sh[0], y - sh[1], z - sh[2])
    inner = (slice(3, -3), slice(3, -3), slice(3, -3))
    assert np.abs(got[inner] - want[inner]).max() <= 1e-12
    if sp2 >= 2.5e-5:
        assert np.abs(got[inner] - L(x, y, z)[inner]).max() > 1e-4
    assert np.all(got[:, :, 0] == 0) and np.all(got[:, :, -1] == 0)
    assert np.all(got[0] == 0) and np.all(got[-1] == 0) and np.all(got[:, -1] == 0)
    o.close()


def test_resolve_dilution_per_fine_cell():
    """Still water, no counters, a lysed blob above y0 + 0.26 on the fine
    grid: the resolve keeps its sum times exactly exp(-0.034 dt)."""
    o = tank(2, ny=40, depth=0.4)
    nz, ny, nx = o.splat_shape()
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.sqrt((i - 15.5) ** 2 + (j - 64) ** 2 + (k - 15.5) ** 2)
    g = np.clip(1.0 - d / 6.0, 0.0, None) * 0.3
    o.view(oracle.F_CONC)[1] = g
    dt = 0.01
    o.stage(oracle.ST_CONC_PRED, dt)
    o.stage(oracle.ST_CONC_RES, dt)
    out = o.get(oracle.F_CONC)
    assert out[1].sum() == pytest.approx(math.exp(-0.034 * dt) * g.sum(), rel=1e-12)
    assert np.all(out[0] == 0) and np.all(out[3] == 0)
    o.close()


def test_resolve_counter_split_per_fine_cell():
    """The lysed counter at a fine cell 103.75 mm above the floor joins the
    lysed channel whole; at fine row 4 (11.25 mm, capture 1) 0.28 of it
    stays lysed and 0.68 enters the bed with the row fade of row 4."""
    o = tank(2, ny=24, depth=0.12)
    o.parcels_set(np.zeros((1, 8)))
    o.stage(oracle.ST_PARCELS, 0.0)
    cnt = o.splat_view(1)
    cnt[...] = 0
    cnt[15, 41, 16] = 3000
    cnt[15, 4, 16] = 5000
    o.stage(oracle.ST_CONC_PRED, 1e-3)
    o.stage(oracle.ST_CONC_RES, 1e-3)
    out = o.get(oracle.F_CONC)
    assert out[1][15, 41, 16] == pytest.approx(3000 / S, rel=1e-12)
    assert out[1][15, 4, 16] == pytest.approx(0.28 * 5000 / S, rel=1e-12)
    fade = 1 - 0.38 * oracle.smoothstep(3.5, 7.0, 4.0)
    assert out[3][15, 4, 16] == pytest.approx(0.68 * 5000 / S * fade, rel=1e-12)
    assert out[0][15, 4, 16] == pytest.approx(out[3][15, 4, 16], rel=1e-12)
    o.close()
