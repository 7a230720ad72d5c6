"""Pins of the oracle's parcel emission (PARTICLE_EMIT,
src/shaders/particles.ts:53-186) and 4-channel concentration resolve
(CONC_RESOLVE, src/shaders/advect.ts:81-298).

These kernels are heuristics defined by their constants, so the pins are
the invariants those constants must keep: emission is exact in count, ring
slots and per-drop shares, moves an overwritten parcel's mass into the lysed
counter at the documented cell, and lays parcels on the R3 sequence within
the neck geometry; the resolve conserves lysed mass under its diffusion
blend, scales it by exactly its dilution factor, splits the lysed counter
between the lysed channel and the floor bed in the stated proportions,
forms the intact channel from the counters, and keeps its walls, early-outs
and ranges.  Parity unpinned beyond these: the oxygenation/shear channel's
interior values (checked only through the device parity) and the bed's
spreading between floor rows."""
import math

import numpy as np
import pytest

import oracle

S = 65536.0


def tank(nx=16, ny=24, nz=16, hx=0.008, depth=0.024, surface=0.0):
    """still water, 1 mm cells, the coupling and the 4-channel concentration on"""
    o = oracle.Oracle(nx, ny, nz, 2 * hx / nx, levels=2)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_DYE):
        o.set(f, np.zeros(o.shape))
    o.wave_enable(n=16, half_extent=0.05, half_xz=hx, depth=depth, surface_y=surface)
    o.conc_enable(True)
    o.stage(oracle.ST_F2C)
    return o


# ------------------------------------------------------------------ emission
def wide(depth=0.2):
    """a 0.2 m wide tank, so the necks' parcels (up to ~35 mm from a drop's
    axis) stay inside the wall sponge and are not retired"""
    return tank(nx=16, ny=32, nz=16, hx=0.1, depth=depth)


def emit(o, drops, cursor, count, ring=64, P=None):
    if P is None:
        P = np.zeros((ring, 8))
    o.parcels_set(P)
    o.set_drops(drops)
    o.parcels_emit(cursor, count)
    o.stage(oracle.ST_PARCELS, 0.0)
    return o.parcels()


DROP_A = (0.001, -0.02, -0.0005, 0.02, -0.3, 0.0, 0.9, 1.3)   # draining slug (slug weight 1)
DROP_B = (-0.045, -0.02, 0.03, 0.02, -0.3, 0.0, 0.3, 2.7)    # 55 mm away: no overlap


def test_emission_count_ring_and_shares():
    """count parcels from cursor wrap round the ring; every other slot is
    untouched; each new parcel has the initial mass 0.134 and a stretch of
    exactly 3.6 mm; the drops' shares follow their drain rates (0.9 : 0.3
    gives 3 : 1 of 40)."""
    o = wide()
    ring = 64
    P0 = np.zeros((ring, 8))
    P0[:, 3] = 0.0
    P0[:, 7] = 5.0
    Q = emit(o, [DROP_A, DROP_B], cursor=50, count=40, ring=ring, P=P0)
    slots = [(50 + g) % ring for g in range(40)]
    new = np.zeros(ring, dtype=bool)
    new[slots] = True
    assert np.all(Q[new, 3] == 0.134)
    assert np.allclose(np.linalg.norm(Q[new, 4:7], axis=1), 0.0036, rtol=1e-13)
    assert np.array_equal(Q[~new], P0[~new])
    # distance from each drop's axis: the first 30 of 40 belong to drop A
    da = np.hypot(Q[slots, 0] - DROP_A[0], Q[slots, 2] - DROP_A[2])
    db = np.hypot(Q[slots, 0] - DROP_B[0], Q[slots, 2] - DROP_B[2])
    # (drop B's neck, drain 0.3, reaches 13 mm plus 4 mm of wake drift; A's
    # reaches 33 mm and lies 55 mm away)
    owner = np.where(db < 0.019, 1, 0)
    assert np.all(da[owner == 0] <= 0.034)
    assert list(owner[:30]) == [0] * 30 and list(owner[30:]) == [1] * 10
    # the parcel stage then splats them (dt 0: no motion): the intact counter
    # holds their mass to rounding, the lysed counter nothing
    assert abs(int(o.splat(0).sum()) - 40 * 0.134 * S) <= 40 * 16
    assert o.splat(1).sum() == 0


def test_emission_moves_an_overwritten_parcel_into_the_lysed_counter():
    """A live parcel at the slot being written: round(65536 m) goes to the
    lysed counter at floor(uvw N) of its position, or of the floor deposit
    point (x, y_min + 0.012, z) when it lies more than 0.32 m under the
    surface (particles.ts:102-117)."""
    o = tank(nx=16, ny=32, nz=16, hx=0.1, depth=0.5)
    nz, ny, nx = o.shape
    P = np.zeros((8, 8))
    P[3] = [0.0023, -0.0107, -0.0041, 0.0917, 0.002, 0, 0, 3.0]
    P[4] = [-0.0031, -0.4, 0.0012, 0.05, 0.002, 0, 0, 3.0]
    o.parcels_set(P)
    o.set_drops([DROP_A])
    o.parcels_emit(3, 2)
    # the emit runs before the parcel step, which also splats the new parcels
    # into the intact counter; the lysed counter holds only the transfers
    o.stage(oracle.ST_PARCELS, 0.0)
    lys = o.splat(1)
    h, dy = 0.2 / nx, 0.5 / ny

    def cell(x, y, z):
        return (int(math.floor((z + 0.1) / h)), int(math.floor((y + 0.5) / dy)),
                int(math.floor((x + 0.1) / h)))
    want = np.zeros(o.shape, dtype=np.int64)
    want[cell(0.0023, -0.0107, -0.0041)] += round(0.0917 * S)
    want[cell(-0.0031, -0.5 + 0.012, 0.0012)] += round(0.05 * S)
    assert np.array_equal(lys.astype(np.int64), want)


def test_emission_lies_on_the_r3_sequence_in_the_neck():
    """A slug (drain 0.9: no wake drift, no thin-wake shrink) well under the
    surface (no nascent core fill): parcel g sits at angle
    2 pi fract((slot + 137 seed) 0.8191725134 + j) about the drop axis with
    |j| <= 0.032 the hash jitter, at radius up to r_neck 1.08 with
    r_neck = clamp(0.84 r + 0.014, 0.020, 0.056), and between 1.4 and 13.5 mm
    (plus the cavity sag, at most 8.5 mm) below the neck top 3.8 mm under the
    calm surface; the stretch points down (y component -0.8 / |t|)."""
    o = wide()
    ring = 256
    Q = emit(o, [DROP_A], cursor=0, count=ring, ring=ring)
    x0, z0, seed = DROP_A[0], DROP_A[2], DROP_A[7]
    r = np.hypot(Q[:, 0] - x0, Q[:, 2] - z0)
    rneck = min(max(0.84 * 0.02 + 0.014, 0.020), 0.056)
    assert np.all(r <= rneck * (1 + 0.036 + 0.026 + 0.018) + 1e-12)
    th = np.arctan2(Q[:, 2] - z0, Q[:, 0] - x0) / (2 * np.pi) % 1.0
    seq = (np.arange(ring, dtype=np.float32) + np.float32(seed) * np.float32(137.0))
    base = (seq.astype(np.float64) * 0.8191725134) % 1.0
    dev = (th - base + 0.5) % 1.0 - 0.5
    assert np.abs(dev).max() <= 0.032 + 2e-3  # jitter, plus the f32 product's rounding
    top = 0.0 - 0.0038
    assert np.all(Q[:, 1] <= top - 0.0014 + 1e-12)
    assert np.all(Q[:, 1] >= top - 0.0085 - 0.0135 * 3.6 - 1e-12)
    st = Q[:, 4:7]
    assert np.all(st[:, 1] < 0)
    assert np.all((Q[:, 7] >= 0) & (Q[:, 7] <= 0.44))


def _hash33_f32(px, py, pz):
    """hash33 (src/shaders/common.ts:76-80) in numpy float32, every
    operation rounded to single precision"""
    f = np.float32

    def fr(x):
        return (x - np.floor(x)).astype(f)

    a0, b0, c0 = fr(px * f(0.1031)), fr(py * f(0.1030)), fr(pz * f(0.0973))
    t = (a0 * (b0 + f(33.33)) + b0 * (a0 + f(33.33))) + c0 * (c0 + f(33.33))
    a, b, c = a0 + t, b0 + t, c0 + t
    return fr((a + b) * c), fr((a + a) * b), fr((b + a) * a)


def test_emission_angle_is_the_single_precision_r3_value():
    """The same slug (no wake drift, so a parcel lies exactly on its ray
    from the drop axis): each parcel's angle about the axis equals
    2 pi r3.x with r3.x = fract(seq 0.8191725134 + hash33(g, 60 t, seed).x
    0.032) formed in float32 as the reference's f32 shader does
    (particles.ts:120-126) -- a one-ulp change of the sum at seq ~ 180
    (1.5e-5) moves the angle by 1e-4 rad, far above the tolerance."""
    f = np.float32
    o = wide()
    ring = 256
    Q = emit(o, [DROP_A], cursor=0, count=ring, ring=ring)
    x0, z0, seed = DROP_A[0], DROP_A[2], f(DROP_A[7])
    g = np.arange(ring, dtype=f)
    seq = g + seed * f(137.0)
    h = _hash33_f32(g, np.full(ring, f(0.0)) * f(60.0), np.full(ring, seed))
    r3x = (seq * f(0.8191725134) + h[0] * f(0.032))
    r3x = (r3x - np.floor(r3x)).astype(f)
    want = 2 * np.pi * r3x.astype(np.float64)
    got = np.arctan2(Q[:, 2] - z0, Q[:, 0] - x0)
    dev = (got - want + np.pi) % (2 * np.pi) - np.pi
    assert np.abs(dev).max() <= 1e-9


# ------------------------------------------------------------------ resolve
def set_conc(o, r=None, g=None, b=None, a=None):
    c = o.view(oracle.F_CONC)
    for q, x in enumerate((r, g, b, a)):
        c[q] = 0.0 if x is None else x


def resolve(o, dt):
    o.stage(oracle.ST_CONC_PRED, dt)
    o.stage(oracle.ST_CONC_RES, dt)
    return o.get(oracle.F_CONC)


def blob(shape, centre, rad):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d = np.sqrt((i - centre[0]) ** 2 + (j - centre[1]) ** 2 + (k - centre[2]) ** 2)
    return np.clip(1.0 - d / rad, 0.0, None) * 0.3


def test_resolve_lysed_mass_is_diluted_exactly():
    """Still water, no counters, a lysed blob well above the floor zone and
    inside the wall sponge: the predictor copies it (static branch), the
    MacCormack correction and limiter leave it, the 6-neighbour diffusion
    blend conserves its sum, and the dilution exp(-dt mix(0.016, 0.034,
    smoothstep(y0 + 0.045, y0 + 0.26, y))) takes exactly exp(-0.034 dt) of it
    where the blob lies above y0 + 0.26 (a 0.4 m tank)."""
    o = tank(ny=40, depth=0.4)
    g = blob(o.shape, (7.5, 32, 7.5), 3.0)
    set_conc(o, g=g)
    dt = 0.01
    out = resolve(o, dt)
    assert out[1].sum() == pytest.approx(math.exp(-0.034 * dt) * g.sum(), rel=1e-12)
    assert np.all(out[0] == 0) and np.all(out[3] == 0)
    beta = min(dt * 5.6, 0.105)
    assert np.abs(out[1] - g * math.exp(-0.034 * dt)).max() > 0.1 * beta * g.max() * 0.05


def test_resolve_lysed_counter_split_near_the_floor():
    """The lysed counter's mass dm: above the capture band (y > y0 + 0.058)
    it all joins the lysed channel; low in it (y < y0 + 0.018, capture 1)
    0.28 dm stays lysed and 0.68 dm (times the bed's row fade
    1 - 0.38 smoothstep(3.5, 7, row)) enters the floor bed, which the intact
    channel also holds (advect.ts:229-233, 263-284)."""
    o = tank(ny=24, depth=0.12)  # 5 mm rows
    set_conc(o)
    o.parcels_set(np.zeros((1, 8)))
    o.stage(oracle.ST_PARCELS, 0.0)  # (allocates and clears the counters)
    cnt = np.zeros(o.shape, dtype=np.uint32)
    cnt[7, 20, 8] = 3000   # y0 + 102.5 mm: capture 0
    cnt[7, 2, 8] = 5000    # y0 + 12.5 mm: capture 1, floor row 2
    np.ctypeslib.as_array(oracle.lib().orc_splat(o.h, 1), shape=(cnt.size,))[:] = cnt.ravel()
    out = resolve(o, 1e-3)
    ss = oracle.smoothstep
    assert out[1][7, 20, 8] == pytest.approx(3000 / S, rel=1e-12)
    assert out[1][7, 2, 8] == pytest.approx(0.28 * 5000 / S, rel=1e-12)
    fade = 1 - 0.38 * ss(3.5, 7.0, 2.0)
    assert out[3][7, 2, 8] == pytest.approx(0.68 * 5000 / S * fade, rel=1e-12)
    assert out[0][7, 2, 8] == pytest.approx(out[3][7, 2, 8], rel=1e-12)  # intact = parcels + bed
    assert out[3][7, 20, 8] == 0.0


def test_resolve_intact_from_the_counters():
    """Still water: a cell whose intact counter holds m and whose neighbours
    hold none gets 0.64 m (the shoulder blend with its empty neighbourhood);
    with a lysed halo present the neighbours load their neighbourhood and
    get 0.18 m / 6; the walls and the top row stay 0."""
    o = tank()
    set_conc(o)
    cnt = np.zeros(o.shape, dtype=np.uint32)
    cnt[8, 12, 8] = 6 * 65536 // 10
    sp = oracle.lib().orc_splat(o.h, 0)
    o.parcels_set(np.zeros((1, 8)))
    np.ctypeslib.as_array(sp, shape=(cnt.size,))[:] = cnt.ravel()
    out = resolve(o, 1e-3)
    m = cnt[8, 12, 8] / S
    assert out[0][8, 12, 8] == pytest.approx(0.64 * m, rel=1e-12)
    assert out[0].sum() == pytest.approx(0.64 * m, rel=1e-12)
    o2 = tank()
    halo = np.full(o2.shape, 1e-3)
    set_conc(o2, g=halo)
    o2.parcels_set(np.zeros((1, 8)))
    np.ctypeslib.as_array(oracle.lib().orc_splat(o2.h, 0), shape=(cnt.size,))[:] = cnt.ravel()
    out2 = resolve(o2, 1e-3)
    assert out2[0][8, 12, 9] == pytest.approx(0.18 * m / 6, rel=1e-12)
    for a in range(4):
        assert np.all(out2[a][:, :, 0] == 0) and np.all(out2[a][:, -1, :] == 0)
        assert np.all(out2[a][0] == 0) and np.all(out2[a][-1] == 0)


def test_resolve_ranges_and_early_outs():
    """Empty input resolves to 0 everywhere; on a random state every channel
    stays finite, the oxygenation/shear channel in [0, 1] and 0 wherever the
    total concentration is at most 1e-5."""
    o = tank()
    set_conc(o)
    o.parcels_set(np.zeros((1, 8)))
    assert np.all(resolve(o, 1e-3) == 0)
    rng = np.random.default_rng(3)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        o.set(f, rng.standard_normal(o.shape) * 0.05)
    o.stage(oracle.ST_F2C)
    set_conc(o, r=rng.random(o.shape) * 0.2, g=rng.random(o.shape) * 0.05,
             a=rng.random(o.shape) * 0.1)
    out = resolve(o, 2e-3)
    assert np.all(np.isfinite(out))
    assert out[2].min() >= 0 and out[2].max() <= 1
    tot = out[0] + out[1]
    assert np.all(out[2][tot <= 1e-5] == 0)
