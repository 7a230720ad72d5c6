"""Pins of the oracle's free-surface wave coupling (CPU only).

The oracle restates the reference's wave passes (src/shaders/wave.ts:56-388)
and evalTopDirichletPressure (src/shaders/fluid.ts:16-78).  The reference
holds no numeric test of them, so each part is pinned here against what the
mathematics fixes: the FFT against numpy's; the whole wave step, on a domain
small enough that the radial absorber and border fade never act, against
the closed-form evolution of one Fourier mode under the gravity-capillary
dispersion relation; the drop forcing against the zero integral of its polar
Ricker wavelet and the closed form of its film deposit; the lid pressure
against its hydrostatic closed form and the zero lateral dipole of its
azimuthal modulation; the divergence and projection against p_top moved
between the right-hand side and the ghost.
"""
import math

import numpy as np
import pytest

import oracle


def make(nx=16, ny=16, nz=16, levels=2, dx=1e-3, **kw):
    return oracle.Oracle(nx, ny, nz, dx, levels=levels, **kw)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [16, 64, 512])
def test_fft_is_numpys_unitary_fft(n):
    """orc_fft = unitary DFT (1/sqrt(n) per transform, wave.ts:243-247) with
    sign -1 forward and +1 inverse; a wrong twiddle sign, a bit-reversal
    slip or a lost scale fails it."""
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle.fft(z, -1)
    assert np.allclose(f, np.fft.fft(z) / math.sqrt(n), atol=1e-12 * n)
    assert np.allclose(oracle.fft(z, +1), np.fft.ifft(z) * math.sqrt(n), atol=1e-12 * n)
    assert np.allclose(oracle.fft(f, +1), z, atol=1e-12 * n)


# ---------------------------------------------------------------- propagation
def omega_of(k, depth):
    """gravity-capillary dispersion with the reference's coefficients
    (g_eff = 0.84, sigma/rho = 0.0024, wave.ts:296-298)"""
    return math.sqrt(max((0.84 * k + 0.0024 * k ** 3) * math.tanh(min(k * depth, 12.0)), 1e-8))


@pytest.mark.parametrize("m,eta0,w0", [(3, 0.01, 0.0), (5, 0.0, 0.02), (11, 0.004, -0.01)])
def test_wave_step_evolves_one_mode_in_closed_form(m, eta0, w0):
    """On a 1 m domain (every texel within 0.71 m of the centre, so the
    absorber (r > 0.90) and border fade (r > 1.22) are 1) with no drops, film
    or Marangoni, one step maps eta = a cos(kx), w = b cos(kx) to
    eta' = D (a cos wdt + b sin wdt / w) cos(kx), w' = D (-a w sin wdt + b cos wdt) cos(kx),
    D = exp(-(0.36 + 0.00018 k^2) dt) (below the high-k roll-off): the FFT
    scale, the Hermitian unpacking and repacking, the k = 2 pi m / L map and
    the phase rotation all enter."""
    n, ext, depth, dt = 64, 0.5, 1.28, 0.01
    o = make()
    o.wave_enable(n=n, half_extent=ext, depth=depth)
    L = 2 * ext
    x = (np.arange(n) + 0.5) / n * L - ext
    k = 2 * math.pi * m / L
    mode = np.broadcast_to(np.cos(k * (x + ext)), (n, n))  # varies along x (the fast axis)
    W = o.view(oracle.F_WAVE)
    W[..., 0] = eta0 * mode
    W[..., 1] = w0 * mode
    o.stage(oracle.ST_WAVE, dt)
    om = omega_of(k, depth)
    D = math.exp(-(0.36 + 0.00018 * k * k) * dt)
    eta1 = D * (eta0 * math.cos(om * dt) + w0 * math.sin(om * dt) / om) * mode
    w1 = D * (-eta0 * om * math.sin(om * dt) + w0 * math.cos(om * dt)) * mode
    got = o.get(oracle.F_WAVE)
    assert np.allclose(got[..., 0], eta1, atol=1e-12)
    assert np.allclose(got[..., 1], w1, atol=1e-12)
    # 9-point Laplacian of the mode in the interior: 0.6 (cos kh - 1) eta (n/256)^2
    h = L / n
    lap = 0.6 * (math.cos(k * h) - 1.0) * eta1 * (n / 256.0) ** 2
    assert np.allclose(got[2:-2, 2:-2, 2], lap[2:-2, 2:-2], atol=1e-12)
    assert np.all(got[..., 3] == 0.0)


def test_wave_step_removes_the_mean_and_the_nyquist_mode():
    """DC (volume conservation) and the Nyquist row/column are zeroed in the
    spectrum (wave.ts:275-279): a constant and a checkerboard vanish."""
    n = 32
    o = make()
    o.wave_enable(n=n, half_extent=0.5)
    W = o.view(oracle.F_WAVE)
    ix = np.arange(n)
    W[..., 0] = 0.01 + 0.003 * (-1.0) ** ix[None, :]
    o.stage(oracle.ST_WAVE, 0.01)
    assert np.abs(o.get(oracle.F_WAVE)[..., 0]).max() < 1e-15


# ---------------------------------------------------------------- forcing
def test_drop_kick_and_film_in_closed_form():
    """One drop with emergence -0.5 radii (crater gate open, capillary
    window shut) on a flat surface: the force pass kicks d eta/dt by
    -dt (6.8 bell (r/0.056)(0.72 + 2.65 s)) exp(-q^2)(1 - q^2), a polar Ricker
    wavelet of zero plane integral (wave.ts:102-125), and deposits the film
    0.65 exp(-r^2/rs^2) clamp(0.42 drain), decayed by 0.88^(dt/0.0167)
    (wave.ts:127-143).  The step's result is that kick carried by the
    spectral propagation, written with numpy's FFT over the two real fields
    (the packed form of the reference is pinned by the one-mode test)."""
    n, ext, dt = 128, 0.4, 0.004
    o = make()
    o.wave_enable(n=n, half_extent=ext, depth=1.28)
    rd, drain, vy = 0.04, 0.5, -0.3
    e = -0.5
    o.set_drops([(0.0, 0.30 + e * rd, 0.0, rd, vy, 0.0, drain, 1.0)])
    o.stage(oracle.ST_WAVE, dt)
    W = o.get(oracle.F_WAVE)
    x = (np.arange(n) + 0.5) / n * 2 * ext - ext
    X, Z = np.meshgrid(x, x)
    r2 = X ** 2 + Z ** 2
    rs = max(0.018, 0.6 * rd)
    film = 0.65 * np.exp(-r2 / rs ** 2) * min(max(0.42 * drain, 0.0), 1.0) * 0.88 ** (dt / 0.0166667)
    assert np.allclose(W[..., 3], film, atol=1e-13)

    excav = oracle.smoothstep(-0.78, -0.04, e)
    bell = math.sin(min(max((1.05 - e) / 1.83, 0.0), 1.0) * 3.14159265) * excav
    speed = min(max(-vy, 0.06), 0.65)
    q2 = r2 / min(max(rd * 1.05, 0.024), 0.082) ** 2
    shape = np.exp(-q2) * (1 - q2)
    assert abs(shape.sum()) < 1e-6 * np.abs(shape).sum()  # zero-integral wavelet
    kick = -dt * (6.8 * bell * (rd / 0.056)) * (0.72 + 2.65 * speed) * shape
    L = 2 * ext
    m = np.fft.fftfreq(n, 1.0 / n)
    KX, KZ = np.meshgrid(2 * np.pi * m / L, 2 * np.pi * m / L)
    k = np.sqrt(KX ** 2 + KZ ** 2)
    om = np.sqrt(np.maximum((0.84 * k + 0.0024 * k ** 3) * np.tanh(np.minimum(k * 1.28, 12.0)), 1e-8))
    D = np.exp(-(0.36 + 0.00018 * k ** 2) * dt) * (1.0 - np.vectorize(oracle.smoothstep)(335.0, 465.0, k))
    keep = np.ones((n, n))
    keep[0, 0] = 0.0
    keep[n // 2, :] = keep[:, n // 2] = 0.0
    V = np.fft.fft2(kick)
    eta1 = np.fft.ifft2(keep * D * np.sin(om * dt) / om * V).real
    w1 = np.fft.ifft2(keep * D * np.cos(om * dt) * V).real
    assert np.abs(eta1).max() > 1e-6
    assert np.abs(W[..., 0] - eta1).max() <= 1e-10 * np.abs(eta1).max()
    assert np.abs(W[..., 1] - w1).max() <= 1e-10 * np.abs(w1).max()


# ---------------------------------------------------------------- lid pressure
def test_top_pressure_hydrostatic_closed_form():
    """A uniform surface eta = c (Laplacian 0), no drops: p_top = 2.45 *
    0.155 c dt/h inside the wall taper, where h = 2 half_xz / nx
    (fluid.ts:36-47, 73-76); the taper rises over 0.72..0.96 half_xz."""
    nx = nz = 16
    o = make(nx=nx, nz=nz)
    o.wave_enable(n=32, half_xz=0.48, depth=1.28)
    W = o.view(oracle.F_WAVE)
    W[..., 0] = 0.02
    dt = 0.01
    o.stage(oracle.ST_TOP_PRESSURE, dt)
    p = o.get(oracle.F_PTOP)
    h = 2 * 0.48 / nx
    want = 2.45 * 0.155 * 0.02 * dt / h
    c = (np.arange(nx) + 0.5) / nx * 0.96 - 0.48
    inner = np.abs(c) < 0.72 * 0.48
    assert np.allclose(p[np.ix_(inner, inner)], want, rtol=1e-13)
    assert np.all(p[0, :] < want) and np.all(p[:, 0] < want)  # tapered at the walls


def test_top_pressure_drop_term_has_no_lateral_dipole():
    """The drop's neck overpressure carries the m = 4, 6, 8 azimuthal
    modulation (fluid.ts:57-63), which has no m = 1 part: the pressure field
    around a drop at the tank centre has zero first moment (no lateral push)."""
    nx = nz = 64
    o = make(nx=nx, nz=nz, ny=16)
    o.wave_enable(n=32, half_xz=0.16, depth=1.0)
    o.set_drops([(0.0, 0.25, 0.0, 0.03, -0.3, 0.0, 0.6, 0.77)])
    o.stage(oracle.ST_TOP_PRESSURE, 0.01)
    p = o.get(oracle.F_PTOP)
    c = (np.arange(nx) + 0.5) / nx * 0.32 - 0.16
    X, Z = np.meshgrid(c, c)
    assert p.max() > 0
    mx = (p * X).sum() / np.abs(p).sum()
    mz = (p * Z).sum() / np.abs(p).sum()
    assert abs(mx) < 2e-4 and abs(mz) < 2e-4


def test_divergence_and_projection_carry_the_lid_pressure():
    """p_top moves to the top row's right-hand side (fluid.ts:548-551) and is
    the ghost of the top face's gradient (fluid.ts:896-898): with p = 0,
    D(u) on the top row drops by p_top and V on the top face drops by p_top."""
    o = make(nx=16, ny=16, nz=16)
    rng = np.random.default_rng(3)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        o.set(f, rng.standard_normal(o.shape) * 0.01)
    o.stage(oracle.ST_DIV)
    d0 = o.get(oracle.F_RHS)
    o.wave_enable(n=16, half_xz=0.008, depth=0.016)
    pt = rng.standard_normal((16, 16)) * 1e-3
    o.view(oracle.F_PTOP)[...] = pt
    o.stage(oracle.ST_DIV)
    d1 = o.get(oracle.F_RHS)
    assert np.array_equal(d1[:, :-1, :], d0[:, :-1, :])
    assert np.allclose(d1[:, -1, :], d0[:, -1, :] - pt, atol=1e-15)
    v0 = o.get(oracle.F_V)
    o.set(oracle.F_P, np.zeros(o.shape))
    o.stage(oracle.ST_PROJECT)
    v1 = o.get(oracle.F_V)
    assert np.allclose(v1[:, -1, :], v0[:, -1, :] - pt, atol=1e-15)
    assert np.array_equal(v1[:, :-1, :], v0[:, :-1, :])


# ---------------------------------------------------------------- advection scene terms
def scene_pair(drops=(), wave=None, **geom):
    """two oracles on one random velocity: coupling with the advection's scene
    terms (fluid.ts:198-321) and without them"""
    rng = np.random.default_rng(11)
    out = []
    for scene in (1, 0):
        o = make(nx=16, ny=24, nz=16)
        for f in (oracle.F_U, oracle.F_V, oracle.F_W):
            o.set(f, rng.standard_normal(o.shape) * 0.01 if scene else out[0][1][f])
        g = dict(n=32, half_extent=0.05, half_xz=0.008, depth=0.024, surface_y=0.0, marangoni=0.6)
        g.update(geom)
        o.wave_enable(scene=scene, **g)
        o.set_drops(list(drops))
        if wave is not None:
            o.view(oracle.F_WAVE)[...] = wave
        out.append((o, {f: o.get(f) for f in (oracle.F_U, oracle.F_V, oracle.F_W)}))
    return out[0][0], out[1][0]


def advect_pair(a, b, dt):
    """face->centre and advection on both; the no-scene oracle traces with the
    scene's centre velocity (its top-row stagnation gate, fluid.ts:944-959),
    so the two differ only by the advection's own scene terms"""
    for o in (a, b):
        o.stage(oracle.ST_F2C)
    b.set(oracle.F_CV, a.get(oracle.F_CV))
    for o in (a, b):
        o.stage(oracle.ST_ADVECT, dt)


def sponge_of(o):
    """the sponge multiplier mix(0.84, 1, wallFade) per cell (fluid.ts:312-321)"""
    nz, ny, nx = o.shape
    hx, y0 = 0.008, 0.0 - 0.024
    x = -hx + (np.arange(nx) + 0.5) / nx * 2 * hx
    z = -hx + (np.arange(nz) + 0.5) / nz * 2 * hx
    y = y0 + (np.arange(ny) + 0.5) / ny * 0.024
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    ss = np.vectorize(oracle.smoothstep)
    fade = (1 - ss(hx * 0.68, hx * 0.94, np.sqrt(X ** 2 + Z ** 2))) * ss(y0 + 0.004, y0 + 0.052, Y)
    return 0.84 + (1.0 - 0.84) * fade


def test_scene_sponge_closed_form():
    """No drops and a calm surface: below the top three rows the scene terms
    only multiply each advected face value by the sponge factor."""
    a, b = scene_pair()
    dt = 0.002
    advect_pair(a, b, dt)
    sp = sponge_of(a)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W):
        ga, gb = a.get(f)[:, :-3, :], b.get(f)[:, :-3, :]
        assert np.allclose(ga, gb * sp[:, :-3, :], rtol=1e-14, atol=1e-300)


def test_scene_kinematic_top_and_marangoni_closed_form():
    """A surface moving at d eta/dt = g everywhere with a film rising
    linearly in x (film = f0 + s x): on the top three rows the x-faces gain
    -m 0.026 (1 - smoothstep(0.45, 0.88, film)) s fade dt (the film gradient by
    centred differences of the bilinear samples is exact for a linear film),
    and the top two rows' y-faces blend towards clamp(0.35 g) by
    (1 - smoothstep(0.02, 0.14, |g|)) times 0.26 / 0.72 (fluid.ts:287-310)."""
    n, E = 32, 0.05
    x = (np.arange(n) + 0.5) / n * 2 * E - E
    W = np.zeros((n, n, 4))
    g, f0, s = 0.05, 0.4, 3.0
    W[..., 1] = g
    W[..., 3] = f0 + s * x[None, :]
    a, b = scene_pair(wave=W)
    dt = 0.002
    advect_pair(a, b, dt)
    sp = sponge_of(a)
    nz, ny, nx = a.shape
    hx = 0.008
    cx = -hx + (np.arange(nx) + 0.5) / nx * 2 * hx
    ss = oracle.smoothstep
    ua, ub = a.get(oracle.F_U), b.get(oracle.F_U)
    va, vb = a.get(oracle.F_V), b.get(oracle.F_V)
    for j in range(ny - 3, ny):
        fade = (j - (ny - 3) + 1) / 3.0
        cmc = np.array([1 - ss(0.45, 0.88, f0 + s * xc) for xc in cx])
        want = (ub[:, j, :] + (-0.6 * 0.026 * cmc * s * fade) * dt) * sp[:, j, :]
        assert np.allclose(ua[:, j, :-1], want[:, :-1], rtol=1e-12, atol=1e-15)
    calm = 1 - ss(0.020, 0.14, abs(g))
    tgt = min(max(g * 0.35, -0.045), 0.045)
    for j, lw in ((ny - 2, 0.26), (ny - 1, 0.72)):
        want = (vb[:, j, :] * (1 - calm * lw) + tgt * calm * lw) * sp[:, j, :]
        assert np.allclose(va[:, j, :], want, rtol=1e-12, atol=1e-15)


def test_scene_neck_draft_on_the_drop_axis():
    """The y-face on a drop's axis (r = 0): the neck profile's centre value
    mix(0.68, 0.14, slug), the vertical mask at the face height, the lobe at
    theta = 0 and the flux drive give the draft target and blend of
    fluid.ts:213-245 in closed form."""
    hx, nx = 0.008, 16
    # a drop whose axis is the centre of cell (8, *, 8): cell centres at -hx + (i + 1/2) h
    h = 2 * hx / nx
    x0 = -hx + 8.5 * h
    rd, vy, drain, seed = 0.02, -0.3, 0.4, 0.9
    a, b = scene_pair(drops=[(x0, 0.01, x0, rd, vy, 0.0, drain, seed)])
    dt = 0.002
    advect_pair(a, b, dt)
    sp = sponge_of(a)
    ss = oracle.smoothstep
    nz, ny, _ = a.shape
    slug = ss(0.14, 0.76, drain)
    radial = 0.68 + (0.14 - 0.68) * slug
    lobe = 1 + 0.036 * math.cos(seed * 1.73) + 0.026 * math.cos(-seed * 2.31)
    flux = (min(max(0.045 + drain * 0.22, 0), 0.2) * ss(0, 0.05, drain)) * (1 - slug) + \
        min(max(max(drain, 0) ** 0.66 * 0.84, 0), 1.24) * slug
    impact = min(max(-vy, 0.10), 0.65)
    va, vb = a.get(oracle.F_V), b.get(oracle.F_V)
    for j in range(ny - 6, ny - 3):  # below the kinematic top rows
        fy = -0.024 + (j + 1.0) / ny * 0.024
        mask = radial * ss(-0.058, -0.014, fy)
        shear = 1 + (0.36 - 1) * ss(-0.024, -0.002, fy)
        target = -min(max(0.14 + 0.64 * impact, 0.15), 0.39) * flux * lobe * shear
        blend = min(max(dt * (14 + 8 * slug) * mask, 0), 1)
        want = (vb[8, j, 8] * (1 - blend) + target * blend) * sp[8, j, 8]
        assert va[8, j, 8] == pytest.approx(want, rel=1e-12, abs=1e-15)


def test_scene_flare_along_the_drop_axis():
    """The x-faces on the line z = z_d through a drop's axis (theta = 0 or pi)
    and the z-faces on the line x = x_d (theta = +-pi/2, where cos(6 theta + a)
    = -cos a): the radial flare blends each face towards sign(offset) times
    the flare speed by 0.65 clamp(dt mix(14, 22, slug)
    radial(r / (r_base azim(theta))) vmask) (fluid.ts:247-285), in closed form
    at every distance the neck reaches."""
    hx, nx = 0.008, 16
    h = 2 * hx / nx
    x0 = -hx + 8.5 * h
    rd, vy, drain, seed = 0.02, -0.3, 0.4, 0.9
    a, b = scene_pair(drops=[(x0, 0.01, x0, rd, vy, 0.0, drain, seed)])
    dt = 0.002
    advect_pair(a, b, dt)
    sp = sponge_of(a)
    ss = oracle.smoothstep
    nz, ny, _ = a.shape
    slug = ss(0.14, 0.76, drain)
    rbase = min(max(rd * (0.80 + 0.28 * slug) + 0.012 + 0.018 * slug, 0.018), 0.088)
    azim_x = 1 + 0.036 * math.cos(1.73 * seed) + 0.026 * math.cos(-2.31 * seed) + \
        0.018 * math.sin(0.89 * seed)                 # theta = 0, pi
    azim_z = 1 + 0.036 * math.cos(1.73 * seed) - 0.026 * math.cos(-2.31 * seed) + \
        0.018 * math.sin(0.89 * seed)                 # theta = +-pi/2

    def radial(rr):
        sl = (0.14 + min(rr * rr, 1.20)) * (1 - ss(0.78, 1.15, rr))
        st = (0.68 + 0.32 * min(rr * rr, 1.0)) * (1 - ss(0.74, 1.10, rr))
        return st + (sl - st) * slug
    lines = ((oracle.F_U, azim_x, lambda j, q: (8, j, q)),   # x-face of cell (q, j, 8)
             (oracle.F_W, azim_z, lambda j, q: (q, j, 8)))   # z-face of cell (8, j, q)
    for f, azim, at in lines:
        fa, fb = a.get(f), b.get(f)
        moved = 0
        for j in range(ny - 3):  # below the Marangoni and kinematic-top rows
            cy = -0.024 + (j + 0.5) / ny * 0.024
            vmask = ss(-0.058, -0.014, cy)
            f0 = 0.104 * (0.35 + 0.65 * slug)
            flare = (f0 + (-0.020 - f0) * ss(-0.066, -0.012, cy)) * min(max(drain, 0), 1.3)
            for q in range(nx - 1):
                off = (q - 7.5) * h  # face offset from the axis along the line
                rr = abs(off) / max(rbase * azim, 1e-4)
                want = fb[at(j, q)]
                if vmask > 0.001 and abs(off) < rbase * 1.34 and rr < 1.22:
                    bf = min(max(dt * (14 + 8 * slug) * radial(rr) * vmask, 0), 1)
                    want = want * (1 - 0.65 * bf) + 0.65 * bf * math.copysign(1.0, off) * flare
                    moved += bf > 0
                got = fa[at(j, q)]
                assert got == pytest.approx(want * sp[at(j, q)], rel=1e-12, abs=1e-15), (f, j, q)
        assert moved > 50


# ---------------------------------------------------------------- Casson scene terms
def casson_scene_oracle(scene=1, nx=16, ny=24, nz=16, geom=None, **clock):
    """an oracle with the wave coupling on a deep tank (the unfurling gate is
    open below 0.36 m under the surface), zero dye and velocity"""
    o = make(nx=nx, ny=ny, nz=nz)
    g = dict(n=32, half_extent=0.5, half_xz=0.1, depth=0.6, surface_y=0.0, marangoni=0.0)
    g.update(geom or {})
    o.wave_enable(scene=scene, **g, **clock)
    for f in (oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_DYE):
        o.set(f, np.zeros(o.shape))
    return o


def test_curl_of_a_rigid_rotation_is_twice_the_rate():
    """Faces sampled from the rigid rotation v = Omega x r (grid units) give
    centre velocities on the same rotation away from the walls, whose
    central-difference curl is exactly 2 Omega with |omega| = 2 |Omega|
    (FLUID_CURL, fluid.ts:331-369)."""
    o = casson_scene_oracle()
    nz, ny, nx = o.shape
    Om = np.array([0.03, -0.02, 0.05])
    c0 = np.array([nx / 2.0, ny / 2.0, nz / 2.0])
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")

    def rot(x, y, z):
        r = np.stack([x - c0[0], y - c0[1], z - c0[2]], axis=-1)
        return np.cross(Om, r)
    o.set(oracle.F_U, rot(i + 1.0, j + 0.5, k + 0.5)[..., 0])  # face positions
    o.set(oracle.F_V, rot(i + 0.5, j + 1.0, k + 0.5)[..., 1])
    o.set(oracle.F_W, rot(i + 0.5, j + 0.5, k + 1.0)[..., 2])
    o.stage(oracle.ST_CASSON, 0.002)
    cu = o.get(oracle.F_CURL)[2:-2, 2:-2, 2:-2]
    assert np.allclose(cu[..., :3], 2 * Om, rtol=0, atol=1e-13)
    assert np.allclose(cu[..., 3], 2 * np.linalg.norm(Om), rtol=0, atol=1e-13)


def test_curl_clamps_its_neighbours_at_the_walls():
    """v_y = a x (grid units) on every y-face: omega = (0, 0, a) inside and
    a / 2 on the two x-wall columns, where the clamped neighbour turns the
    central difference one-sided (loadVelC, fluid.ts:338-343)."""
    o = casson_scene_oracle()
    nz, ny, nx = o.shape
    a = 0.07
    o.set(oracle.F_V, np.broadcast_to(a * (np.arange(nx) + 0.5), o.shape).copy())
    o.stage(oracle.ST_CASSON, 0.002)
    cu = o.get(oracle.F_CURL)[:, 1:, :]  # (the floor row's centre sees the solid face)
    want = np.full(nx, a)
    want[0] = want[-1] = a / 2
    assert np.allclose(cu[..., 2], want, rtol=0, atol=1e-14)
    assert np.allclose(cu[..., :2], 0.0, atol=1e-14)


def test_vorticity_confinement_closed_form():
    """v_y = a xc^2 - 1.5 with xc the cell's x offset from the middle:
    omega = (0, 0, 2 a xc), N = (sign xc, 0, 0) and the confinement pushes
    each y-face by -dt scale gate(|omega|) |omega| with
    gate(m) = smoothstep(0.18, 0.95, m) / (1 + 0.28 m) (evalVortConf,
    fluid.ts:394-423), averaged with the cell above (:470), so the rows
    next to the inactive ones take half; nothing on the x- and z-faces.  The scene's other terms are off (no dye, no clamp reached)."""
    dt, scale, a = 0.004, 1.3, 0.05
    got = []
    for scene in (1, 0):
        o = casson_scene_oracle(scene=scene, vort_conf=scale)
        nz, ny, nx = o.shape
        xc = np.arange(nx) + 0.5 - nx / 2
        o.set(oracle.F_V, np.broadcast_to(a * xc ** 2 - 1.5, o.shape).copy())
        o.stage(oracle.ST_CASSON, dt)
        got.append([o.get(f) for f in (oracle.F_U, oracle.F_V, oracle.F_W)])
    du, dv, dw = (s - n for s, n in zip(*got))
    assert np.all(du == 0) and np.all(dw == 0)
    m = 2 * a * np.abs(xc)
    f = np.array([-scale * oracle.smoothstep(0.18, 0.95, x) / (1 + 0.28 * x) * x
                  if 2 <= i <= nx - 3 else 0.0 for i, x in enumerate(m)])
    assert np.count_nonzero(f) >= 8
    zin = slice(2, nz - 2)
    assert np.allclose(dv[zin, 2:ny - 4, :], dt * f, rtol=1e-12, atol=1e-16)
    for j in (1, ny - 4):  # one of the two cells of the face is inactive
        assert np.allclose(dv[zin, j, :], 0.5 * dt * f, rtol=1e-12, atol=1e-16)
    assert np.all(dv[:, 0, :] == 0) and np.all(dv[:, ny - 3:, :] == 0)
    assert np.all(dv[:2] == 0) and np.all(dv[nz - 2:] == 0)


def _unfurl_increment(**clock):
    o = casson_scene_oracle(nx=32, ny=48, nz=32, geom=dict(half_xz=0.2), **clock)
    o.set(oracle.F_DYE, np.ones(o.shape))
    ref = casson_scene_oracle(scene=0, nx=32, ny=48, nz=32, geom=dict(half_xz=0.2))
    ref.set(oracle.F_DYE, np.ones(ref.shape))
    for s in (o, ref):
        s.stage(oracle.ST_CASSON, 0.01)
    return [o.get(f) - ref.get(f) for f in (oracle.F_U, oracle.F_V, oracle.F_W)]


def test_unfurling_forcing_is_solenoidal():
    """Deep in the column (gate and vertical modulation 1) the unfurling
    forcing is the sum of three transverse plane waves (k_i . d_i = 0,
    fluid.ts:488-503), so its central-difference divergence at the cell
    centres vanishes to O((h k)^2) against each term's size; a wrong sign or
    component in any polarisation leaves an O(1) divergence."""
    du, dv, dw = _unfurl_increment(sim_time=0.7)
    nz, ny, nx = du.shape
    hx, hy = 0.4 / nx, 0.6 / ny
    # centres 0.43 m or more below the surface, two cells from the x/z walls
    # (whose last faces are zeroed)
    jt = max(j for j in range(ny) if 0.6 - (j + 0.5) * hy >= 0.43)
    K, J, I = slice(2, nz - 2), slice(1, jt), slice(2, nx - 2)

    def cdiff(f, axis, h):
        return (np.roll(f, -1, axis) - np.roll(f, 1, axis)) / (2 * h)
    terms = [cdiff(du, 2, hx)[K, J, I], cdiff(dv, 1, hy)[K, J, I], cdiff(dw, 0, hx)[K, J, I]]
    div = sum(terms)
    size = np.sqrt(sum(np.mean(t ** 2) for t in terms))
    assert size > 0
    assert np.sqrt(np.mean(div ** 2)) <= 0.01 * size


def test_unfurling_phase_takes_time_and_schlieren_together():
    """The forcing's phase is 0.36 simTime + 1.7 schlierenStr (fluid.ts:492):
    a schlieren offset s equals a time shift of 1.7 s / 0.36."""
    a = _unfurl_increment(sim_time=0.5, schlieren=0.2)
    b = _unfurl_increment(sim_time=0.5 + 1.7 * 0.2 / 0.36, schlieren=0.0)
    c = _unfurl_increment(sim_time=0.5, schlieren=0.0)
    for x, y, z in zip(a, b, c):
        assert np.allclose(x, y, rtol=1e-9, atol=1e-15)
        assert not np.allclose(x, z, rtol=1e-3, atol=0)


def test_casson_scene_clamps_and_keeps_walls():
    """Faces beyond 1.8 m/s come out at +-1.8 (fluid.ts:505) and the wall
    faces at 0 (:506-507); without the scene the clamp is not applied."""
    for scene in (1, 0):
        o = casson_scene_oracle(scene=scene)
        o.set(oracle.F_U, np.full(o.shape, 2.5))
        o.set(oracle.F_V, np.full(o.shape, -3.0))
        o.stage(oracle.ST_CASSON, 0.002)
        u, v = o.get(oracle.F_U), o.get(oracle.F_V)
        assert np.all(u[:, :, -1] == 0)
        if scene:
            assert u[:, :, :-1].max() == 1.8 and v.min() == -1.8
        else:
            assert u.max() > 1.8 and v.min() < -1.8


def test_scene_top_row_stagnation_gate():
    """FLUID_MAC_TO_CENTER's top-row gate (fluid.ts:944-959): a downward centre
    velocity in the top two rows is scaled by mix(0.45, 1, smoothstep(0.002,
    0.018, depth below the surface)) (0.50 and 0.95 for 10 mm rows), except
    inside a draining drop's footprint (within 0.78 ra of its axis, drain past
    0.065: factor 1); upward centres and lower rows keep 0.5 (vD + vU)."""
    ss = oracle.smoothstep
    o = casson_scene_oracle(geom=dict(depth=0.24, half_xz=0.1))
    nz, ny, nx = o.shape
    V = np.full(o.shape, -0.1)
    V[:, :, :4] = 0.1  # upward in the first four columns
    o.set(oracle.F_V, V)
    hx, h = 0.1, 0.2 / nx
    xd = -hx + 12.5 * h  # a drop over cell column (12, 12)
    rd = 0.01
    o.set_drops([(xd, 0.0, xd, rd, -0.3, 0.0, 0.1, 1.2)])
    o.stage(oracle.ST_F2C)
    cv = o.get(oracle.F_CV)
    ra = min(max(rd * 1.15 + 0.032, 0.046), 0.092)
    x = -hx + (np.arange(nx) + 0.5) * h
    for j in range(1, ny):
        depth_below = -(-0.24 + (j + 0.5) * 0.01)
        stag = 0.45 + 0.55 * ss(0.002, 0.018, depth_below)
        for k in range(nz):
            for i in range(nx):
                want = V[k, j, i]
                if j >= ny - 2 and want < 0:
                    r = math.hypot(x[i] - xd, x[k] - xd)
                    act = (1 - ss(0.78 * ra, 1.20 * ra, r)) * ss(0.008, 0.065, 0.1)
                    want *= stag + (1 - stag) * act
                assert cv[k, j, i] == pytest.approx(want, rel=1e-14, abs=1e-16), (k, j, i)
    top = cv[:, ny - 1, 4:]
    # (negative values: the least held back is the minimum)
    assert np.isclose(top.min(), -0.1)
    assert np.isclose(top.max(), -0.1 * (0.45 + 0.55 * ss(0.002, 0.018, 0.005)))


def test_scene_neck_draft_on_the_diagonals():
    """The neck draft of the y-faces on the two diagonals through a drop's
    axis, theta = pi/4 (5 pi/4) and -pi/4 (3 pi/4), where the azimuthal
    modulation takes closed values: cos(4 theta + a) = -cos a,
    cos(6 theta - b) = -+sin b, sin(8 theta + c) = sin c, so
      azim  = 1 - 0.036 cos(1.73 s) -+ 0.026 sin(2.31 s) + 0.018 sin(0.89 s)
      lobe  = 1 - 0.036 cos(1.73 s) -+ 0.026 sin(2.31 s)
    (- on the main diagonal, + on the other).  With r = sqrt(2) |q| h the
    draft target, mask and blend of fluid.ts:198-245 follow as on the axis.
    A wrong multiple of theta in any of the three modes, or a swapped sine
    and cosine, changes azim or lobe on one of the diagonals."""
    hx, nx = 0.008, 16
    h = 2 * hx / nx
    x0 = -hx + 8.5 * h
    rd, vy, drain, seed = 0.02, -0.3, 0.4, 0.9
    a, b = scene_pair(drops=[(x0, 0.01, x0, rd, vy, 0.0, drain, seed)])
    dt = 0.002
    advect_pair(a, b, dt)
    sp = sponge_of(a)
    ss = oracle.smoothstep
    nz, ny, _ = a.shape
    slug = ss(0.14, 0.76, drain)
    rbase = min(max(rd * (0.80 + 0.28 * slug) + 0.012 + 0.018 * slug, 0.018), 0.088)
    flux = (min(max(0.045 + drain * 0.22, 0), 0.2) * ss(0, 0.05, drain)) * (1 - slug) + \
        min(max(max(drain, 0) ** 0.66 * 0.84, 0), 1.24) * slug
    impact = min(max(-vy, 0.10), 0.65)

    def radial(rr):
        sl = (0.14 + min(rr * rr, 1.20)) * (1 - ss(0.78, 1.15, rr))
        st = (0.68 + 0.32 * min(rr * rr, 1.0)) * (1 - ss(0.74, 1.10, rr))
        return st + (sl - st) * slug
    va, vb = a.get(oracle.F_V), b.get(oracle.F_V)
    checked = 0
    for sgn, cell in ((-1.0, lambda q: (8 + q, 8 + q)), (1.0, lambda q: (8 - q, 8 + q))):
        azim = 1 - 0.036 * math.cos(1.73 * seed) + sgn * 0.026 * math.sin(2.31 * seed) + \
            0.018 * math.sin(0.89 * seed)
        lobe = 1 - 0.036 * math.cos(1.73 * seed) + sgn * 0.026 * math.sin(2.31 * seed)
        for q in range(-7, 8):
            if q == 0:
                continue
            k, i = cell(q)
            r = math.sqrt(2.0) * abs(q) * h
            rad = radial(r / max(rbase * azim, 1e-4))
            for j in range(ny - 6, ny - 3):
                fy = -0.024 + (j + 1.0) / ny * 0.024
                mask = rad * ss(-0.058, -0.014, fy)
                want = vb[k, j, i]
                if mask > 0.001:
                    shear = 1 + (0.36 - 1) * ss(-0.024, -0.002, fy)
                    target = -min(max(0.14 + 0.64 * impact, 0.15), 0.39) * flux * lobe * shear
                    blend = min(max(dt * (14 + 8 * slug) * mask, 0), 1)
                    want = want * (1 - blend) + target * blend
                    checked += 1
                assert va[k, j, i] == pytest.approx(want * sp[k, j, i], rel=1e-12, abs=1e-15), \
                    (sgn, q, j)
    assert checked > 40
