"""Drop-release fitting through the adjoint (mac3d_adjoint_release,
mac3d_fit_release): the gradient with respect to spawnAerialDrop's
parameters (src/sim/dropKinematics.ts:354-391) against central differences
of the device's own FP64 loss and of the FP64 CPU oracle's, and a fit that
recovers a release from the dye field it leaves.

Workload: 48x64x48 cells of 1 mm, dt 0.5 ms, FP64 device mode; the default
release map (the reference tank's 0.96 m on 48 cells, so a scene cell is
20 mm) with vel_scale 2.5, so the hero drop's 0.22 m/s contact speed moves
its blob ~3 cells over the run.  Loss L = 1/2 ||dye_N - target||^2."""
import numpy as np
import pytest

import oracle
from mac3d import MacSolver, lib, scenarios
from mac3d.lib import config_for

pytestmark = pytest.mark.gpu

TRUE = dict(x=0.05, z=-0.03, radius=0.07, start_y=0.495, vel_y=-0.075)
GUESS = dict(x=0.08, z=-0.05, radius=0.078, start_y=0.495, vel_y=-0.15)


def work(steps=12, precision=2):
    return scenarios.Workload("release", 48, 64, 48, 1e-3, 5e-4, steps, 3, 8, precision)


def rmap(w):
    m = lib.release_default_map(config_for(w))
    m.vel_scale = 2.5
    return m


def run_dye(w, m, rel):
    blobs, _ = lib.release_blobs(m, rel)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(blobs)
        s.step_nostats(w.dt, w.steps)
        return s.download("DYE")


def test_release_gradient_against_central_differences():
    """Every release parameter: the adjoint against the device's FP64 central
    difference at h = 2e-7 m (1e-5 of a scene cell) and 1e-6 m/s, within
    1e-5 relative; and the release gradient is the blob gradient times the
    map's Jacobian."""
    w = work()
    m = rmap(w)
    target = run_dye(w, m, [GUESS])
    with MacSolver(config_for(w)) as s:
        g = s.adjoint_release(m, [TRUE], w.dt, w.steps, target, checkpoint_every=4)
    _, J = lib.release_blobs(m, [TRUE])
    chained = J[0].T @ np.array(g["d_blob"][0])
    assert np.allclose(g["d_release"][0], chained, rtol=1e-13, atol=1e-13 * np.abs(chained).max())

    def loss(rel):
        d = run_dye(w, m, [rel])
        return 0.5 * float(np.sum((d - target) ** 2))

    assert abs(g["loss"] - loss(TRUE)) <= 1e-12 * g["loss"]
    for b, name in enumerate(lib.RELEASE_PARAMS):
        h = 1e-6 if name == "vel_y" else 2e-7
        lo, hi = dict(TRUE), dict(TRUE)
        lo[name] -= h
        hi[name] += h
        fd = (loss(hi) - loss(lo)) / (2 * h)
        ad = g["d_release"][0][b]
        print(f"{name:8s} adjoint {ad:+.10e} fd {fd:+.10e} ({abs(ad - fd) / abs(fd):.1e})")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (name, ad, fd)


def test_release_gradient_against_oracle_differences():
    """The same gradient from central differences of the FP64 CPU oracle's
    loss, on a shorter run (6 steps), within 1e-5 relative: at h = 1e-5 of a
    cell the position and radius differences carry the blob edge's
    smoothstep and the trilinear kinks (the device's own differences above
    show 1e-6 to 8e-6 on those parameters)."""
    w = work(steps=6)
    m = rmap(w)
    target = run_dye(w, m, [GUESS])
    with MacSolver(config_for(w)) as s:
        g = s.adjoint_release(m, [TRUE], w.dt, w.steps, target, checkpoint_every=3)

    def oloss(rel):
        blobs, _ = lib.release_blobs(m, [rel])
        o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, cycles=w.cycles)
        o.init_blobs(blobs)
        o.step(w.dt, w.steps)
        d = o.get(oracle.F_DYE)
        o.close()
        return 0.5 * float(np.sum((d - target) ** 2))

    for name in ("x", "radius", "vel_y"):
        h = 1e-6 if name == "vel_y" else 2e-7
        lo, hi = dict(TRUE), dict(TRUE)
        lo[name] -= h
        hi[name] += h
        fd = (oloss(hi) - oloss(lo)) / (2 * h)
        ad = g["d_release"][0][lib.RELEASE_PARAMS.index(name)]
        print(f"{name:8s} adjoint {ad:+.10e} oracle fd {fd:+.10e} ({abs(ad - fd) / abs(fd):.1e})")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (name, ad, fd)


def test_fit_recovers_the_release():
    """From a guess 1.5 cells off in x, 1 cell in z, 11% off in radius and
    with twice the release speed (16% more contact speed), the fit of x, z,
    radius and vel_y (start_y held) to the target dye of the true release
    gets the blob centre within 0.02 cells, the radius within 0.2% and the
    contact speed within 0.2%, the loss below 1e-6 of its start."""
    w = work()
    m = rmap(w)
    target = run_dye(w, m, [TRUE])
    with MacSolver(config_for(w)) as s:
        fit, rep = s.fit_release(m, [GUESS], w.dt, w.steps, target, max_iters=30)
    print("\nfit", fit[0], rep)
    (bt,), _ = lib.release_blobs(m, [TRUE])
    (bf,), _ = lib.release_blobs(m, fit)
    assert rep["loss"] <= 1e-6 * rep["loss0"]
    assert abs(bf.cx - bt.cx) <= 0.02 and abs(bf.cz - bt.cz) <= 0.02
    assert abs(bf.radius / bt.radius - 1) <= 2e-3
    assert abs(bf.w0 / bt.w0 - 1) <= 2e-3
    assert fit[0]["start_y"] == GUESS["start_y"]


def test_fit_recovers_the_release_fp32():
    """The same fit with the FP32 step and adjoint (FP32 storage of the run's
    states; the gradient carries ~1e-3 of relative noise), against a target
    from the FP64 step: the blob centre within 0.01 cells, radius and contact
    speed within 0.1%, the loss below 1e-6 of its start (measured: 2e-8
    cells, 1e-8, 6e-7 and 5e-14)."""
    w64 = work()
    w32 = work(precision=0)
    m = rmap(w64)
    target = run_dye(w64, m, [TRUE])
    with MacSolver(config_for(w32)) as s:
        fit, rep = s.fit_release(m, [GUESS], w32.dt, w32.steps, target, max_iters=30)
    print("\nfit fp32", fit[0], rep)
    (bt,), _ = lib.release_blobs(m, [TRUE])
    (bf,), _ = lib.release_blobs(m, fit)
    assert rep["loss"] <= 1e-6 * rep["loss0"]
    assert abs(bf.cx - bt.cx) <= 0.01 and abs(bf.cz - bt.cz) <= 0.01
    assert abs(bf.radius / bt.radius - 1) <= 1e-3
    assert abs(bf.w0 / bt.w0 - 1) <= 1e-3


def test_fit_and_release_errors():
    w = work(steps=2)
    m = rmap(w)
    target = np.zeros(w.nx * w.ny * w.nz)
    with MacSolver(config_for(w)) as s:
        with pytest.raises(lib.Mac3dError):  # a parameter of a second release
            s.fit_release(m, [TRUE], w.dt, w.steps, target, mask=1 << 7)
        with pytest.raises(lib.Mac3dError):
            s.adjoint_release(m, [dict(TRUE, radius=float("inf"))], w.dt, w.steps, target)
        with pytest.raises(lib.Mac3dError):
            s.fit_release(m, [TRUE], -w.dt, w.steps, target)
        with pytest.raises(lib.Mac3dError):  # five mask bits per release: at most 12
            s.fit_release(m, [TRUE] * 13, w.dt, w.steps, target)
        # nothing to move: a zero-iteration fit returns the guess and its loss
        fit, rep = s.fit_release(m, [TRUE], w.dt, w.steps, target, max_iters=0)
        assert fit[0] == TRUE and rep["iters"] == 0 and rep["evals"] == 1
