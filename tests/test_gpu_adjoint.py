"""Adjoint of the MAC step against central finite differences (FP64 device
mode) and against finite differences of the FP64 CPU oracle.

Loss: L = 1/2 ||dye_N - target||^2, target from a run with the blob centre
moved by (+3, -3, +2) cells and w0 by +0.1 m/s.  Parameters: blob centre,
radius, initial downward speed w0 (cell units, m/s) and the blood density
rho_b (kg/m^3).

The loss is only C1 in the blob parameters (smoothstep edge of the blob,
MacCormack limiter switches, trilinear cell crossings), so a central
difference at h = 1e-3 carries kink errors of up to a few 1e-4 relative;
as h -> 0 the differences converge onto the adjoint, which the strict
checks below pin at h = 1e-5 (cell units and m/s).  rho_b moves the loss
by only ~1e-9 of itself per 1e-5 kg/m^3, where single limiter flips show,
so its strict difference uses h = 1e-2 kg/m^3 (the loss is smooth in rho_b
at that scale: the h = 1e-3 difference agrees to ~1e-8).
"""
import copy

import numpy as np
import pytest

import oracle
from mac3d import MacSolver, scenarios
from mac3d.lib import config_for

pytestmark = pytest.mark.gpu
NAMES = ["cx", "cy", "cz", "radius", "w0"]


def strict_h(name):
    return 1e-2 if name == "rho_b" else 1e-5


def _perturbed(blobs):
    b = copy.deepcopy(blobs)
    b[0].cx += 3.0
    b[0].cy -= 3.0
    b[0].cz += 2.0
    b[0].w0 += 0.1
    return b


def _run_dye(w, blobs, cfg_kw):
    with MacSolver(config_for(w, **cfg_kw)) as s:
        s.init_blobs(blobs)
        s.step_nostats(w.dt, w.steps)
        return s.download("DYE")


def _loss(w, blobs, target, cfg_kw):
    d = _run_dye(w, blobs, cfg_kw)
    return 0.5 * float(np.sum((d - target) ** 2))


def _fd(lossfn, w, which, h, cfg_kw):
    def at(delta):
        b = copy.deepcopy(w.blobs)
        kw = dict(cfg_kw)
        if which == "rho_b":
            kw["rho_b"] = 1060.0 + delta
        else:
            setattr(b[0], which, getattr(b[0], which) + delta)
        return lossfn(b, kw)
    return (at(h) - at(-h)) / (2 * h)


def _check(w, K, strict_rel, loose_rel, cfg_kw=None):
    cfg_kw = cfg_kw or {}
    target = _run_dye(w, _perturbed(w.blobs), cfg_kw)
    with MacSolver(config_for(w, **cfg_kw)) as s:
        g = s.adjoint(w.blobs, w.dt, w.steps, target, checkpoint_every=K)
    lref = _loss(w, w.blobs, target, cfg_kw)
    assert abs(g["loss"] - lref) <= 1e-12 * max(lref, 1.0)
    lossfn = lambda b, kw: _loss(w, b, target, kw)  # noqa: E731
    rows = []
    for p, name in enumerate(NAMES + ["rho_b"]):
        ad = g["d_rho_b"] if name == "rho_b" else g["d_blob"][0][p]
        rows.append((name, ad, _fd(lossfn, w, name, strict_h(name), cfg_kw),
                     _fd(lossfn, w, name, 1e-3, cfg_kw)))
    print("\n" + "\n".join(f"{n:7s} adjoint {a:+.10e}  fd(strict) {f5:+.10e} ({abs(a - f5) / abs(f5):.1e})"
                           f"  fd(1e-3) {f3:+.10e} ({abs(a - f3) / abs(f3):.1e})"
                           for n, a, f5, f3 in rows))
    for n, a, f5, f3 in rows:
        assert abs(a - f5) <= strict_rel * abs(f5), (n, a, f5)
        if loose_rel is not None:
            assert abs(a - f3) <= loose_rel * abs(f3), (n, a, f3)
    return g, target


def test_adjoint_tiny_fp64():
    """32x48x32, 6 steps, checkpoints every 4 steps."""
    w = scenarios.scaled(scenarios.workload("tiny", 2), steps=6, precision=2)
    _check(w, 4, strict_rel=1e-6, loose_rel=1e-3)


def test_adjoint_against_oracle_fd():
    """The same gradient from central differences of the FP64 CPU oracle."""
    w = scenarios.scaled(scenarios.workload("tiny", 2), steps=6, precision=2)
    target = _run_dye(w, _perturbed(w.blobs), {})
    with MacSolver(config_for(w)) as s:
        g = s.adjoint(w.blobs, w.dt, w.steps, target, checkpoint_every=3)

    def oloss(b, kw):
        o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, cycles=w.cycles, **kw)
        o.init_blobs(b)
        o.step(w.dt, w.steps)
        d = o.get(oracle.F_DYE)
        o.close()
        return 0.5 * float(np.sum((d - target) ** 2))

    for name in ("cy", "w0", "rho_b"):
        fd = _fd(oloss, w, name, strict_h(name), {})
        ad = g["d_rho_b"] if name == "rho_b" else g["d_blob"][0][NAMES.index(name)]
        assert abs(ad - fd) <= 1e-6 * abs(fd), (name, ad, fd)


def test_adjoint_workload_64x96x64():
    """The adjoint workload: 64x96x64, dx 1 mm, 50 steps of 0.5 ms,
    checkpoints every 10 steps (FP64 device mode).  With the MacCormack
    limiter the loss jumps where a backtrace point changes texel cell (the
    clamp bounds are the 8 texels around it), so only the h -> 0 differences
    are a fair check; the h = 1e-3 ones are printed."""
    w = scenarios.scaled(scenarios.workload("adjoint", 1), precision=2)
    _check(w, 10, strict_rel=1e-5, loose_rel=None)


def test_adjoint_workload_64x96x64_smooth_transport():
    """Same workload with the MacCormack correction off (pure
    semi-Lagrangian dye transport, no limiter switches): central differences
    at h = 1e-3 agree with the adjoint to 1e-4."""
    w = scenarios.scaled(scenarios.workload("adjoint", 1), precision=2)
    _check(w, 10, strict_rel=1e-5, loose_rel=1e-4, cfg_kw={"maccormack": 0.0})


def test_checkpoint_interval_invariant():
    w = scenarios.scaled(scenarios.workload("tiny", 2), steps=7, precision=2)
    target = _run_dye(w, _perturbed(w.blobs), {})
    grads = []
    for K in (1, 3, 7, 20):
        with MacSolver(config_for(w)) as s:
            grads.append(s.adjoint(w.blobs, w.dt, w.steps, target, checkpoint_every=K))
    for g in grads[1:]:
        assert abs(g["d_rho_b"] - grads[0]["d_rho_b"]) <= 1e-10 * abs(grads[0]["d_rho_b"])
        for p in range(5):
            a, b = g["d_blob"][0][p], grads[0]["d_blob"][0][p]
            assert abs(a - b) <= 1e-10 * abs(b)


def test_fp32_adjoint_close_to_fp64():
    """FP32 primal and adjoint against the FP64 gradient on the adjoint
    workload (single-precision storage of 50 steps of state)."""
    w64 = scenarios.scaled(scenarios.workload("adjoint", 1), precision=2)
    w32 = scenarios.scaled(w64, precision=0)
    target = _run_dye(w64, _perturbed(w64.blobs), {})
    with MacSolver(config_for(w64)) as s:
        g64 = s.adjoint(w64.blobs, w64.dt, w64.steps, target, checkpoint_every=10)
    with MacSolver(config_for(w32)) as s:
        g32 = s.adjoint(w32.blobs, w32.dt, w32.steps, target, checkpoint_every=10)
    print("\nfp32 vs fp64:", [(n, g32["d_blob"][0][p], g64["d_blob"][0][p]) for p, n in enumerate(NAMES)],
          g32["d_rho_b"], g64["d_rho_b"])
    for p in range(5):
        a, b = g32["d_blob"][0][p], g64["d_blob"][0][p]
        assert abs(a - b) <= 1e-3 * abs(b), (NAMES[p], a, b)
    assert abs(g32["d_rho_b"] - g64["d_rho_b"]) <= 1e-3 * abs(g64["d_rho_b"])
