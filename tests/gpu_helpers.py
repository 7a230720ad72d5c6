"""Shared set-up of the GPU parity tests: one HIP solver and one FP64 oracle
on the same seeded state."""
import numpy as np

import oracle
from mac3d import MacSolver, scenarios
from mac3d.lib import config_for


def oracle_for(w, **kw):
    # the mixed mode smooths with over-relaxation 1.25 (the device default)
    if w.precision == 1:
        kw.setdefault("omega", 1.25)
    return oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, nu1=w.nu1, nu2=w.nu2,
                         cycles=w.cycles, solve_to_tol=(w.precision == 1), tol=w.ir_tol,
                         max_cycles=30, **kw)


# physical parameters shared by mac3d_config and the oracle's parameters
CFG_KEYS = ("vel_damp", "maccormack", "rho_b", "tau_y", "mu_inf", "g", "alpha_max")


def pair(name="tiny", seed=1, precision=None, warm_steps=2, **over):
    """HIP solver and oracle with identical state; warm_steps oracle steps
    first so the velocity field is not the trivial initial one.  Keywords
    name workload fields, or (CFG_KEYS) physical parameters given to both."""
    w = scenarios.workload(name, seed)
    if precision is not None:
        w = scenarios.scaled(w, precision=precision)
    cfg = {k: over.pop(k) for k in list(over) if k in CFG_KEYS}
    for k, v in over.items():
        setattr(w, k, v)
    o = oracle_for(w, **cfg)
    o.init_blobs(w.blobs)
    if warm_steps:
        o.step(w.dt, warm_steps)
    s = MacSolver(config_for(w, **cfg))
    s.init_blobs(w.blobs)
    copy_state(o, s)
    return w, s, o


FIELDS = ("U", "V", "W", "P", "DYE")


def copy_state(o, s, fields=FIELDS):
    for f in fields:
        s.upload(f, o.get(getattr(oracle, "F_" + f)))
