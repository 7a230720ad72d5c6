#!/usr/bin/env python3
"""Full-size accuracy and large-case runs (not collected by pytest).

  python tests/report_workloads.py accuracy [--steps 200] [--every 20]
      192x256x192 workload, plain FP32 and mixed precision on the device,
      each against the FP64 oracle (the mixed one solving to the same
      tolerance with the same smoother): per-field relative L2 of u, v, w,
      p, dye and the device's FP64 max-divergence, printed every `every`
      steps.
  python tests/report_workloads.py large [--steps 200]
      256x384x256, 5-level V-cycle: FP32 and mixed steps/s and max-div, then
      the adjoint over the run with sqrt(N) checkpointing, and a central
      difference in rho_b (h = 1e-2) against it.
  python tests/report_workloads.py rho_b [--steps 200]
      the large case's dL/d rho_b (FP64) against central differences at
      h = 1e-1, 1e-2, 1e-3, with the MacCormack correction on and off.

Prints one JSON object per line.
"""
import argparse
import copy
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "drop-of-blood_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from mac3d import MacSolver, scenarios  # noqa: E402
from mac3d.lib import config_for  # noqa: E402


def emit(**kw):
    print(json.dumps(kw), flush=True)


def rel(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


def accuracy(steps, every):
    threads = oracle.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0))
    for name in ("ref", "mixed"):
        w = scenarios.workload(name, 1)
        kw = dict(solve_to_tol=True, tol=w.ir_tol, max_cycles=30, omega=1.2) if w.precision == 1 \
            else dict(cycles=w.cycles)
        o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, **kw)
        o.init_blobs(w.blobs)
        with MacSolver(config_for(w)) as s:
            s.init_blobs(w.blobs)
            done = 0
            t_cpu = 0.0
            while done < steps:
                n = min(every, steps - done)
                st = s.step(w.dt, n)
                t0 = time.perf_counter()
                o.step(w.dt, n)
                t_cpu += time.perf_counter() - t0
                done += n
                errs = {f: rel(s.download(f.upper()), o.get(getattr(oracle, "F_" + f.upper())))
                        for f in ("u", "v", "w", "p", "dye")}
                omax, _ = o.div_norms()
                emit(workload=name, step=done, rel_l2=errs, device_max_div=st["max_div"],
                     oracle_max_div=omax, device_vcycles_last_call=st["vcycles"],
                     oracle_threads=threads, oracle_s_per_step=t_cpu / done)


def large(steps):
    out = {}
    for prec in (0, 1):
        w = scenarios.scaled(scenarios.workload("large", 1), precision=prec)
        with MacSolver(config_for(w)) as s:
            s.init_blobs(w.blobs)
            s.step_nostats(w.dt, 3)
            t0 = time.perf_counter()
            st = s.step(w.dt, 40)
            dt = time.perf_counter() - t0
            emit(workload="large", precision=prec, steps_per_s=40 / dt, max_div=st["max_div"],
                 vcycles_per_step=st["vcycles"] / 40, rel_res=st["rel_res"])
            out[prec] = 40 / dt
    emit(workload="large", mixed_speedup=out[1] / out[0])
    # adjoint over `steps` steps with sqrt(N) checkpoints: FP32 primal for the
    # time, FP64 for the check against a central difference in rho_b (an FP32
    # difference of a 200-step loss at a 1e-5 relative step is rounding noise)
    w32 = scenarios.scaled(scenarios.workload("large", 1), precision=0, steps=steps)
    K = max(1, int(round(math.sqrt(steps))))
    with MacSolver(config_for(w32)) as s:
        s.init_blobs(w32.blobs)
        s.step_nostats(w32.dt, steps)
        t32 = s.download("DYE")
    with MacSolver(config_for(w32)) as s:
        t0 = time.perf_counter()
        s.adjoint(w32.blobs, w32.dt, steps, t32 * 0.5, checkpoint_every=K)
        emit(workload="large", adjoint_fp32_seconds=time.perf_counter() - t0, steps=steps,
             checkpoint_every=K)
    w = scenarios.scaled(w32, precision=2)
    tb = copy.deepcopy(w.blobs)
    tb[0].cx += 3.0
    tb[0].w0 += 0.1
    with MacSolver(config_for(w)) as s:
        s.init_blobs(tb)
        s.step_nostats(w.dt, steps)
        target = s.download("DYE")
    with MacSolver(config_for(w)) as s:
        t0 = time.perf_counter()
        g = s.adjoint(w.blobs, w.dt, steps, target, checkpoint_every=K)
        ta = time.perf_counter() - t0

    def loss(rho_b):
        with MacSolver(config_for(w, rho_b=rho_b)) as s:
            s.init_blobs(w.blobs)
            s.step_nostats(w.dt, steps)
            d = s.download("DYE")
        return 0.5 * float(np.sum((d.astype(np.float64) - target) ** 2))
    fd = (loss(1060.0 + 1e-2) - loss(1060.0 - 1e-2)) / 2e-2
    emit(workload="large", adjoint_steps=steps, checkpoint_every=K, adjoint_seconds=ta,
         loss=g["loss"], d_rho_b=g["d_rho_b"], fd_rho_b=fd,
         rel_err_rho_b=abs(g["d_rho_b"] - fd) / abs(fd), d_blob0=g["d_blob"][0])


def rho_b_differences(steps):
    """dL/d rho_b of the large case (FP64, sqrt(N) checkpoints) against
    central differences at h = 1e-1, 1e-2 and 1e-3 kg/m^3, with the
    MacCormack correction on and off."""
    w = scenarios.scaled(scenarios.workload("large", 1), precision=2, steps=steps)
    K = max(1, int(round(math.sqrt(steps))))
    tb = copy.deepcopy(w.blobs)
    tb[0].cx += 3.0
    tb[0].w0 += 0.1
    for mc in (None, 0.0):
        kw = {} if mc is None else {"maccormack": mc}
        with MacSolver(config_for(w, **kw)) as s:
            s.init_blobs(tb)
            s.step_nostats(w.dt, steps)
            target = s.download("DYE")
        with MacSolver(config_for(w, **kw)) as s:
            g = s.adjoint(w.blobs, w.dt, steps, target, checkpoint_every=K)

        def loss(rho_b):
            with MacSolver(config_for(w, rho_b=rho_b, **kw)) as s:
                s.init_blobs(w.blobs)
                s.step_nostats(w.dt, steps)
                d = s.download("DYE")
            return 0.5 * float(np.sum((d.astype(np.float64) - target) ** 2))
        l0 = loss(1060.0)
        for h in (1e-1, 1e-2, 1e-3):
            lp, lm = loss(1060.0 + h), loss(1060.0 - h)
            fd = (lp - lm) / (2 * h)
            emit(workload="large", steps=steps, maccormack="off" if mc == 0.0 else "on", h=h,
                 d_rho_b=g["d_rho_b"], fd_rho_b=fd, rel_err=abs(g["d_rho_b"] - fd) / abs(fd),
                 curvature=(lp - 2 * l0 + lm) / (h * h), loss=l0, adjoint_loss=g["loss"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("what", choices=["accuracy", "large", "rho_b"])
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--every", type=int, default=20)
    a = ap.parse_args()
    if a.what == "accuracy":
        accuracy(a.steps, a.every)
    elif a.what == "rho_b":
        rho_b_differences(a.steps)
    else:
        large(a.steps)


if __name__ == "__main__":
    main()
