#!/usr/bin/env python3
"""Benchmark of the MAC step on one MI355X.

Workload: 192x256x192 cells, dx 0.5 mm, dt 0.25 ms, three seeded blood blobs
(radius 7+-1 cells, w in [0.3, 0.8] m/s), 3-level V-cycle with 2+2 red-black
sweeps and 8 cycles per step, plain FP32.  `value` is FP32 steps/s with all
fields resident in HBM.  The same line carries the mixed-precision run
(FP64 refinement to ||r||/||b|| <= 1e-8, FP16 level>=2 storage) and its
speed-up, the smoother's roofline figures (HIP events on the solver's
stream), the FP32 step with the free-surface wave coupling on (a 512 x 512
wave grid, two active drops), the same with the reference scene's advection
and Casson terms, the parcel stage, the reference's whole scene at its own
shape (96x128x96 MAC grid, 192x256x192 concentration, 393,216 parcels), and
the CPU baseline timed on this host: the FP64
JavaScript transcription of the step (Node worker threads) and the FP64 C
oracle (all threads and one).

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "drop-of-blood_amd"))

HBM_PEAK_GBPS = 8000.0
# bytes per level-0 cell: the level-0 smoothing kernel runs two red-black
# sweeps per launch and must move 12 B per cell (read p and rhs once, write p
# once), the roofline's algorithmic bytes; the two-colour-pass algorithm
# moves 24 B per cell per sweep (read p and rhs, write p, per colour pass),
# reported as a sweep-equivalent rate
SWEEP_BYTES_PER_CELL = 24.0
SWEEPS_PER_LAUNCH = 2
KERNEL_BYTES_PER_CELL = 12.0
# algorithmic bytes per wave texel of one wave step: force pass (read the
# state 16, write eta + i deta/dt 8 and the film 4), four FFT passes (8 in,
# 8 out each), propagation (8 in, 8 out), finalize (8 + 4 in, 16 out)
WAVE_BYTES_PER_TEXEL = 28.0 + 4 * 16.0 + 16.0 + 28.0
WAVE_N = 512
# two drops entering the surface at the tank centre (x, y, z, radius, vel_y,
# osc, drain, seed; metres and m/s, surface at y = 0.30)
WAVE_DROPS = [(0.006, 0.295, -0.004, 0.020, -0.4, 0.0, 0.55, 1.3),
              (-0.012, 0.305, 0.010, 0.018, -0.2, 0.0, 0.20, 2.1)]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PROFILE_STEPS = 3  # steps timed kernel by kernel (HIP events) after the timed run


def run_gpu(workload, steps, warmup, profile_steps=PROFILE_STEPS, wave=False, scene=False):
    from mac3d import MacSolver
    from mac3d.lib import config_for
    cfg = config_for(workload)
    with MacSolver(cfg) as s:
        s.init_blobs(workload.blobs)
        if wave:
            s.wave_enable(n=WAVE_N, marangoni=0.8, scene=int(scene))
            s.set_drops(WAVE_DROPS)
        if warmup:
            s.step_nostats(workload.dt, warmup)
        t0 = time.perf_counter()
        st = s.step(workload.dt, steps)  # blocks until the stream has drained
        wall = time.perf_counter() - t0
        # stats add one divergence-norm pass after the timed kernels; the
        # kernel span of the K steps is st["ms"]
        prof = {}
        if profile_steps:
            s.profile(True)
            s.step_nostats(workload.dt, profile_steps)
            from mac3d.lib import KC
            for name in KC:
                ms, n = s.profile_read(name)
                if n:
                    prof[name] = (ms, n)
            s.profile(False)
        return st, wall, prof


PARCELS = 393216  # the reference's persistent parcel count (PARTICLE_COUNT)
# parcel state read and written once per step (two float4 each way)
PARCEL_BYTES = 64.0


def run_parcels(workload, steps, order):
    """The parcel stage on the scene-term set-up: PARCELS parcels in a cloud
    under the two drops, advanced and splatted `steps` times through the
    flow after two MAC steps; HIP-event time of the stage per call.
    order "cell": parcels stored in the order of their cells (neighbours in
    memory sample neighbouring cells, as parcels emitted together do);
    "random": uniformly shuffled (every lane of a wave gathers elsewhere)."""
    import numpy as np
    from mac3d import MacSolver
    from mac3d.lib import config_for
    with MacSolver(config_for(workload)) as s:
        s.init_blobs(workload.blobs)
        s.wave_enable(n=WAVE_N, marangoni=0.8, scene=1)
        s.set_drops(WAVE_DROPS)
        s.step_nostats(workload.dt, 2)
        hx, top = 0.5 * workload.nx * workload.dx, 0.30
        depth = workload.ny * workload.dx
        rng = np.random.default_rng(5)
        P = np.zeros((PARCELS, 8), dtype=np.float32)
        r, th = rng.uniform(0, 0.5 * hx, PARCELS), rng.uniform(0, 2 * np.pi, PARCELS)
        P[:, 0], P[:, 2] = r * np.cos(th), r * np.sin(th)
        P[:, 1] = rng.uniform(top - 0.8 * depth, top - 0.004, PARCELS)
        P[:, 3] = 0.134
        d = rng.standard_normal((PARCELS, 3))
        P[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True) * \
            np.where(rng.random(PARCELS) < 0.4, 0.010, 0.002)[:, None]
        P[:, 7] = rng.integers(0, 3000, PARCELS)
        if order == "cell":
            ix = ((P[:, 0] + hx) / workload.dx).astype(np.int64)
            iy = ((P[:, 1] - (top - depth)) / workload.dx).astype(np.int64)
            iz = ((P[:, 2] + hx) / workload.dx).astype(np.int64)
            P = P[np.argsort((iz * workload.ny + iy) * workload.nx + ix, kind="stable")]
        s.parcels_upload(P)
        s.stage("PARCELS", workload.dt)
        s.profile(True)
        for _ in range(steps):
            s.stage("PARCELS", workload.dt)
        ms, n = s.profile_read("PARCELS")
        s.profile(False)
        live = int((s.parcels_download()[:, 3] > 0).sum())
        return ms / max(n, 1), live


# The reference's own scene at its shape: VEL_GRID 96x128x96 with 10 mm
# cells in the 0.96 x 1.28 x 0.96 m box under the surface at y = 0.30, dt
# 1/60 s, velDamp 0.9993, the 512-texel wave, the scene terms with the
# reference's physics, the 393,216-parcel ring with 2000 emissions per step
# from two draining drops, and the 4-channel concentration on CONC_GRID
# 192x256x192 (twice the MAC grid per axis)
REF_SCENE_DROPS = [(0.03, 0.285, -0.02, 0.024, -0.35, 0.0, 0.62, 1.3),
                   (-0.11, 0.31, 0.07, 0.018, -0.9, 0.0, 0.21, 2.4)]
REF_SCENE_EMIT = 2000


def run_reference_scene(steps, warmup):
    """steps/s and ms/step of the whole physics step of the reference's
    scene (pass A and pass B) from a quiescent start, after `warmup` steps;
    the emission request goes to the device before each step"""
    import numpy as np
    from mac3d import MacSolver, scenarios
    from mac3d.lib import config_for
    w = scenarios.scaled(scenarios.workload("tiny", 1), nx=96, ny=128, nz=96, dx=0.01,
                         dt=1.0 / 60.0, levels=3, blobs=[])
    with MacSolver(config_for(w, vel_damp=0.9993)) as s:
        s.init_blobs([])
        s.wave_enable(n=WAVE_N, half_xz=0.48, depth=1.28, surface_y=0.30, marangoni=1.0,
                      scene=1, physics=1)
        s.set_drops(REF_SCENE_DROPS)
        s.conc_enable(True, refine=2)
        s.parcels_upload(np.zeros((PARCELS, 8), dtype=np.float32))
        q = 0
        for _ in range(warmup):
            s.parcels_emit(REF_SCENE_EMIT * q, REF_SCENE_EMIT)
            s.step_nostats(w.dt, 1)
            q += 1
        t0 = time.perf_counter()
        for _ in range(steps):
            s.parcels_emit(REF_SCENE_EMIT * q, REF_SCENE_EMIT)
            s.step_nostats(w.dt, 1)
            q += 1
        wall = time.perf_counter() - t0
        live = int((s.parcels_download()[:, 3] > 0).sum())
        return steps / wall, live


def copy_probe(workload, times=5):
    """k_convert<float, float> over one field: reads and writes 4 B per cell."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    import numpy as np
    with MacSolver(config_for(workload)) as s:
        out = np.empty((workload.nz, workload.ny, workload.nx), dtype=np.float32)
        for _ in range(times):
            s.download_into("U", out)


def smoother_cache_probe(workload, reps=20):
    """The level-0 two-sweep pass with the device cache warm and cold: the
    pass re-run on the same x and b (both fit the 256 MB Infinity Cache), and
    the same with a 256x384x256 face->centre pass of another context (about
    600 MB read and written) between launches, which evicts them.  Average
    launch microseconds of each, from the library's HIP events."""
    import numpy as np
    from mac3d import MacSolver, scenarios
    from mac3d.lib import config_for
    rng = np.random.default_rng(3)
    big = scenarios.scaled(scenarios.workload("large", 1), precision=0)
    out = {}
    with MacSolver(config_for(workload)) as s, MacSolver(config_for(big)) as flusher:
        shp = s.level_shape(0)
        s.upload("MGX", rng.standard_normal(shp).astype(np.float32), 0)
        s.upload("MGB", rng.standard_normal(shp).astype(np.float32), 0)
        for mode in ("warm", "cold"):
            s.mg("SMOOTH", 0, 2)  # untimed
            s.profile(True)
            for _ in range(reps):
                if mode == "cold":
                    flusher.stage("FACE_TO_CENTRE")
                s.mg("SMOOTH", 0, 2)
            ms, n = s.profile_read("SMOOTH2_L0")
            s.profile(False)
            out[mode] = 1e3 * ms / max(n, 1)
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, steps, threads):
    """Seconds per step of the FP64 oracle on `threads` host threads, timed
    over `steps` steps after the first (the first touches the pages)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    used = oracle.set_threads(threads)
    o = oracle.Oracle(workload.nx, workload.ny, workload.nz, workload.dx, levels=workload.levels,
                      nu1=workload.nu1, nu2=workload.nu2, cycles=workload.cycles)
    o.init_blobs(workload.blobs)
    o.step(workload.dt, 1)
    t0 = time.perf_counter()
    o.step(workload.dt, steps)
    dt = time.perf_counter() - t0
    o.close()
    return dt / steps, used


def js_baseline(name, steps, threads):
    """The FP64 JavaScript transcription of the step (oracle/js), the CPU path
    in the reference's language family, on `threads` Node worker threads:
    `steps` steps of the workload, the last one's time (the first includes
    the compiler's warm-up).  None when node is missing or the run fails."""
    node = shutil.which("node")
    if node is None:
        return None
    try:
        r = subprocess.run([node, "--max-old-space-size=8192",
                            os.path.join(ROOT, "oracle", "js", "parallel.js"), name, str(steps),
                            str(threads)], capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log(f"[bench] js baseline failed: {r.stderr[-400:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-wave", action="store_true")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the smoothing pass's warm/cold cache probe")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, 5 device copies of one FP32 field (a plain copy "
                         "kernel of known size, to calibrate HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the MAC step is "
              "built and measured per GPU; multi-GPU slab decomposition is design-only.",
              file=sys.stderr)
        sys.exit(2)

    from mac3d import scenarios
    w32 = scenarios.workload("ref", args.seed)
    log(f"[bench] fp32 {w32.nx}x{w32.ny}x{w32.nz} steps={args.steps} warmup={args.warmup}")
    st32, wall32, prof = run_gpu(w32, args.steps, args.warmup)
    sps32 = args.steps / wall32
    log(f"[bench] fp32 {sps32:.2f} steps/s (kernel span {st32['ms']:.1f} ms) "
        f"max_div {st32['max_div']:.3e}")

    mixed = None
    if not args.no_mixed:
        wm = scenarios.workload("mixed", args.seed)
        log("[bench] mixed")
        stm, wallm, profm = run_gpu(wm, args.steps, args.warmup, profile_steps=0)
        spsm = args.steps / wallm
        mixed = {"steps_per_s": round(spsm, 3), "speedup_vs_fp32": round(spsm / sps32, 3),
                 "vcycles_per_step": round(stm["vcycles"] / args.steps, 3),
                 "max_div": stm["max_div"], "rel_res": stm["rel_res"],
                 "converged": bool(stm["converged"])}
        log(f"[bench] mixed {spsm:.2f} steps/s, {mixed['vcycles_per_step']} cycles/step")

    wave = None
    if not args.no_wave:
        log(f"[bench] fp32 with the wave coupling ({WAVE_N}^2 wave grid)")
        stw, wallw, profw = run_gpu(w32, args.steps, args.warmup, wave=True)
        spsw = args.steps / wallw
        ms, n = profw.get("WAVE", (0.0, 0))
        wave_us = 1e3 * ms / PROFILE_STEPS if n else None
        wbytes = WAVE_BYTES_PER_TEXEL * WAVE_N * WAVE_N
        wave = {"steps_per_s": round(spsw, 3), "n": WAVE_N,
                "wave_us_per_step": round(wave_us, 2) if wave_us else None,
                "wave_bytes_per_step": wbytes,
                "wave_gbps": round(wbytes / (wave_us * 1e-6) / 1e9, 1) if wave_us else None,
                "max_div": stw["max_div"]}
        log(f"[bench] wave-coupled fp32 {spsw:.2f} steps/s, wave step {wave['wave_us_per_step']} us")
        # the same with the reference scene's terms in the advection and the
        # Casson pass (curl pass, vorticity confinement, unfurling forcing)
        sts, walls, profs = run_gpu(w32, args.steps, args.warmup, wave=True, scene=True)

        def per_step_us(prof, name):
            ms, n = prof.get(name, (0.0, 0))
            return round(1e3 * ms / PROFILE_STEPS, 2) if n else None
        wave["scene"] = {"steps_per_s": round(args.steps / walls, 3),
                         "advect_us_per_step": per_step_us(profs, "ADVECT"),
                         "casson_us_per_step": per_step_us(profs, "CASSON"),
                         "advect_us_per_step_no_scene": per_step_us(profw, "ADVECT"),
                         "casson_us_per_step_no_scene": per_step_us(profw, "CASSON"),
                         "max_div": sts["max_div"]}
        log(f"[bench] scene-term fp32 {wave['scene']['steps_per_s']:.2f} steps/s")
        wave["parcels"] = {"count": PARCELS}
        for order in ("cell", "random"):
            ms_p, live = run_parcels(w32, max(5, min(args.steps, 20)), order)
            wave["parcels"][order] = {
                "live_after": live, "us_per_step": round(1e3 * ms_p, 2),
                "parcels_per_s": round(PARCELS / (ms_p * 1e-3), 1),
                "state_gbps": round(PARCEL_BYTES * PARCELS / (ms_p * 1e-3) / 1e9, 1)}
            log(f"[bench] parcels {PARCELS} ({order} order): {1e3 * ms_p:.1f} us per step")
        n_rs = max(10, min(args.steps, 40))
        sps_rs, live_rs = run_reference_scene(n_rs, 5)
        wave["reference_scene"] = {
            "steps_per_s": round(sps_rs, 3), "ms_per_step": round(1e3 / sps_rs, 4),
            "steps": n_rs, "warmup": 5, "live_parcels_after": live_rs,
            "config": "96x128x96 MAC grid (10 mm), dt 1/60 s, 3-level V-cycle 2+2 x 8, "
                      "512^2 wave, scene terms and the reference's physics, 393216-parcel "
                      f"ring with {REF_SCENE_EMIT} emissions per step, 4-channel "
                      "concentration on 192x256x192, FP32, quiescent start"}
        log(f"[bench] reference scene at its shape: {sps_rs:.2f} steps/s "
            f"({1e3 / sps_rs:.3f} ms/step), {live_rs} live parcels")

    roof = None
    if "SMOOTH2_L0" in prof:
        ms, n = prof["SMOOTH2_L0"]
        avg_s = ms / n / 1e3
        # achieved: the bytes the fused two-sweep launch must move (read x
        # and b once, write x once: 12 B per cell), so frac is a fraction of
        # the HBM peak; the sweep-equivalent rate charges the two-colour-pass
        # algorithm's 24 B per cell per sweep and can exceed the peak
        kbytes = KERNEL_BYTES_PER_CELL * w32.cells
        ach = kbytes / avg_s / 1e9
        sweep_bytes = SWEEP_BYTES_PER_CELL * SWEEPS_PER_LAUNCH * w32.cells
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": "k_rbgs_sweep2 (level 0, two red-black sweeps per launch)",
                "avg_launch_us": round(avg_s * 1e6, 2), "launches": n,
                "bytes_per_launch": kbytes,
                "sweep_equivalent_bytes_per_launch": sweep_bytes,
                "sweep_equivalent_gbps": round(sweep_bytes / avg_s / 1e9, 1)}
        traffic_file = os.path.join(ROOT, "profiles", "rbgs2_l0_192x256x192_traffic.json")
        if os.path.exists(traffic_file):
            with open(traffic_file) as f:
                t = json.load(f)
            roof["traffic"] = t.get("bytes_per_launch")
            if roof["traffic"]:
                roof["traffic_gbps"] = round(roof["traffic"] / avg_s / 1e9, 1)
                roof["traffic_frac"] = round(roof["traffic"] / avg_s / 1e9 / HBM_PEAK_GBPS, 4)
    if roof is not None and not args.no_cold:
        cc = smoother_cache_probe(w32)
        kb = roof["bytes_per_launch"]
        roof["cache_probe"] = {
            "warm_avg_launch_us": round(cc["warm"], 2), "cold_avg_launch_us": round(cc["cold"], 2),
            "cold_achieved": round(kb / (cc["cold"] * 1e-6) / 1e9, 1),
            "cold_frac": round(kb / (cc["cold"] * 1e-6) / 1e9 / HBM_PEAK_GBPS, 4),
            "how": "20 launches on the same x and b (warm), and 20 each after a 256x384x256 "
                   "face->centre pass of another context (~600 MB) evicts the Infinity Cache "
                   "(cold)"}
        log(f"[bench] level-0 smoothing pass: warm {cc['warm']:.1f} us, cold {cc['cold']:.1f} us")
    kernel_ms = {k: round(v[0] / v[1], 4) for k, v in prof.items()}
    if args.copy_probe:
        copy_probe(w32)

    cpu = None
    if not args.no_cpu:
        # the box's CPU share is 16 threads (OMP_NUM_THREADS there)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline: FP64 oracle, 2 steps on {threads} threads, 1 on 1 thread")
        s_all, used = cpu_baseline(w32, 2, threads)
        s_one, _ = cpu_baseline(w32, 1, 1)
        c_port = {"kind": "port", "language": "C (OpenMP), FP64",
                  "seconds_per_step_all_threads": round(s_all, 3), "threads": used,
                  "seconds_per_step_single_thread": round(s_one, 3),
                  "single_thread_steps_per_s": round(1.0 / s_one, 5),
                  "sample": f"2 steps on {used} OpenMP threads ({2 * s_all:.1f} s) and 1 step on "
                            f"1 thread ({s_one:.1f} s), after one untimed step"}
        log(f"[bench] cpu (C) {s_all:.2f} s/step on {used} threads, {s_one:.2f} s/step on 1")
        js = js_baseline("ref", 3, threads)
        js1 = js_baseline("tiny", 20, 1)
        if js is not None:
            warm = js["per_step"][1:]
            spp = sum(warm) / len(warm)
            log(f"[bench] cpu (JavaScript) {spp:.2f} s/step on {js['threads']} worker threads")
            cpu = {"value": round(1.0 / spp, 5), "unit": "steps/s",
                   "cores": js["threads"], "kind": "port",
                   "cpu_model": js.get("cpu_model") or cpu_model(),
                   "host_threads_visible": js.get("cpus", os.cpu_count()),
                   "language": f"JavaScript (Node {js['node']}, worker_threads), FP64",
                   "seconds_per_step_all_threads": round(spp, 3),
                   "seconds_per_step_first": round(js["per_step"][0], 3),
                   "sample": "the 192x256x192 FP32 workload's step (8 V-cycles) in the FP64 "
                             "JavaScript transcription (oracle/js/parallel.js): 3 steps on "
                             f"{js['threads']} worker threads, the last 2 timed "
                             f"({spp:.2f} s/step; the first, with the compiler's warm-up, "
                             f"{js['per_step'][0]:.2f} s)",
                   "c_oracle": c_port}
            if js1 is not None:
                cpu["single_thread_32x48x32"] = {
                    "seconds_20_steps": round(js1["seconds"], 3),
                    "steps_per_s": round(20 / js1["seconds"], 3),
                    "sample": "the 32x48x32 workload (one blob, 8 V-cycles per step), 20 steps "
                              "on one thread, timed in full (compiler warm-up included)"}
        else:
            cpu = {"value": round(1.0 / s_all, 5), "unit": "steps/s", "cores": used,
                   "kind": "port", "cpu_model": cpu_model(),
                   "host_threads_visible": os.cpu_count(), "c_oracle": c_port,
                   "seconds_per_step_all_threads": round(s_all, 3),
                   "seconds_per_step_single_thread": round(s_one, 3),
                   "sample": "the 192x256x192 FP32 workload's step (8 V-cycles) in the FP64 C "
                             "oracle (node is not installed): " + c_port["sample"]}

    line = {
        "metric": "MAC steps/sec @192x256x192 (mixed vs FP32, div<=1e-6); HBM GB/s % peak",
        "value": round(sps32, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall32 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded blobs)",
        "config": {"workload": "192x256x192 MAC step, 3 seeded blobs, FP32, 3-level V-cycle "
                               "2+2 RBGS x 8 cycles", "nx": w32.nx, "ny": w32.ny, "nz": w32.nz,
                   "dx": w32.dx, "dt": w32.dt, "seed": args.seed, "parallelism": "single GPU"},
        "max_div": st32["max_div"],
        "mixed": mixed,
        "wave": wave,
        "roofline": roof,
        "kernel_avg_ms": kernel_ms,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
