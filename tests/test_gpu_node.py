"""The JavaScript host (js/MacSolver.js over the N-API addon) steps
the tiny workload on the GPU and agrees with the FP64 oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from mac3d import scenarios

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_node_host_steps_tiny():
    steps = 4
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "step_tiny.js"), str(steps)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["badDt"] is not None and "dt" in out["badDt"]
    assert out["stats"]["steps"] == steps and out["stats"]["maxDiv"] <= 1e-6
    w = scenarios.workload("tiny", 1)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, cycles=w.cycles)
    o.init_blobs(w.blobs)
    o.step(w.dt, steps)
    for name, fid in (("u", oracle.F_U), ("v", oracle.F_V), ("w", oracle.F_W),
                      ("p", oracle.F_P), ("dye", oracle.F_DYE)):
        want = np.linalg.norm(o.get(fid))
        assert abs(out["norms"][name] - want) <= 1e-5 * want + 1e-12, name


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_node_backward_matches_c_abi():
    from mac3d import MacSolver
    from mac3d.lib import config_for
    steps = 3
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "step_tiny.js"), str(steps),
                        "backward"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])["backward"]
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=2)
    with MacSolver(config_for(w)) as s:
        g = s.adjoint(w.blobs, w.dt, steps, np.zeros((w.nz, w.ny, w.nx)), checkpoint_every=2)
    assert abs(js["loss"] - g["loss"]) <= 1e-12 * g["loss"]
    for p, n in enumerate(["cx", "cy", "cz", "radius", "w0"]):
        assert abs(js["dBlobs"][0][n] - g["d_blob"][0][p]) <= 1e-9 * abs(g["d_blob"][0][p]) + 1e-12
    assert abs(js["dRhoB"] - g["d_rho_b"]) <= 1e-9 * abs(g["d_rho_b"]) + 1e-15


def test_node_wave_coupling_matches_c_abi():
    """waveEnable / setDrops through the N-API host give the C ABI's fields
    and wave state after the same steps."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    steps = 3
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "step_tiny.js"), str(steps),
                        "wave"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])["wave"]
    w = scenarios.workload("tiny", 1)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.wave_enable(n=64, half_extent=0.06, half_xz=0.016, depth=0.048, surface_y=0.0,
                      marangoni=0.8)
        s.set_drops([(0.004, -0.003, -0.002, 0.02, -0.4, 0.0, 0.55, 1.3),
                     (-0.01, 0.01, 0.006, 0.018, -0.2, 0.0, 0.2, 2.1)])
        s.step(w.dt, steps)
        got = {"v": s.download("V"), "p": s.download("P"), "dye": s.download("DYE"),
               "wave": s.download("WAVE"), "ptop": s.download("PTOP")}
    assert js["wave"] > 0 and js["ptop"] > 0
    for k, a in got.items():
        assert js[k] == pytest.approx(float(np.sqrt(np.sum(a * a))), rel=1e-12), k


def test_node_parcel_stage_matches_c_abi():
    """parcelsUpload / parcelsStep / splatDownload through the N-API host
    give the C ABI's parcels and counters bit for bit (512 parcels laid out by
    the same formula after the same coupled steps)."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    steps = 2
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "step_tiny.js"), str(steps),
                        "parcels"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])["parcels"]
    w = scenarios.workload("tiny", 1)
    q = np.arange(512)
    P = np.zeros((512, 8))
    P[:, 0] = -0.008 + 0.016 * ((q * 37) % 64) / 64
    P[:, 1] = -0.04 + 0.035 * ((q * 11) % 128) / 128
    P[:, 2] = -0.008 + 0.016 * ((q * 23) % 64) / 64
    P[:, 3] = 0.1
    P[:, 4] = np.where(q % 3, 0.002, 0.006)
    P[:, 5] = np.where(q % 3, 0.0, 0.004)
    P[:, 7] = (q * 7) % 3000
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.wave_enable(n=64, half_extent=0.06, half_xz=0.016, depth=0.048, surface_y=0.0)
        s.step(w.dt, steps)
        s.parcels_upload(P)
        s.stage("PARCELS", w.dt)
        out = s.parcels_download().astype(np.float64)
        intact, lysed = int(s.splat_download(0).sum()), int(s.splat_download(1).sum())
    assert intact > 0
    assert js["intact"] == intact and js["lysed"] == lysed
    assert js["p"] == pytest.approx(float(np.sqrt(np.sum(out * out))), rel=1e-12)


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_node_concentration_on_the_finer_grid_matches_c_abi():
    """concEnable(true, 2) through the N-API host: the concentration, its
    splat counters and the fields after coupled steps with emission equal the
    C ABI's run of the same calls (the counters bit for bit)."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    steps = 2
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "step_tiny.js"), str(steps),
                        "conc2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])["conc2"]
    w = scenarios.workload("tiny", 1)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.wave_enable(n=64, half_extent=0.25, half_xz=0.1, depth=0.2, surface_y=0.0,
                      marangoni=0.8, scene=1, physics=1)
        s.set_drops([(0.004, -0.003, -0.002, 0.02, -0.4, 0.0, 0.85, 1.3)])
        s.conc_enable(True, refine=2)
        s.parcels_upload(np.zeros((4096, 8)))
        for q in range(steps):
            s.parcels_emit(300 * q, 300)
            s.step(w.dt, 1)
        conc = s.download("CONC")
        v = s.download("V").astype(np.float32).astype(np.float64)
        intact, lysed = int(s.splat_download(0).sum()), int(s.splat_download(1).sum())
        assert js["cells"] == 8 * int(np.prod(s.shape)) == conc.shape[0] * conc.shape[1] * conc.shape[2]
    assert intact > 0 and js["intact"] == intact and js["lysed"] == lysed
    assert js["conc"] == pytest.approx(float(np.sqrt(np.sum(conc * conc))), rel=1e-12)
    assert js["v"] == pytest.approx(float(np.sqrt(np.sum(v * v))), rel=1e-12)


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
@pytest.mark.parametrize("mode", ["fp32", "mixed"])
def test_node_checkpoint_round_trip(mode, tmp_path):
    """saveCheckpoint -> step -> loadCheckpoint -> step through the N-API
    host repeats the steps bit for bit; the same bytes loaded into a C ABI
    solver give the same fields after the same steps."""
    from mac3d import MacSolver
    from mac3d.lib import config_for
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "checkpoint_tiny.js"),
                        str(tmp_path), mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["identical"] is True
    assert out["badLoad"] is not None and "status 1" in out["badLoad"]  # EINVAL: truncated
    ck = np.fromfile(str(tmp_path / "ck.bin"), dtype=np.uint8)
    js_fields = np.fromfile(str(tmp_path / "fields.bin"), dtype=np.float32)
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=1 if mode == "mixed" else 0)
    with MacSolver(config_for(w)) as s:
        s.checkpoint_load(ck)
        s.step(w.dt, 2)
        mine = np.concatenate([s.download(f, dtype=np.float32).ravel()
                               for f in ("U", "V", "W", "P", "DYE")])
    assert np.array_equal(mine, js_fields)


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_node_release_gradient_and_fit_match_c_abi():
    """releaseMap / releaseBlobs / backwardRelease / fitRelease through the
    N-API host against the same calls through the C ABI: the map and blobs
    bit for bit, the gradient to 1e-10 (the adjoint's scatters are atomics,
    so the last bits depend on the order), two fit iterations to 1e-8; a
    release with missing fields is refused."""
    from mac3d import MacSolver, lib
    steps, iters = 6, 2
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "release_tiny.js"), str(steps),
                        str(iters)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])
    assert js["bad"] is not None and "status 1" in js["bad"]
    w = scenarios.Workload("release", 32, 48, 32, 1e-3, 5e-4, steps, 3, 8, 2)
    from mac3d.lib import config_for
    cfg = config_for(w)
    m = lib.release_default_map(cfg)
    m.vel_scale = 2.5
    names = dict(origin_x="originX", origin_y="originY", origin_z="originZ", cell="cell",
                 surface_y="surfaceY", gravity="gravity", vel_scale="velScale",
                 submerge="submerge", max_offset="maxOffset", r_min="rMin", r_max="rMax")
    for c, j in names.items():
        assert getattr(m, c) == js["map"][j], c
    truth = dict(x=0.05, z=-0.03, radius=0.09, start_y=0.495, vel_y=-0.075)
    guess = dict(x=0.08, z=-0.05, radius=0.1, start_y=0.495, vel_y=-0.15)
    (bg,), _ = lib.release_blobs(m, [guess])
    for n in lib.BLOB_PARAMS:
        assert getattr(bg, n) == js["blobs"][0][n], n
    blobs, _ = lib.release_blobs(m, [truth])
    with MacSolver(cfg) as s:
        s.init_blobs(blobs)
        s.step(w.dt, steps)
        target = s.download("DYE")
        g = s.adjoint_release(m, [guess], w.dt, steps, target, checkpoint_every=3)
        fit, rep = s.fit_release(m, [guess], w.dt, steps, target, max_iters=iters,
                                 checkpoint_every=3)
    jsn = dict(x="x", z="z", radius="radius", start_y="startY", vel_y="velY")
    assert js["grad"]["loss"] == pytest.approx(g["loss"], rel=1e-12)
    for b, n in enumerate(lib.RELEASE_PARAMS):
        assert js["grad"]["dReleases"][0][jsn[n]] == pytest.approx(
            g["d_release"][0][b], rel=1e-10, abs=1e-10 * max(map(abs, g["d_release"][0]))), n
        assert js["fit"]["releases"][0][jsn[n]] == pytest.approx(fit[0][n], rel=1e-8, abs=1e-12), n
    assert js["fit"]["iters"] == rep["iters"] and js["fit"]["loss"] == pytest.approx(rep["loss"], rel=1e-8)
