"""The JavaScript FP64 transcription (oracle/js/mac3d_oracle.js), the CPU
path in the reference's own language family, against the C oracle on the
same seeded workload: both restate the same kernels in FP64, so after two
steps of the 32x48x32 workload (8 V-cycles each) every field agrees to
1e-11 relative L2 (the summation orders differ only in the coarse solve's
conjugate-gradient reductions)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from mac3d import scenarios

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_js_oracle_matches_c_oracle(tmp_path):
    steps = 2
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "oracle_run.js"), "tiny",
                        str(steps), str(tmp_path)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout.strip().splitlines()[-1])
    w = scenarios.workload("tiny", 1)
    o = oracle.Oracle(w.nx, w.ny, w.nz, w.dx, levels=w.levels, cycles=w.cycles)
    o.init_blobs(w.blobs)
    o.step(w.dt, steps)
    got = np.fromfile(str(tmp_path / "fields.bin"), dtype=np.float64).reshape(5, w.nz, w.ny, w.nx)
    for q, f in enumerate((oracle.F_U, oracle.F_V, oracle.F_W, oracle.F_P, oracle.F_DYE)):
        want = o.get(f)
        err = np.linalg.norm(got[q] - want) / np.linalg.norm(want)
        assert err <= 1e-11, (q, err)
    mx, l2 = o.div_norms()
    assert js["maxDiv"] == pytest.approx(mx, rel=1e-6, abs=1e-14)
    o.close()


@pytest.mark.skipif(shutil.which("node") is None, reason="node not installed")
def test_js_oracle_threads_match_one_thread(tmp_path):
    """The worker-thread step (oracle/js/parallel.js, a share of the z-planes
    per worker, a barrier between stages) gives the single-threaded step's
    fields bit for bit: within one colour pass and one stage no cell reads
    another that the same pass writes."""
    one = tmp_path / "one"
    one.mkdir()
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "oracle_run.js"), "tiny", "2",
                        str(one)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    par = tmp_path / "par.bin"
    r = subprocess.run(["node", os.path.join(ROOT, "oracle", "js", "parallel.js"), "tiny", "2", "3",
                        str(par)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["threads"] == 3 and len(out["per_step"]) == 2
    assert np.array_equal(np.fromfile(str(par)), np.fromfile(str(one / "fields.bin")))
