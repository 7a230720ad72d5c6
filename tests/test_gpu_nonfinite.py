"""The non-finite sentinel of mac3d_step (MAC3D_ENONFINITE): a NaN or an
infinity in the velocity is reported by the step's FP64 divergence norms
(stats requested) or by the mixed mode's residual norms, and the caller
recovers by loading a checkpoint saved before it.

(Every texel index the step forms from a velocity is clamped to the grid
or read through a range-checked buffer resource, so a non-finite velocity
gives non-finite values, not out-of-range accesses.)"""
import numpy as np
import pytest

from mac3d import MacSolver, scenarios
from mac3d.lib import Mac3dError, config_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_velocity_is_reported_and_recovered(prec, bad):
    w = scenarios.scaled(scenarios.workload("tiny", 1), precision=prec)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        s.step(w.dt, 2)
        ck = s.checkpoint_save()
        good = s.step(w.dt, 1)
        assert np.isfinite(good["max_div"])
        s.checkpoint_load(ck)
        u = s.download("U")
        u[5, 10, 7] = bad
        s.upload("U", u)
        with pytest.raises(Mac3dError) as e:
            s.step(w.dt, 1)
        assert e.value.code == 5 and "finite" in str(e.value)
        s.checkpoint_load(ck)
        again = s.step(w.dt, 1)
        assert again["max_div"] == good["max_div"]
        assert np.isfinite(s.download("P")).all()


def test_fp32_step_without_stats_does_not_test():
    """Without stats the FP32 step reads nothing back (its captured sequence
    replays with no host test): the state goes non-finite silently, and the
    next step with stats reports it."""
    w = scenarios.workload("tiny", 1)
    with MacSolver(config_for(w)) as s:
        s.init_blobs(w.blobs)
        u = s.download("U")
        u[3, 3, 3] = np.nan
        s.upload("U", u)
        s.step_nostats(w.dt, 1)
        with pytest.raises(Mac3dError) as e:
            s.step(w.dt, 1)
        assert e.value.code == 5
