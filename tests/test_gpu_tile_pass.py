"""The coarse levels' two-sweep tile pass against the z-marching two-sweep
pass, bit for bit, through the C ABI (needs an MI355X).

Levels >= 1 of at least 512 cells smooth pairs of red-black sweeps with one
3-D tile per block: 16x8, 32x8 or 48x8 cells by the level's size and width,
8, 6 or 2 planes deep (the deepest that still gives every CU of a 256-CU
device two blocks); level 0's z-marching pass (MAC3D_MG_SMOOTH2_MARCH) runs
the same cell arithmetic, so the two must agree exactly: from a given guess
and from a zero guess, omega 1 and 1.2, FP32 and FP16 storage, on grids that
fill their tiles and on grids with partial tiles in every direction, with
each of the nine tile shapes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


CASES = [
    # (level-0 dims, coarsenings, level, precision); the level's tile shape
    ((192, 256, 192), 3, 1, 0),   # 96x128x96: 48x8x6
    ((192, 256, 192), 3, 2, 0),   # 48x64x48: 16x8x2
    ((192, 256, 192), 3, 2, 1),   # 48x64x48, FP16 storage (mixed): 16x8x2
    ((200, 264, 200), 3, 1, 0),   # 100x132x100: 16x8x8, partial tiles in x, y, z
    ((200, 264, 200), 3, 2, 1),   # 50x66x50, FP16: 16x8x2, partial tiles
    ((192, 264, 200), 3, 1, 0),   # 96x132x100: 48x8x6, partial in y and z
    ((256, 384, 256), 5, 1, 1),   # 128x192x128 (the large case's level 1, mixed: FP32):
                                  # 32x8x8
    ((256, 384, 256), 5, 2, 1),   # 64x96x64, FP16: 16x8x6
    ((256, 384, 256), 5, 3, 1),   # 32x48x32, FP16: 16x8x2, the large case's level 3
    ((288, 256, 192), 3, 1, 0),   # 144x128x96: 48x8x8
    ((128, 256, 192), 3, 1, 0),   # 64x128x96: 32x8x6
    ((128, 272, 128), 3, 1, 0),   # 64x136x64: 32x8x2, partial in y
    ((192, 128, 192), 3, 1, 0),   # 96x64x96: 48x8x2
    ((32, 48, 32), 3, 1, 0),      # 16x24x16: the tiny workload's level 1, 16x8x2
    ((32, 48, 32), 3, 2, 1),      # 8x12x8, FP16: 16x8x2 tiles over the whole level
]


@pytest.mark.parametrize("dims,levels,level,prec", CASES)
@pytest.mark.parametrize("omega", [1.0, 1.2])
def test_tile_pass_equals_marching_pass(dims, levels, level, prec, omega):
    from mac3d import MacSolver
    from mac3d.lib import config_for
    from mac3d import scenarios
    w = scenarios.scaled(scenarios.workload("tiny", 1), nx=dims[0], ny=dims[1], nz=dims[2],
                         levels=levels, precision=prec)
    print(f"[tile2] dims={dims} levels={levels} level={level} prec={prec} omega={omega}",
          flush=True)
    rng = np.random.default_rng(level * 7 + prec + int(omega * 10))
    with MacSolver(config_for(w, sor_omega=omega)) as s:
        shp = s.level_shape(level)
        assert np.prod(shp) >= 512
        x = rng.standard_normal(shp)
        b = rng.standard_normal(shp)
        res = {}
        for name, op, n in [("tile", "SMOOTH", 2), ("march", "SMOOTH2_MARCH", 0),
                            ("ztile", "SMOOTH_ZERO", 2), ("zmarch", "SMOOTH2_MARCH", 1)]:
            s.upload("MGX", x, level)
            s.upload("MGB", b, level)
            s.mg(op, level, n)
            res[name] = s.download("MGX", level, dtype=np.float32)
        assert np.array_equal(res["tile"], res["march"])
        assert np.array_equal(res["ztile"], res["zmarch"])
        assert not np.array_equal(res["tile"], res["ztile"])
