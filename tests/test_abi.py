"""C ABI checks that need no GPU: the library loads, exports every function
include/mac3d.h declares, and rejects bad configurations before touching
the device."""
import ctypes as C
import os
import re

import pytest

import mac3d
from mac3d import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mac3d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mac3d_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = mac3d.load_library()
    names = declared()
    assert "mac3d_step" in names and "mac3d_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(L.EXPORTS)


def test_version_and_defaults():
    lib = mac3d.load_library()
    assert lib.mac3d_version().startswith(b"mac3d")
    c = L.default_config()
    assert (c.nx, c.ny, c.nz, c.levels, c.nu1, c.nu2, c.cycles) == (32, 48, 32, 3, 2, 2, 8)
    assert c.rho_b == 1060.0 and c.tau_y == 0.005 and c.mu_inf == 0.0035
    assert abs(c.maccormack - 0.45 * 0.995) < 1e-15
    assert c.gamma_reg == 0.40


@pytest.mark.parametrize("bad", [
    dict(nx=30),                  # not divisible by 2^levels
    dict(levels=0),
    dict(levels=9),
    dict(dx=0.0),
    dict(dx=float("nan")),
    dict(precision=7),
    dict(nu1=-1),
    dict(nx=512, nz=512, ny=8, levels=3),  # coarsest plane too large
])
def test_create_rejects_bad_config(bad):
    lib = mac3d.load_library()
    c = L.default_config(**bad)
    h = C.c_void_p()
    rc = lib.mac3d_create(C.byref(c), C.byref(h))
    assert rc == 1 and not h.value
    assert len(lib.mac3d_last_error(None)) > 0


def test_null_handles():
    lib = mac3d.load_library()
    assert lib.mac3d_destroy(None) == 0
    assert lib.mac3d_step(None, 1e-3, 1, None) == 1
    assert lib.mac3d_upload(None, 0, 0, None, 0, 0) == 1


@pytest.mark.parametrize("table,prefix", [("FIELD", "MAC3D_FIELD_"), ("STAGE", "MAC3D_STAGE_"),
                                          ("MG", "MAC3D_MG_"), ("KC", "MAC3D_KC_")])
def test_host_id_tables_match_the_header(table, prefix):
    """The Python host's field, stage, multigrid-op and kernel-class ids are
    the header's #defines, name for name."""
    src = open(os.path.join(ROOT, "include", "mac3d.h")).read()
    defs = {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define " + prefix + r"([A-Z0-9_]+)\s+(\d+)", src)}
    defs.pop("COUNT", None)
    assert getattr(L, table) == defs
