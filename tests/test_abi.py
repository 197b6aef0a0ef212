"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import os
import re

import paragraph_amd as pg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "paragraph.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(pg_\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ["pg_ctx_create", "pg_year_step", "pg_graph_load", "pg_bfs", "pg_version"]:
        assert must in names


def test_library_exports_all_declared_symbols():
    L = ctypes.CDLL(pg.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_version_and_growth_rule():
    assert pg.lib().pg_version() == 1
    # reference rule: ceil(n * rate) in double precision each year
    assert pg.growth_schedule(50000, 0.03, 3) == [1500, 1545, 1592]
    s = pg.growth_schedule(2_000_000, 0.03, 10)
    assert sum(s) == 687_837 and s[0] == 60000 and s[-1] == 78287


def test_create_without_device_fails_cleanly():
    import pytest
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present")
    cfg = pg.Config(ctypes.sizeof(pg.Config), 0, 100, 100, 10, 4, 16, 1)
    h = ctypes.c_void_p()
    rc = pg.lib().pg_ctx_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == 6  # PG_ENODEV
    assert b"device" in pg.lib().pg_last_error(None)


def test_bad_config_rejected():
    cfg = pg.Config(ctypes.sizeof(pg.Config), 0, 100, 100, 10, 65, 16, 1)  # max_cites > 64
    h = ctypes.c_void_p()
    assert pg.lib().pg_ctx_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
    cfg = pg.Config(ctypes.sizeof(pg.Config) - 4, 0, 100, 100, 10, 4, 16, 1)
    assert pg.lib().pg_ctx_create(ctypes.byref(cfg), ctypes.byref(h)) == 1
