"""Error paths that must leave a context consistent (needs an MI355X).

- A failed growth of the per-year edge buffers (an injected out-of-memory)
  returns PG_ENOMEM and keeps the old buffers: the same year then runs
  correctly, edge set equal to the oracle's.
- A checkpoint whose graph state is inconsistent but whose checksum was
  recomputed (ids out of range, a backward CSR that is not the transpose, a
  wrong in-degree, offsets out of order) is refused with PG_EIO before any
  of it reaches the device.
"""
import os
import struct
import sys

import numpy as np
import pytest

import paragraph_amd as pg
from oracle import binding as ob

pytestmark = pytest.mark.gpu

HEADER = struct.Struct("<8sIIQqqqqiiiiQ")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_failed_edge_buffer_growth_is_recoverable():
    n0, D, p = 20000, 600, 1e-2
    none = np.zeros(0, np.int32)
    g = pg.SeedGraph(np.zeros(n0, np.int32), none, none, np.zeros(n0 + D, np.float32))
    # max_cites=1: the per-year edge buffers start at D entries, and this
    # year's ~120k ER edges must grow them
    ctx = pg.Context(n0 + D, 200_000, D, max_cites=1, bfs_cap=16, seed=9)
    ctx.load(g)
    log("ER year with an injected allocation failure")
    os.environ["PG_TEST_FAIL_ALLOC_BYTES"] = str(64 * 1024)
    try:
        with pytest.raises(pg.PGError) as e:
            ctx.year_er(1, 0, D, p)
    finally:
        del os.environ["PG_TEST_FAIL_ALLOC_BYTES"]
    assert e.value.code == 2  # PG_ENOMEM
    assert ctx.size() == (n0, 0)
    st = ctx.year_er(1, 0, D, p)
    off, tgt = ob.er_year(n0, D, p, 9, 0)
    assert st.edges_added == tgt.size > 100_000
    d = ctx.download()
    assert np.array_equal(d["fwd_adj"], tgt)
    assert np.array_equal(d["fwd_off"][n0:] - d["fwd_off"][n0], off)
    assert np.array_equal(np.diff(d["bwd_off"]), d["in_deg"])


def fnv1a(buf: bytes) -> int:
    h = 0xcbf29ce484222325
    for c in buf:
        h = ((h ^ c) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def split(raw: bytes):
    h = list(HEADER.unpack_from(raw, 0))
    n, e, nl = h[4], h[5], h[6]
    arrays, pos = {}, HEADER.size
    for k, cnt, dt in [("birth", n, np.int32), ("in_deg", n, np.int32),
                       ("fwd_off", n + 1, np.int32), ("fwd_adj", e, np.int32),
                       ("bwd_off", n + 1, np.int32), ("bwd_adj", e, np.int32),
                       ("logfit", nl, np.float32)]:
        arrays[k] = np.frombuffer(raw, dt, cnt, pos).copy()
        pos += cnt * 4
    return h, arrays, raw[pos:]


def join(h, arrays, user: bytes) -> bytes:
    body = b"".join(arrays[k].tobytes() for k in
                    ("birth", "in_deg", "fwd_off", "fwd_adj", "bwd_off", "bwd_adj", "logfit"))
    body += user
    h = list(h)
    h[-1] = fnv1a(body)
    return HEADER.pack(*h) + body


def test_checkpoint_with_inconsistent_state_is_refused(tmp_path):
    g = pg.seed_graph(8, 1500, 5.0, n_final=1600)
    ctx = pg.Context(1600, g.src.size + 100 * 6 + 16, 100, max_cites=6, seed=8)
    ctx.load(g)
    ctx.year_pa(1, 0, 100, 6, 2, 1.0, 5.0)
    good = str(tmp_path / "good.pgck")
    ctx.checkpoint_save(good, b"blob")
    raw = open(good, "rb").read()
    h, arr, user = split(raw)
    assert join(h, arr, user) == raw  # the format is understood
    n = h[4]
    v = int(np.argmax(np.diff(arr["fwd_off"])))  # a paper with citations
    t = int(arr["fwd_adj"][arr["fwd_off"][v]])
    cases = {}
    a = {k: x.copy() for k, x in arr.items()}
    a["fwd_adj"][arr["fwd_off"][v]] = n + 5
    cases["forward id out of range"] = a
    a = {k: x.copy() for k, x in arr.items()}
    a["fwd_adj"][arr["fwd_off"][v]] = v
    cases["self-citation"] = a
    a = {k: x.copy() for k, x in arr.items()}
    a["in_deg"][t] += 1
    cases["in-degree"] = a
    a = {k: x.copy() for k, x in arr.items()}
    b0, b1 = arr["bwd_off"][t], arr["bwd_off"][t + 1]
    assert b1 - b0 >= 1
    a["bwd_adj"][b0] = (a["bwd_adj"][b0] + 1) % n
    cases["transpose"] = a
    a = {k: x.copy() for k, x in arr.items()}
    a["fwd_off"][v + 1], a["fwd_off"][v] = arr["fwd_off"][v], arr["fwd_off"][v + 1]
    cases["offsets"] = a
    other = pg.Context(1600, g.src.size + 100 * 6 + 16, 100, max_cites=6, seed=8)
    for name, a in cases.items():
        path = str(tmp_path / "bad.pgck")
        open(path, "wb").write(join(h, a, user))
        log("crafted checkpoint:", name)
        with pytest.raises(pg.PGError) as e:
            other.checkpoint_load(path)
        assert e.value.code == 7, name  # PG_EIO
        assert "inconsistent graph state" in str(e.value), name
    # the untouched file still loads
    assert other.checkpoint_load(good) == b"blob"
    assert other.size() == ctx.size()
