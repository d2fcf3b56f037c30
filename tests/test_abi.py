"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/hitleiden.h declares (no compute calls), the Python
mirror reports the absence of a device loudly, and the generators are seeded."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hitleiden.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hl_[a-z_]+)\s*\(", text)))


def test_header_declares_api():
    syms = declared_symbols()
    for s in ("hl_create", "hl_destroy", "hl_load", "hl_apply_batch", "hl_read_partition", "hl_read_hierarchy",
              "hl_modularity", "hl_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import hit_leiden_amd as hl
    path = hl.library_path()
    assert os.path.exists(path), "libhitleiden.so not built"
    L = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_mirror_binds_all_symbols():
    import hit_leiden_amd as hl
    L = hl.lib()
    for s in declared_symbols():
        assert getattr(L, s).restype is not None or s == "hl_destroy"


def test_create_without_device_fails_loudly(monkeypatch):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    import hit_leiden_amd as hl
    with pytest.raises(hl.AccelerationError):
        hl.HitLeiden(hl.RunConfig())


def test_invalid_config_rejected():
    import hit_leiden_amd as hl
    with pytest.raises(hl.InvalidInput):
        hl.HitLeiden(hl.RunConfig(levels=0))
    with pytest.raises(hl.InvalidInput):
        hl.HitLeiden(hl.RunConfig(resolution=-1.0))


def test_config_flags_checked_before_the_device():
    """hl_config's key_cache and hub_tables take 0 or 1 and decide_rounds at
    most 2^20; anything else is an invalid config (reported before any device
    is looked for)."""
    import hit_leiden_amd as hl
    L = hl.lib()
    for field, bad in (("key_cache", 2), ("hub_tables", 2), ("decide_rounds", (1 << 20) + 1)):
        c = hl._Config()
        c.gamma, c.mode, c.levels = 1.0, 1, 2
        setattr(c, field, bad)
        h = ctypes.c_void_p()
        assert L.hl_create(ctypes.byref(c), ctypes.byref(h)) == hl.HL_INVALID_INPUT
        assert not h.value


def _header_struct(name):
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    macros = dict(re.findall(r"#define\s+(\w+)\s+(\d+)", text))
    body = re.search(r"typedef struct \{([^}]*)\}\s*" + name + ";", text).group(1)
    fields = re.findall(r"([\w]+\s*\*?)\s*(\w+)(?:\[(\w+)\])?;", body)
    return [(t.replace(" ", ""), f, int(macros.get(k, k)) if k else 1) for t, f, k in fields]


@pytest.mark.parametrize("struct,mirror", [("hl_config", "_Config"), ("hl_batch_stats", "BatchStats"),
                                           ("hl_request", "_Request"), ("hl_serve_stats", "ServeStats")])
def test_struct_layout_matches_header(struct, mirror):
    """Each ctypes mirror has its header struct's fields, in order, with the
    same sizes and the same total size, so both sides agree on every offset."""
    import hit_leiden_amd as hl
    base = {"double": 8, "int32_t": 4, "uint32_t": 4, "uint64_t": 8}
    size = {t: 8 if t.endswith("*") else base[t] for t, _, _ in _header_struct(struct)}
    want = _header_struct(struct)
    got = getattr(hl, mirror)._fields_
    assert [f for _, f, _ in want] == [f for f, _ in got]
    for (t, f, k), (_, ct) in zip(want, got):
        assert ctypes.sizeof(ct) == k * size[t], f
    # C layout of the header's fields (natural alignment) against the mirror's
    off = 0
    for t, _, k in want:
        a = size[t]
        off = (off + a - 1) // a * a + k * size[t]
    off = (off + 7) // 8 * 8
    assert ctypes.sizeof(getattr(hl, mirror)) == off


def test_graph_input_shapes():
    import hit_leiden_amd as hl
    g = hl.GraphInput.from_edges(4, [(0, 1), (1, 2, 2.5)])
    assert g.weights.tolist() == [1.0, 2.5]
    with pytest.raises(hl.InvalidInput):
        hl.GraphInput(3, [0, 1], [1])


def test_generators_are_seeded_and_exact():
    from hit_leiden_amd import gen
    a = gen.small_config(seed=7)
    b = gen.small_config(seed=7)
    assert a.n == 9877 and a.m == 21778
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v)
    assert len(a.batches) == 5 and all(len(x[0]) == 1000 for x in a.batches)
    # canonical, no self-loops, no duplicates
    assert (a.u < a.v).all()
    keys = a.u.astype(np.uint64) << 32 | a.v
    assert len(np.unique(keys)) == len(keys)
    # batch inserts are new edges
    s = set(keys.tolist())
    for u, v, al in a.batches[:1]:
        for x, y in zip(u.tolist(), v.tolist()):
            assert ((min(x, y) << 32) | max(x, y)) not in s


def test_planted_deletions_cancel_exactly():
    from hit_leiden_amd import gen
    w = gen.planted(3000, 30, 16.0, 0.3, seed=5)
    w.make_batches(2, 50, 20, mode=1, wlo=0.5, whi=1.5, seed=6)
    wt = {(int(a), int(b)): float(x) for a, b, x in zip(w.u, w.v, w.w)}
    u, v, al = w.batches[0]
    dels = al < 0
    assert dels.sum() == 20
    for a, b, x in zip(u[dels], v[dels], al[dels]):
        assert wt[(int(min(a, b)), int(max(a, b)))] + x == 0.0
