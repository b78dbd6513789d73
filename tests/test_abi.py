"""The C ABI library: it loads, it exports every declared entry point, and its
host-side checks raise the reference's error classes and messages
(src/index.cpp:550-611, 881-892) before any GPU work."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

import common
import katoann


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(common.ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(common.ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(katoann_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(katoann.LIB_PATH)
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_and_truncated_files(tmp_path):
    ix = katoann.Index()
    p = str(tmp_path / "none")
    with pytest.raises(RuntimeError, match="cannot open " + re.escape(p) + ".meta"):
        ix.load(p)
    with open(p + ".meta", "wb") as f:
        f.write(struct.pack("<2I", 8, 2))
    with pytest.raises(RuntimeError, match="truncated"):
        ix.load(p)
    with open(p + ".meta", "wb") as f:
        f.write(struct.pack("<5I", 0, 2, 4, 0, 0))
    with pytest.raises(RuntimeError, match="bad header"):
        ix.load(p)
    with open(p + ".meta", "wb") as f:
        f.write(struct.pack("<5I", 8, 2, 4, 1, 0))
    with pytest.raises(RuntimeError, match="only the L2 metric"):
        ix.load(p)
    with open(p + ".meta", "wb") as f:
        f.write(struct.pack("<5I", 8, 2, 4, 0, 1))
    with pytest.raises(RuntimeError, match="cannot read"):
        ix.load(p)
    np.zeros((2, 8), np.float32).tofile(p + ".centroids")
    with pytest.raises(RuntimeError, match="cannot open .*invlist"):
        ix.load(p)
    with open(p + ".invlist", "wb") as f:
        f.write(struct.pack("<I", 3))
    with pytest.raises(RuntimeError, match="truncated .*invlist"):
        ix.load(p)


def test_dimc_checks(tmp_path):
    from katoann import formats
    p = str(tmp_path / "ix")
    rng = np.random.default_rng(0)
    v = rng.standard_normal((6, 8)).astype(np.float32)
    formats.write_index(p, v[:2], [np.array([0, 1, 2]), np.array([3, 4, 5])], [v[:3], v[3:]], 6, False,
                        np.eye(8, dtype=np.float32))
    raw = open(p + ".dimc", "rb").read()
    open(p + ".dimc", "wb").write(struct.pack("<I", 0) + raw[4:])
    ix = katoann.Index()
    with pytest.raises(RuntimeError, match="bad header in .*dimc"):
        ix.load(p)
    open(p + ".dimc", "wb").write(raw[:8] + struct.pack("<I", 9) + raw[12:])
    with pytest.raises(RuntimeError, match="does not match the index"):
        ix.load(p)
    open(p + ".dimc", "wb").write(raw[:40])
    with pytest.raises(RuntimeError, match="truncated rotation"):
        ix.load(p)
    open(p + ".dimc", "wb").write(raw)
    with pytest.raises(ValueError, match="prefix_dims must be a multiple of 64"):
        ix.load(p, prefix_dims=96)
    with pytest.raises(ValueError, match="precision is"):
        ix.load(p, precision="fp8")


def test_search_argument_checks():
    ix = katoann.Index()
    with pytest.raises(RuntimeError, match="search before load"):
        ix.search(np.zeros((1, 0), np.float32))
    assert ix.engine == "gfx950"
    assert katoann.amx_available() is False


def test_build_argument_checks(tmp_path):
    """Index::build's checks (src/index.cpp:419-432), raised before any GPU work."""
    p = str(tmp_path / "b")
    with pytest.raises(ValueError, match="dimension must be 1 to 4096"):
        katoann.build(np.zeros((10, 4097), np.float32), p, nlist=2)
    with pytest.raises(ValueError, match="need at least nlist vectors"):
        katoann.build(np.zeros((10, 8), np.float32), p, nlist=11)
    with pytest.raises(ValueError, match="need at least nlist vectors"):
        katoann.build(np.zeros((10, 8), np.float32), p, nlist=0)
    with pytest.raises(ValueError, match="train_size is limited to 8388608"):
        katoann.build(np.zeros(((1 << 23) + 1, 1), np.float32), p, nlist=2, train_size=(1 << 23) + 1)
    assert not os.path.exists(p + ".meta")


def test_server_needs_a_loaded_index():
    with pytest.raises(RuntimeError, match="not loaded"):
        katoann.Server(katoann.Index())


def test_tuning_knob_ranges():
    """The ranking knobs reject values outside their ranges before any GPU work."""
    ix = katoann.Index()
    with pytest.raises(ValueError, match="band factor"):
        ix.set_band_factor(13)
    with pytest.raises(ValueError, match="band factor"):
        ix.set_band_factor(0)
    with pytest.raises(ValueError, match="rank kernel"):
        ix.set_rank_kernel(2, 2)
    with pytest.raises(ValueError, match="rank kernel"):
        ix.set_rank_kernel(1, 4)
    with pytest.raises(ValueError, match="queries per ranking item"):
        ix.set_rank_kernel(1, 0, 0)
    with pytest.raises(ValueError, match="seed rows"):
        ix.set_seed_rows(1 << 21)
    ix.set_band_factor(12)
    ix.set_rank_kernel(1, 3, 4096)
    ix.set_seed_rows(0)
    ix.set_pool_allowance(0)
