"""CPU-side checks of libfastrecon_hip.so: it loads, exports every symbol
the C header declares, and its host helpers (KB LUT, trajectory
normalisation) match the golden fixture and the oracle bit for bit.
No GPU compute calls."""
import ctypes
import os
import re

import numpy as np
import pytest

from fastrecon_amd import hip
from fastrecon_amd import synthetic as syn

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = hip.load()
    names = _declared("fastrecon_hip.h")
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(hip.EXPORTS) == names


def test_kb_lut_matches_golden():
    g = np.load(os.path.join(ROOT, "tests", "golden", "kb_lut_golden.npz"))
    for key, w, og, L in [("w4_og2_L512", 4.0, 2.0, 512), ("w4_og1p5_L512", 4.0, 1.5, 512),
                          ("w4_og2_L32", 4.0, 2.0, 32), ("w3_og2_L512", 3.0, 2.0, 512)]:
        lut, beta = hip.kb_lut(w, og, L)
        np.testing.assert_array_equal(lut.view(np.int32), g[key].view(np.int32))
        assert np.float32(beta) == g[key + "_beta"][0]


def test_normalize_matches_oracle(orc):
    for traj, N in [(syn.radial2d(50, 128)[0], 256), (syn.kooshball3d(300, 64)[0], 128)]:
        a, ba = hip.normalize_traj(traj, N)
        b, bb = orc.normalize_traj(traj, N)
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
        np.testing.assert_array_equal(ba, bb)


def test_invalid_arguments_rejected_on_host():
    L = hip.load()
    out = np.zeros(4, np.float32)
    assert L.fr_kb_lut(4.0, 2.0, 1, out.ctypes.data_as(ctypes.c_void_p), None) == 1
    assert L.fr_kb_lut(-1.0, 2.0, 4, out.ctypes.data_as(ctypes.c_void_p), None) == 1
    assert L.fr_traj_normalize(None, 10, 2, 2, 64, None) == 1
    assert L.fr_ctx_create(0, None) == 1


def test_python_sources_compile():
    """Every Python file of the package, the trainer, bench and entry point
    byte-compiles (catches errors in modules no CPU test imports)."""
    files = [os.path.join(ROOT, f) for f in ("bench.py", "__graft_entry__.py")]
    for d in ("fastrecon_amd", "oracle", "tests"):
        for dp, _, fs in os.walk(os.path.join(ROOT, d)):
            files += [os.path.join(dp, f) for f in fs if f.endswith(".py")]
    for f in files:
        with open(f) as fh:
            compile(fh.read(), f, "exec")


def test_checkpoint_metadata():
    """The shipped learned-DCF checkpoint loads without pickle and carries
    the row count and training recipe the bench reports."""
    import json
    z = np.load(os.path.join(ROOT, "fastrecon_amd", "ldcf_kooshball.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    assert isinstance(meta["rows"], int) and meta["rows"] > 0
    assert meta["steps"] > 0 and meta["losses"][-1][1] > 0
    assert z["params"].dtype == np.float32 and z["params"].size == 8 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def test_tune_cache_round_trip(tmp_path):
    """The autotune cache file through the C ABI (host only): a miss on a
    missing file, store and hit, a second key, replacement of a key's line,
    a key differing only in M misses, malformed lines are skipped."""
    lib = hip.load()
    path = str(tmp_path / "tune.txt")

    def key(M=13107200, C=32):
        return hip.fr_tune_key(arch=b"gfx950", cus=256, dim=3, N=256, M=M, C=C, type=0, frames=0, kwidth=4.0, L=512)

    assert hip.tune_cache_lookup(lib, path, key()) is None
    a = dict(tile_x=4, tile_y=2, tile_z=2, slabs=0, accumulate=6, split_taps=1024, exact_order=0,
             items_per_wave=2, xcd_run=1024, stores=2, waves_per_simd=4, tile_writes=0)
    hip.tune_cache_store(lib, path, key(), a, 4.5, 5.25)
    assert hip.tune_cache_lookup(lib, path, key()) == (a, 4.5, 5.25)
    b = dict(a, tile_x=8, tile_y=2, tile_z=1, accumulate=5, items_per_wave=0)
    hip.tune_cache_store(lib, path, key(C=16), b, 9.0, 9.5)
    a2 = dict(a, xcd_run=64)
    hip.tune_cache_store(lib, path, key(), a2, 4.25, 5.0)
    assert hip.tune_cache_lookup(lib, path, key()) == (a2, 4.25, 5.0)
    assert hip.tune_cache_lookup(lib, path, key(C=16)) == (b, 9.0, 9.5)
    assert hip.tune_cache_lookup(lib, path, key(M=13107201)) is None
    lines = open(path).read().splitlines()
    assert sum("C=32" in l for l in lines) == 1 and sum("C=16" in l for l in lines) == 1
    with open(path, "a") as f:
        f.write("garbage line without separators\n" + lines[-1].split(" | ")[0] + " | 1 2 | best_ms=1 default_ms=1\n")
    assert hip.tune_cache_lookup(lib, path, key(C=16)) == (b, 9.0, 9.5)
