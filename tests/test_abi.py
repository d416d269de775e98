"""C-ABI checks that need no GPU: the library loads, exports every function
include/saber_hip.h declares, and fails loudly (no CPU fallback) without a
device."""
import ctypes as C
import os
import re

import pytest

from saber_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "saber_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(saber_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _ffi.lib()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"{n} declared in saber_hip.h but not exported"
        assert n in _ffi._SIGS, f"{n} has no ctypes signature"


def test_version_and_variant_table():
    lib = _ffi.lib()
    assert b"gfx950" in lib.saber_hip_version()
    import saber_amd as S
    v = S.band_variants()
    assert len(v) >= 8 and len(v) % 2 == 0
    for L, Cc, K, B in v:
        assert L in (1, 2, 4, 8, 16) and Cc >= 2 and 1 <= K <= Cc and B in (0, 1)


def test_no_device_is_an_error_not_a_fallback():
    n = C.c_int(-1)
    assert _ffi.lib().saber_hip_device_count(C.byref(n)) == 0
    if n.value > 0:
        pytest.skip("a GPU is present")
    import saber_amd as S
    with pytest.raises(S.HipError):
        S.Context(0)


def test_workload_generator_is_deterministic():
    from saber_amd import workload
    a = workload.band_regions(seed=11, n=500)
    b = workload.band_regions(seed=11, n=500)
    for k in ("qres", "qoff", "sres", "soff"):
        assert (a[k] == b[k]).all()
    assert (a["jobs"] == b["jobs"]).all()
    hw = (a["jobs"]["band_hi"] - a["jobs"]["band_lo"]) // 2
    assert set(hw.tolist()) <= {8, 16, 32, 64}
    lens = (a["qoff"][1:] - a["qoff"][:-1])
    assert lens.min() >= 20 and lens.max() <= 2000
    assert 0.3 < a["homolog"].mean() < 0.5
