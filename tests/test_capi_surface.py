"""C-ABI surface: the library loads, exports every function the header
declares, and reports errors as status codes (no GPU needed)."""
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(REPO, "include", "gswitch_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gsw_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(gsw):
    import ctypes
    lib = ctypes.CDLL(gsw.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(gsw.exported_symbols()) == declared


def test_version_and_errors(gsw):
    assert "gfx950" in gsw.version()
    with pytest.raises(gsw.GswError) as e:
        gsw.rmat(0)
    assert e.value.code == 1
    if gsw.device_count() == 0:
        h = gsw.rmat(6, seed=1)
        with pytest.raises(gsw.GswError) as e:
            gsw.Graph(h)
        assert e.value.code == 2  # no GPU: HIP error, no silent fallback


def test_configs_string_parsing_rejects_nothing(gsw):
    # unknown fields are treated as "x" (free), as prase_configs does
    o = gsw.run_opts("Push-Queue-WM-Fused")
    assert o.configs == b"Push-Queue-WM-Fused"
