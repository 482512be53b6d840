"""The C ABI library loads and exports every symbol include/bsx.h declares
(no GPU needed: nothing here launches work)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "bsx.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(bsx_[a-z0-9_]+)\s*\(", hdr)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("bsx_extend_batch", "bsx_extend_batch_device", "bsx_global_batch", "bsx_global_batch_device",
              "bsx_ksw_extend2", "bsx_create", "bsx_sync"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from bsbolt_amd import bsx
    L = ctypes.CDLL(bsx.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(bsx.EXPORTS)


def test_version_string():
    from bsbolt_amd import bsx
    assert b"gfx950" in bsx.lib().bsx_version()


def test_library_has_no_oracle_dependency():
    # the product library must not link or embed the CPU oracle
    data = open(os.path.join(ROOT, "bsbolt_amd", "libbsx.so"), "rb").read()
    assert b"orc_extend2" not in data and b"liboracle" not in data


def test_bin_shapes_match_library():
    from bsbolt_amd import bsx
    lim = bsx.ext_bin_limits()
    assert list(lim) == [l * c for l, c in bsx.EXT_SHAPES]
