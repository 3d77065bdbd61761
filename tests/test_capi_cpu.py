"""The C-ABI library loads and exports every symbol include/pfsimp.h
declares.  No compute call is made (no GPU here)."""
import ctypes
import os
import re

import pytest

from parafrost_amd import simplifier

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pfsimp.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pfs_[a-z_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(simplifier.SYMBOLS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(simplifier.lib_path())
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_error_paths_without_context():
    L = simplifier.load_library()
    assert L.pfs_version().decode().startswith("pfsimp")
    assert L.pfs_stage_count() >= 9
    # null handles are rejected, never dereferenced
    assert L.pfs_set_option(None, b"phases", b"1") == 2
    assert L.pfs_simplify(None, None) == 2
    assert L.pfs_rewind(None) == 2


def test_library_links_hip_not_oracle():
    """The product library is the HIP build and does not carry the oracle."""
    data = open(simplifier.lib_path(), "rb").read()
    assert b"amdgcn" in data or b"hip" in data.lower()
    assert b"oracle_run" not in data
