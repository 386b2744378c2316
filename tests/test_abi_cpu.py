"""The C-ABI library loads and exports every symbol include/hyteg_amd.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os

import hyteg_amd
from hyteg_amd import _lib


def test_library_is_built():
    assert os.path.exists(hyteg_amd.library_path()), "build with python hyteg_amd/build.py"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(hyteg_amd.library_path())
    declared = _lib.declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    declared = set(_lib.declared_symbols()) - {"hg_last_error", "hg_serve_last_error"}
    assert declared == set(_lib.SIGNATURES)


def test_no_symbol_from_the_oracle_in_the_product():
    # the product library must not contain or link the CPU oracle
    lib = ctypes.CDLL(hyteg_amd.library_path())
    assert not hasattr(lib, "or_p1_apply")
    deps = open(hyteg_amd.library_path(), "rb").read()
    assert b"libhg_oracle" not in deps


def test_error_path_without_context():
    lib = _lib.lib()
    assert lib.hg_ctx_synchronize(None) == -1
    assert lib.hg_last_error(None) == b"null context"
