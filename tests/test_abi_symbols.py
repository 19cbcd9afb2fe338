"""The C-ABI library loads and exports every entry point include/cl2qcd.h declares (CPU only)."""
import ctypes
import os

from cl2qcd_amd import abi


def test_library_built():
    assert os.path.exists(abi.LIB_PATH), "run __graft_entry__.build() first"


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(abi.LIB_PATH)
    declared = abi.header_symbols()
    assert len(declared) > 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_the_header():
    declared = set(abi.header_symbols())
    assert declared == set(abi.SIGNATURES), (declared ^ set(abi.SIGNATURES))
