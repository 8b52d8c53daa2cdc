"""CPU checks of the C ABI: the library loads and exports every symbol the
header declares; no compute call is made without a GPU."""
import ctypes
import os
import re

from gpu_intensities_amd import binding

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(name="gpu_intensities.h"):
    txt = open(os.path.join(ROOT, "include", name)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gi_[a-z_0-9]+)\s*\(", txt)))


def test_library_exists_and_loads():
    assert os.path.exists(binding.LIB_PATH), "run __graft_entry__.build() first"
    ctypes.CDLL(binding.LIB_PATH)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(binding.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 16
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(binding.EXPORTS) == syms


def test_library_is_gfx950_code_object():
    data = open(binding.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_trove_header_symbols_exported():
    lib = ctypes.CDLL(binding.LIB_PATH)
    syms = header_symbols("gi_trove.h")
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(binding.TROVE_EXPORTS) == syms


def test_driver_built():
    exe = os.path.join(ROOT, "gpu-intensities_amd", "gi_intens")
    assert os.path.exists(exe) and os.access(exe, os.X_OK), "run __graft_entry__.build() first"
