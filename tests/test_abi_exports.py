"""The C-ABI library loads and exports every function declared in include/*.h (no GPU
calls are made)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "posit_gpu_amd", "lib", "libposit_gpu_amd.so")


def declared_c_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        if 'extern "C"' not in src:
            continue  # C++-linkage headers are checked by their own test
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b(pga_[a-z0-9_]+)\s*\(", src):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB)
    names = declared_c_functions()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_status_strings_without_device():
    lib = ctypes.CDLL(LIB)
    lib.pga_status_string.restype = ctypes.c_char_p
    assert lib.pga_status_string(0) == b"success"
    assert lib.pga_status_string(1) == b"illegal argument"
    lib.pga_is_initialized.restype = ctypes.c_int
    assert lib.pga_is_initialized() == 0
