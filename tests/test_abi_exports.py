"""The C-ABI library loads and exports every function declared in include/*.h (no GPU
calls are made)."""
import ctypes
import glob
import os
import re

import pytest

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


@pytest.mark.parametrize("cname,pyname", [("pga_factor_task", "FactorTask"), ("pga_server_config", "ServerConfig"),
                                           ("pga_request", "Request"), ("pga_request_result", "RequestResult"),
                                           ("pga_server_stats", "ServerStats")])
def test_struct_layout_matches_binding(tmp_path, cname, pyname):
    """Each ABI struct as the C compiler lays it out equals its ctypes mirror."""
    import subprocess

    import posit_gpu_amd as pga

    cls = getattr(pga, pyname)
    fields = [f for f, _ in cls._fields_]
    src = tmp_path / "layout.c"
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "posit_gpu_amd.h"\nint main(void){\n'
                   f'  printf("%zu", sizeof({cname}));\n' +
                   "".join(f'  printf(" %zu", offsetof({cname}, {f}));\n' for f in fields) +
                   "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(cls)] + [getattr(cls, f).offset for f in fields]
    assert got == want
