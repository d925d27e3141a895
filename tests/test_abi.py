"""The C-ABI library loads and exports every symbol include/*.h declares (CPU only:
no compute call is made)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))
LIB = ROOT / "mrmath_amd" / "libmrmath_gpu.so"


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"(?m)^(?!#)[A-Za-z_][\w \t\*]*?\b(mrm_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ["mrm_mult", "mrm_blocked_mult", "mrm_blocked_mult_t1", "mrm_blocked_mult_t2",
                 "mrm_dev_dgemm", "mrm_last_error"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "build the library first (__graft_entry__.build())"
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_library_loads_and_binding_covers_header():
    from mrmath_amd import _lib
    lib = _lib.load()
    for n in declared_functions():
        assert hasattr(lib, n)
        assert n in _lib.SIGNATURES, f"{n} missing from the ctypes binding"
    assert lib.mrm_version() > 0


def test_exports_are_c_linkage():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    mangled = [l for l in out.splitlines() if "_Z" in l and "mrm_" in l]
    assert not mangled
