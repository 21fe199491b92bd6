"""CPU-side checks of the boundary: the HIP library loads and exports exactly what
include/sf_align.h declares (no device call is made here)."""
import ctypes as C
import os
import re

from spearfish_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "sf_align.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(N.ALIGN_LIB)
    names = header_functions()
    assert len(names) >= 18
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert sorted(N.SIGNATURES) == header_functions()


def test_abi_version_and_errors_without_device():
    lib = N.load()
    assert lib.sf_abi_version() == 1
    assert lib.sf_strerror(N.SF_EINVAL) == b"invalid argument"
    # Argument checks run before any device use.
    h = C.c_void_p()
    assert lib.sf_batch_create(None, C.byref(h)) == N.SF_EINVAL
    assert b"null" in lib.sf_last_error()


def test_oracle_is_not_linked_into_the_product():
    out = os.popen(f"ldd {N.ALIGN_LIB}").read()
    assert "oracle" not in out
    import subprocess
    syms = subprocess.run(["nm", "-D", N.ALIGN_LIB], capture_output=True, text=True).stdout
    assert "sfo_" not in syms
