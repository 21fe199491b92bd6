"""CPU-side checks of the boundary: the HIP libraries load and export exactly what
include/sf_align.h and include/sf_diag.h declare (no device call is made here)."""
import ctypes as C
import os
import re

from spearfish_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions(header="sf_align.h"):
    src = open(os.path.join(ROOT, "include", header)).read()
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


def test_diag_library_is_separate_and_complete():
    lib = C.CDLL(N.DIAG_LIB)
    names = header_functions("sf_diag.h")
    assert sorted(N.DIAG_SIGNATURES) == names
    for name in names:
        assert hasattr(lib, name), name
    # The alignment library carries no diagnostics or test hooks.
    align = C.CDLL(N.ALIGN_LIB)
    for name in names:
        assert not hasattr(align, name), name
    assert "fail_batch" not in open(os.path.join(ROOT, "include", "sf_align.h")).read()


def test_abi_version_and_errors_without_device():
    lib = N.load()
    assert lib.sf_abi_version() == N.ABI_VERSION == 2
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
