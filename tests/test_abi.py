"""The C ABI library loads (no GPU needed) and exports every entry point declared in
include/sds_hip.h; the Python bindings cover them all."""
import ctypes
import os
import re

from sparsedirectsolver import _abi
from sparsedirectsolver._native import HIP_SO, hip_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "sds_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(sds_\w+)\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "sds_factorize" in names and "sds_solve" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(HIP_SO)
    for name in declared():
        assert hasattr(lib, name), name


def test_bindings_cover_the_header():
    assert sorted(_abi.EXPORTS) == declared()
    hip_lib()  # binds every signature without calling into the GPU


def test_calls_without_a_device_fail_cleanly():
    lib = hip_lib()
    assert lib.sds_destroy(None) == _abi.SDS_SUCCESS
    assert lib.sds_upload_symbolic(None, None) == _abi.SDS_INVALID_VALUE
