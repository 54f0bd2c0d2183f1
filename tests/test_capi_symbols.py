"""The C-ABI library loads on a machine without a GPU and exports every function that
include/feat3_hip.h declares (no compute calls are made here)."""
import ctypes
import os

from feat3_amd import capi


def test_header_parses_and_all_symbols_exported():
    decls = capi.parse_header()
    assert len(decls) > 90
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in decls if not hasattr(lib, n)]
    assert not missing, missing


def test_error_strings():
    assert capi.lib.feat3_hip_error_string(0) == b"success"
    assert capi.lib.feat3_hip_error_string(1) == b"invalid argument"


def test_argument_checks_without_gpu():
    # argument validation happens before any HIP call
    assert capi.lib.feat3_mgpcg_create(None, 0, 0) == 1
    assert capi.lib.feat3_mgpcg_init(None) == 4
    assert capi.lib.lafem_hip_apply_csr_f64_u32(None, None, 1.0, None, 0.0, None, None, None, None, 1, 1, 1, 0) == 1


def test_library_is_in_tree():
    assert os.path.dirname(capi.LIB_PATH) == os.path.dirname(os.path.abspath(capi.__file__))
