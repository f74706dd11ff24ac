"""The C-ABI library loads and exports every entry point include/gpunb.h
declares (no compute call: runs without a GPU)."""
import ctypes
import os
import re

from nbody6ppgpu_amd import gpunb

HEADER = os.path.join(os.path.dirname(os.path.dirname(__file__)), "include",
                      "gpunb.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(gpu\w+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_reference_abi():
    names = declared_functions()
    for ref_name in ("gpunb_devinit_", "gpunb_open_", "gpunb_close_",
                     "gpunb_send_", "gpunb_regf_", "gpunb_profile_", "gpupot_"):
        assert ref_name in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(gpunb.library_path()), "build the library first"
    lib = ctypes.CDLL(gpunb.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_front_end_binds_every_symbol():
    lib = gpunb.load()
    for n in declared_functions():
        assert getattr(lib, n).argtypes is not None or n in (), n
