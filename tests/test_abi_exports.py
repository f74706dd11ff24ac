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


def test_regcor_structures_match_the_header_layout():
    """gpunb_regcor_par: 14 doubles then 12 ints; gpunb_regcor_state: 22
    pointers (the C side asserts the same sizes, gpunb_iface.f90 declares
    the same members in the same order)."""
    import ctypes
    from nbody6ppgpu_amd.gpunb import RegcorPar, RegcorState
    assert ctypes.sizeof(RegcorPar) == 14 * 8 + 12 * 4
    assert RegcorPar.nnbopt.offset == 14 * 8
    assert ctypes.sizeof(RegcorState) == 22 * ctypes.sizeof(ctypes.c_void_p)
    assert [f for f, _ in RegcorState._fields_][-2:] == ["h", "list"]
