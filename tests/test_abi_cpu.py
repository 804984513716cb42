"""C-ABI checks that need no GPU: the library loads, exports every function
include/afdtd.h declares (the reference's 26 plus the extensions), the
options struct layout agrees with ctypes, and the argument checks that
precede any device call return the reference's codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afdtd.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"SK_API\s+[\w\s\*]*?\b(fdtd_\w+)\s*\(", txt)))


def test_header_declares_the_26_reference_functions():
    from afdtd import REFERENCE_SYMBOLS
    names = _declared()
    assert len(REFERENCE_SYMBOLS) == 26
    missing = [n for n in REFERENCE_SYMBOLS if n not in names]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    from afdtd import LIB_PATH
    assert os.path.exists(LIB_PATH), "build libafdtd.so first"
    L = C.CDLL(LIB_PATH)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing


def test_options_struct_layout_matches():
    from afdtd import _lib
    o = _lib.default_options()
    assert o.struct_size == C.sizeof(_lib.FdtdOptions)
    assert o.stencil_order == 4 and o.storage == 0 and list(o.refl) == [1.0] * 6


def test_null_and_argument_codes_without_device():
    from afdtd import lib
    L = lib()
    assert L.fdtd_step(None, 1) == -1
    assert L.fdtd_reset(None) == -1
    assert L.fdtd_get_step_count(None) == 0
    assert L.fdtd_get_dt(None) == 0.0
    assert L.fdtd_inject_bridge(None, 0.0, 0.0, 1.0) == -3          # disabled API
    assert L.fdtd_set_bridge_sources(None, 1, None, None) == -1
    a, b, c = C.c_int(7), C.c_int(7), C.c_int(7)
    L.fdtd_get_velocity_dims(None, C.byref(a), C.byref(b), C.byref(c))
    assert (a.value, b.value, c.value) == (0, 0, 0)
    ct = np.zeros(8, np.uint8)
    # invalid sizes / null arrays fail before touching the device (NULL handle)
    assert not L.fdtd_create(0, 2, 2, 0.01, 343.0, 1.21, ct.ctypes.data_as(C.c_void_p), 1,
                             ct.ctypes.data_as(C.c_void_p), 7.0, 0.45, 2.0, 1e-5, 0)
    assert not L.fdtd_create(2, 2, 2, 0.01, 343.0, 1.21, None, 1, None, 7.0, 0.45, 2.0, 1e-5, 0)


def test_python_class_size_checks_match_reference_messages():
    from afdtd import AcousticFDTD
    with pytest.raises(RuntimeError, match="cell_type: expected Nx\\*Ny\\*Nz elements"):
        AcousticFDTD(2, 2, 2, 0.01, cell_type=np.zeros(7, np.uint8), plate_iz=1,
                     plate_active=np.zeros(4, np.uint8))
    with pytest.raises(RuntimeError, match="plate_active: expected Nx\\*Ny elements"):
        AcousticFDTD(2, 2, 2, 0.01, cell_type=np.zeros(8, np.uint8), plate_iz=1,
                     plate_active=np.zeros(3, np.uint8))


def test_drop_in_module_names():
    import _spectral_kernels as sk
    assert (sk.FDTD_AIR, sk.FDTD_WALL, sk.FDTD_PLATE, sk.FDTD_PML) == (0, 1, 2, 3)
    assert sk.AcousticFDTD.__name__ == "AcousticFDTD"
