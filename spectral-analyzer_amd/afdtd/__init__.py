"""MI355X acoustic FDTD engine (HIP kernels behind a C ABI, ctypes front end)."""
from ._lib import (FDTD_AIR, FDTD_WALL, FDTD_PLATE, FDTD_PML, FDTD_OK, FDTD_ERR_NULL,
                   FDTD_ERR_DIM, FDTD_ERR_UNSTABLE, LIB_PATH, REFERENCE_SYMBOLS, lib)
from .solver import AcousticFDTD
from .signals import differentiated_gaussian
from .batch import RoomBatch

__all__ = ["AcousticFDTD", "RoomBatch", "differentiated_gaussian", "lib", "LIB_PATH", "REFERENCE_SYMBOLS",
           "FDTD_AIR", "FDTD_WALL", "FDTD_PLATE", "FDTD_PML", "FDTD_OK", "FDTD_ERR_NULL",
           "FDTD_ERR_DIM", "FDTD_ERR_UNSTABLE"]
