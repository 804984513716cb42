"""Drop-in for the reference module ``_spectral_kernels`` on the FDTD path.

The reference bridge imports ``from _spectral_kernels import AcousticFDTD``
(acoustic_fdtd_bridge.py:81-86), ``AcousticCoEvolver`` (:88-95) and the
FDTD_* / PICKUP_* constants (pybind_kernels.cpp:9208-9216); with this
directory on ``sys.path`` the same imports resolve to the MI355X engine.
"""
from afdtd import AcousticFDTD  # noqa: F401

FDTD_AIR = 0
FDTD_WALL = 1
FDTD_PLATE = 2
FDTD_PML = 3

# The AMR stepper (pybind_kernels.cpp:8877-8920), gradient orders 2 and 8.
from afdtd.amr import AcousticAMR  # noqa: E402,F401

# The instrument co-evolver on the uniform body (pybind_kernels.cpp:9024-9190)
# and its module constants (:9208-9210).
from afdtd.coevolver import (AcousticCoEvolver, PICKUP_HUMBUCKER, PICKUP_PIEZO,  # noqa: E402,F401
                             PICKUP_SINGLE_COIL)
