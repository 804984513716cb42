"""``RoomBatch`` — many independent rooms of uneven size in one sweep.

One fused launch per step covers every tile of every room (for order 2,
one launch of the two-step sweep per two steps); the n-step sequence of a
call is captured in a hipGraph and launched once (fdtd_batch_* in
include/afdtd.h).  No reference counterpart: the
reference steps one ``AcousticFDTD`` handle at a time.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import lib

_STORAGE = {"f32": 0, "bf16": 1, "f16": 2}
_KERNEL = {"auto": _lib.FDTD_KERNEL_AUTO, "fused": _lib.FDTD_KERNEL_FUSED,
           "fused2": _lib.FDTD_KERNEL_FUSED2}
_FAMILY = {_lib.FDTD_KERNEL_FUSED: "fused", _lib.FDTD_KERNEL_FUSED2: "fused2"}


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class RoomBatch:
    def __init__(self, dims: Sequence[Sequence[int]], dx: float, courant: Sequence[float],
                 refl: Optional[np.ndarray] = None, c: float = 343.0, rho_air: float = 1.2,
                 order: int = 2, storage: str = "f32", tile_y: int = 0, seg_x: int = 0,
                 kernel: str = "auto"):
        self._L = lib()
        self._h = None
        self.dims = np.ascontiguousarray(np.asarray(dims, np.int32).reshape(-1, 3))
        self.n = len(self.dims)
        cr = np.ascontiguousarray(np.asarray(courant, np.float32).ravel())
        rf = None if refl is None else np.ascontiguousarray(np.asarray(refl, np.float32).reshape(self.n, 6))
        h = self._L.fdtd_batch_create(self.n, _p(self.dims), float(dx), float(c), float(rho_air),
                                      _p(cr), _p(rf), int(order), _STORAGE[storage],
                                      int(tile_y), int(seg_x))
        if not h:
            raise RuntimeError(f"fdtd_batch_create failed: {_lib.last_error()}")
        self._h = h
        self.per_room = 0
        self._cap = 0
        if kernel != "auto":
            rc = self._L.fdtd_batch_set_kernel(h, _KERNEL[kernel])
            if rc:
                raise RuntimeError(f"fdtd_batch_set_kernel error {rc}: {_lib.last_error()}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.fdtd_batch_destroy(h)
            self._h = None

    @property
    def total_cells(self) -> int:
        t = C.c_longlong()
        self._L.fdtd_batch_cells(self._h, C.byref(t), None)
        return int(t.value)

    @property
    def kernel_family(self) -> str:
        """The sweep the next steps use: "fused2" (two steps per launch) or
        "fused"."""
        return _FAMILY[int(self._L.fdtd_batch_get_kernel_family(self._h))]

    @property
    def step_count(self) -> int:
        return int(self._L.fdtd_batch_get_step_count(self._h))

    def autotune(self, trial_steps: int = 10) -> dict:
        """Pick the sweep, tile height, x-segment and dispatch order by timing
        every candidate on this batch (before the first step; fields stay
        zero)."""
        ty, sg, so, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        rc = self._L.fdtd_batch_autotune(self._h, int(trial_steps), C.byref(ty), C.byref(sg),
                                         C.byref(so), C.byref(ms))
        if rc:
            raise RuntimeError(f"fdtd_batch_autotune error {rc}: {_lib.last_error()}")
        return {"kernel": self.kernel_family, "tile_y": ty.value, "seg_x": sg.value,
                "schedule": ("room order", "largest first")[so.value],
                "ms_per_step": ms.value}

    def set_sources(self, cell_idx, signals):
        idx = np.ascontiguousarray(np.asarray(cell_idx, np.int32).ravel())
        sig = np.ascontiguousarray(np.asarray(signals, np.float32).reshape(self.n, -1))
        rc = self._L.fdtd_batch_set_sources(self._h, _p(idx), _p(sig), int(sig.shape[1]))
        if rc:
            raise RuntimeError(f"fdtd_batch_set_sources error {rc}")

    def set_receivers(self, xyz):
        a = np.ascontiguousarray(np.asarray(xyz, np.float32).reshape(self.n, -1, 3))
        self.per_room = a.shape[1]
        rc = self._L.fdtd_batch_set_receivers(self._h, self.per_room, _p(a))
        if rc:
            raise RuntimeError(f"fdtd_batch_set_receivers error {rc}")

    def record(self, max_steps: int):
        rc = self._L.fdtd_batch_record(self._h, int(max_steps))
        if rc:
            raise RuntimeError(f"fdtd_batch_record error {rc}")
        self._cap = int(max_steps)

    def step(self, n_steps: int, graph: bool = True) -> float:
        """Advance all rooms; returns the device time of the call in ms."""
        ms = C.c_float()
        rc = self._L.fdtd_batch_step(self._h, int(n_steps), 1 if graph else 0, C.byref(ms))
        if rc:
            raise RuntimeError(f"fdtd_batch_step error {rc}: {_lib.last_error()}")
        return float(ms.value)

    def get_traces(self) -> np.ndarray:
        out = np.empty(max(self._cap * self.n * self.per_room, 1), np.float32)
        n = C.c_int()
        rc = self._L.fdtd_batch_get_traces(self._h, _p(out), out.size, C.byref(n))
        if rc:
            raise RuntimeError(f"fdtd_batch_get_traces error {rc}")
        return out[: n.value * self.n * self.per_room].reshape(n.value, self.n, self.per_room).copy()

    def get_pressure(self, room: int) -> np.ndarray:
        Nx, Ny, Nz = (int(v) for v in self.dims[room])
        out = np.empty((Nx, Ny, Nz), np.float32)
        rc = self._L.fdtd_batch_get_pressure(self._h, int(room), _p(out), out.size)
        if rc:
            raise RuntimeError(f"fdtd_batch_get_pressure error {rc}")
        return out
