"""Differentiable FDTD traces: ``torch.autograd.Function`` over the hand-written
adjoint kernels of libafdtd.so (``fdtd_adjoint_*``).

Parameters are per-cell torch tensors on the GPU: ``log_rho``, ``log_K`` and
``eta`` (admittance sum of the domain sides a boundary cell touches,
sum of (1-R)/(1+R); ignored on interior cells).  The forward stores
checkpoints every ``checkpoint_every`` steps; the backward recomputes each
segment and runs the time-reversed adjoint, so memory is
O(N * (steps/K + 4K)) instead of O(N * steps).

The model is the engine's order-2 step for heterogeneous media with
impedance walls; torch tensors are used only as device buffers.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib


def trilinear_stencils(Nx, Ny, Nz, rec_xyz, dtype=np.float64):
    """8-tap (index, weight) stencils of fdtd_precompute_pressure_samplers
    (acoustic_fdtd.cpp:1575-1594), in the requested precision."""
    xyz = np.asarray(rec_xyz, dtype).reshape(-1, 3)
    n = len(xyz)
    idx = np.zeros((n, 8), np.int32)
    w = np.zeros((n, 8), dtype)
    for r, (gx, gy, gz) in enumerate(xyz):
        gx = min(max(gx, 0), Nx - 1); gy = min(max(gy, 0), Ny - 1); gz = min(max(gz, 0), Nz - 1)
        i0, j0, k0 = int(gx), int(gy), int(gz)
        i1 = i0 + 1 if i0 < Nx - 1 else i0
        j1 = j0 + 1 if j0 < Ny - 1 else j0
        k1 = k0 + 1 if k0 < Nz - 1 else k0
        fx, fy, fz = gx - i0, gy - j0, gz - k0
        cs = [(i0, j0, k0), (i1, j0, k0), (i0, j1, k0), (i1, j1, k0),
              (i0, j0, k1), (i1, j0, k1), (i0, j1, k1), (i1, j1, k1)]
        ws = [(1 - fx) * (1 - fy) * (1 - fz), fx * (1 - fy) * (1 - fz), (1 - fx) * fy * (1 - fz),
              fx * fy * (1 - fz), (1 - fx) * (1 - fy) * fz, fx * (1 - fy) * fz,
              (1 - fx) * fy * fz, fx * fy * fz]
        for t, ((i, j, k), wt) in enumerate(zip(cs, ws)):
            idx[r, t] = k + Nz * (j + Ny * i)
            w[r, t] = wt
    return idx, w


class AdjointSolver:
    """Owns one fdtd_adjoint handle (grid, dt/dx, precision, checkpoints)."""

    def __init__(self, Nx, Ny, Nz, dt_over_dx, dtype="float32", checkpoint_every=100):
        import torch
        self._L = lib()
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.N = self.Nx * self.Ny * self.Nz
        self.torch_dtype = {"float32": torch.float32, "float64": torch.float64}[dtype]
        self.np_dtype = {"float32": np.float32, "float64": np.float64}[dtype]
        self.dt_over_dx = float(dt_over_dx)
        h = self._L.fdtd_adjoint_create(self.Nx, self.Ny, self.Nz, self.dt_over_dx,
                                        8 if dtype == "float64" else 4, int(checkpoint_every))
        if not h:
            raise RuntimeError(f"fdtd_adjoint_create failed: {_lib.last_error()}")
        self._h = h
        self.n_rec = 0
        self._keep = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.fdtd_adjoint_destroy(h)
            self._h = None

    def set_sources(self, cell_idx, signals):
        idx = np.ascontiguousarray(np.asarray(cell_idx, np.int32).ravel())
        sig = np.ascontiguousarray(np.asarray(signals, self.np_dtype).reshape(idx.size, -1))
        rc = self._L.fdtd_adjoint_set_sources(self._h, idx.size, idx.ctypes.data_as(C.c_void_p),
                                              sig.ctypes.data_as(C.c_void_p), sig.shape[1])
        if rc:
            raise RuntimeError(f"fdtd_adjoint_set_sources error {rc}")

    def set_receivers(self, rec_xyz):
        idx, w = trilinear_stencils(self.Nx, self.Ny, self.Nz, rec_xyz, self.np_dtype)
        idx = np.ascontiguousarray(idx)
        w = np.ascontiguousarray(w)
        self.n_rec = idx.shape[0]
        rc = self._L.fdtd_adjoint_set_receivers(self._h, self.n_rec,
                                                idx.ctypes.data_as(C.c_void_p),
                                                w.ctypes.data_as(C.c_void_p))
        if rc:
            raise RuntimeError(f"fdtd_adjoint_set_receivers error {rc}")

    def _check(self, t):
        import torch
        if not (t.is_cuda and t.dtype == self.torch_dtype and t.is_contiguous() and t.numel() == self.N):
            raise ValueError("parameters must be contiguous CUDA tensors of the solver's dtype, one value per cell")

    def forward(self, log_rho, log_K, eta, n_steps):
        import torch
        for t in (log_rho, log_K, eta):
            self._check(t)
        traces = torch.empty((int(n_steps), self.n_rec), device=log_rho.device, dtype=self.torch_dtype)
        torch.cuda.synchronize()
        rc = self._L.fdtd_adjoint_forward(self._h, log_rho.data_ptr(), log_K.data_ptr(),
                                          eta.data_ptr(), int(n_steps), traces.data_ptr())
        if rc:
            raise RuntimeError(f"fdtd_adjoint_forward error {rc}: {_lib.last_error()}")
        self._keep = (log_rho, log_K, eta)      # must outlive the backward
        return traces

    def backward(self, grad_traces):
        import torch
        g = grad_traces.contiguous().to(self.torch_dtype)
        lr = self._keep[0]
        gm, gl, ge = (torch.empty_like(lr) for _ in range(3))
        torch.cuda.synchronize()
        rc = self._L.fdtd_adjoint_backward(self._h, g.data_ptr(), gm.data_ptr(), gl.data_ptr(),
                                           ge.data_ptr())
        if rc:
            raise RuntimeError(f"fdtd_adjoint_backward error {rc}: {_lib.last_error()}")
        return gm, gl, ge


def _function():
    import torch

    class FDTDTraces(torch.autograd.Function):
        """traces = F(log_rho, log_K, eta); backward by the adjoint kernels."""

        @staticmethod
        def forward(ctx, log_rho, log_K, eta, solver, n_steps):
            ctx.solver = solver
            return solver.forward(log_rho.detach(), log_K.detach(), eta.detach(), n_steps)

        @staticmethod
        def backward(ctx, grad_traces):
            gm, gl, ge = ctx.solver.backward(grad_traces)
            return gm, gl, ge, None, None

    return FDTDTraces


_FN = None


def simulate_traces(solver: AdjointSolver, log_rho, log_K, eta, n_steps: int):
    """Differentiable receiver traces (n_steps x n_rec)."""
    global _FN
    if _FN is None:
        _FN = _function()
    return _FN.apply(log_rho.contiguous(), log_K.contiguous(), eta.contiguous(), solver, int(n_steps))
