"""Differentiable float64 CPU restatement of the order-2 step (test
infrastructure only: tests import it as the autograd reference for the
adjoint kernels; the product never calls it).

Same model as the engine's order-2 heterogeneous step with impedance walls
(and as the C oracle, fdtd_oracle_impl.h) written with torch tensor ops:
  b = exp(-m), K = exp(l), kappa = K dt/dx,
  xi = (dt/(2dx)) eta sqrt(K b) on domain-edge cells, 0 elsewhere,
  V <- V - (b_L + b_R)/2 dt/dx (P_R - P_L),
  P <- ((1 - xi) P - kappa div V) / (1 + xi),  then point sources, then
  receivers y_n = sum_t w_t P[idx_t].
"""
from __future__ import annotations

import torch
import torch.nn.functional as F


def edge_mask(Nx, Ny, Nz):
    e = torch.zeros((Nx, Ny, Nz), dtype=torch.bool)
    e[0] = e[-1] = True
    e[:, 0] = e[:, -1] = True
    e[:, :, 0] = e[:, :, -1] = True
    return e


def forward_traces(m, l, eta, dtdx, src_idx, sig, rec_idx8, rec_w8, n_steps):
    """m, l, eta: (Nx, Ny, Nz) float64 tensors (may require grad).
    src_idx (S,) long, sig (S, n_steps) float64, rec_idx8 (R, 8) long,
    rec_w8 (R, 8) float64.  Returns traces (n_steps, R)."""
    Nx, Ny, Nz = m.shape
    b = torch.exp(-m)
    K = torch.exp(l)
    kap = K * dtdx
    xi = torch.where(edge_mask(Nx, Ny, Nz), 0.5 * dtdx * eta * torch.sqrt(K * b),
                     torch.zeros_like(m))
    A = (1.0 - xi) / (1.0 + xi)
    E = kap / (1.0 + xi)
    bx = 0.5 * dtdx * (b[:-1] + b[1:])
    by = 0.5 * dtdx * (b[:, :-1] + b[:, 1:])
    bz = 0.5 * dtdx * (b[:, :, :-1] + b[:, :, 1:])
    P = torch.zeros_like(m)
    Vx = torch.zeros((Nx - 1, Ny, Nz), dtype=m.dtype)
    Vy = torch.zeros((Nx, Ny - 1, Nz), dtype=m.dtype)
    Vz = torch.zeros((Nx, Ny, Nz - 1), dtype=m.dtype)
    out = []
    for n in range(n_steps):
        Vx = Vx - bx * (P[1:] - P[:-1])
        Vy = Vy - by * (P[:, 1:] - P[:, :-1])
        Vz = Vz - bz * (P[:, :, 1:] - P[:, :, :-1])
        D = (F.pad(Vx, (0, 0, 0, 0, 0, 1)) - F.pad(Vx, (0, 0, 0, 0, 1, 0))
             + F.pad(Vy, (0, 0, 0, 1)) - F.pad(Vy, (0, 0, 1, 0))
             + F.pad(Vz, (0, 1)) - F.pad(Vz, (1, 0)))
        P = A * P - E * D
        flat = P.reshape(-1).index_add(0, src_idx, sig[:, n])
        P = flat.reshape(Nx, Ny, Nz)
        out.append((flat[rec_idx8] * rec_w8).sum(1))
    return torch.stack(out)
