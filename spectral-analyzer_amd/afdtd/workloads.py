"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

Conventions recorded here are the ones the benchmark and the tests use:
  * heterogeneous media: g ~ N(0,1) per cell, smoothed by a 3x3x3 box
    filter (scipy.ndimage.uniform_filter, mode 'nearest'), renormalised to
    unit variance; rho = rho0 exp(0.1 g_rho), K = rho0 c0^2 exp(0.1 g_K) with
    two independent fields;
  * sources: differentiated Gaussian, sigma = 8 steps, n0 = 48, 1 Pa, at
    seeded interior cells (margin 8 cells);
  * receivers: seeded interior cell centres (margin 8 cells).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .signals import differentiated_gaussian

RHO0 = 1.2
C0 = 343.0


@dataclass
class Workload:
    name: str
    Nx: int
    Ny: int
    Nz: int
    dx: float
    steps: int
    order: int
    courant: Optional[float]
    boundary: str
    refl: List[float]
    n_pml: int = 0
    rho: Optional[np.ndarray] = None
    bulk: Optional[np.ndarray] = None
    src: List[int] = field(default_factory=list)
    rec: Optional[np.ndarray] = None
    sigma: float = 8.0
    n0: float = 48.0

    @property
    def cells(self) -> int:
        return self.Nx * self.Ny * self.Nz

    @property
    def hetero(self) -> bool:
        return self.rho is not None or self.bulk is not None

    @property
    def bytes_per_cell_update(self) -> int:
        """Algorithmic HBM bytes per cell-update of the FP32 step: P, Vx, Vy,
        Vz read and written once (32 B) plus one read of rho and K (8 B)
        when the medium is heterogeneous."""
        return self.bytes_per_cell_update_for("f32")

    def bytes_per_cell_update_for(self, storage: str) -> int:
        """Same count for a field storage type: 8 field accesses of 4 B
        (f32) or 2 B (bf16, f16); per-cell rho and K stay FP32."""
        elem = 4 if storage == "f32" else 2
        return 8 * elem + (8 if self.hetero else 0)

    def signals(self, n: int) -> np.ndarray:
        return np.stack([differentiated_gaussian(n, self.sigma, self.n0, 1.0)
                         for _ in self.src]) if self.src else np.zeros((0, n), np.float32)

    def solver_kwargs(self) -> dict:
        return dict(order=self.order, courant=self.courant, boundary=self.boundary,
                    refl=self.refl, rho=self.rho, bulk=self.bulk)


def _smoothed_lognormal_field(rng, shape, sigma: float) -> np.ndarray:
    from scipy.ndimage import uniform_filter
    g = rng.standard_normal(shape, dtype=np.float32)
    g = uniform_filter(g, size=3, mode="nearest")
    g /= max(float(g.std()), 1e-12)
    return np.exp(sigma * g).astype(np.float32)


def _interior_cells(rng, Nx, Ny, Nz, n, margin=8):
    i = rng.integers(margin, Nx - margin, n)
    j = rng.integers(margin, Ny - margin, n)
    k = rng.integers(margin, Nz - margin, n)
    return i, j, k


def config1(seed: int = 1, parity: bool = False) -> Workload:
    """64^3, dx = 1 cm, air (rho 1.2, c 343).  As written: order 2, Courant
    0.5, impedance walls R = 0.9.  parity=True: the reference's own scheme
    (order 4, reference dt, rigid walls)."""
    rng = np.random.default_rng(seed)
    N = 64
    i, j, k = _interior_cells(rng, N, N, N, 1)
    ri, rj, rk = _interior_cells(rng, N, N, N, 16)
    w = Workload("config1-64cube" + ("-parity" if parity else ""), N, N, N, 0.01, 2000,
                 4 if parity else 2, None if parity else 0.5,
                 "rigid" if parity else "impedance", [1.0] * 6 if parity else [0.9] * 6)
    w.src = [int(k[0] + N * (j[0] + N * i[0]))]
    w.rec = np.stack([ri, rj, rk], 1).astype(np.float32)
    return w


def config2(seed: int = 2, Nx=512, Ny=384, Nz=256, steps=4000) -> Workload:
    """Single large room: 512x384x256, order 2, Courant 0.55, heterogeneous
    rho and K, per-side R ~ U[0.6, 0.98], 4 sources, 64 receivers."""
    rng = np.random.default_rng(seed)
    shape = (Nx, Ny, Nz)
    rho = (RHO0 * _smoothed_lognormal_field(rng, shape, 0.1)).ravel()
    bulk = (RHO0 * C0 * C0 * _smoothed_lognormal_field(rng, shape, 0.1)).ravel()
    refl = [float(x) for x in rng.uniform(0.6, 0.98, 6)]
    w = Workload(f"config2-room-{Nx}x{Ny}x{Nz}-hetero-order2", Nx, Ny, Nz, 0.01, steps, 2,
                 0.55, "impedance", refl, rho=rho, bulk=bulk)
    i, j, k = _interior_cells(rng, Nx, Ny, Nz, 4)
    w.src = [int(k[q] + Nz * (j[q] + Ny * i[q])) for q in range(4)]
    ri, rj, rk = _interior_cells(rng, Nx, Ny, Nz, 64)
    w.rec = np.stack([ri, rj, rk], 1).astype(np.float32)
    return w


def config2_order4_pml(seed: int = 2, Nx=512, Ny=384, Nz=256, steps=4000) -> Workload:
    """The config-2 grid with the reference scheme: homogeneous air, order 4,
    reference dt, PML 10 cells."""
    rng = np.random.default_rng(seed)
    w = Workload(f"room-{Nx}x{Ny}x{Nz}-order4-pml10", Nx, Ny, Nz, 0.01, steps, 4, None,
                 "rigid", [1.0] * 6, n_pml=10)
    i, j, k = _interior_cells(rng, Nx, Ny, Nz, 4, margin=16)
    w.src = [int(k[q] + Nz * (j[q] + Ny * i[q])) for q in range(4)]
    ri, rj, rk = _interior_cells(rng, Nx, Ny, Nz, 64, margin=16)
    w.rec = np.stack([ri, rj, rk], 1).astype(np.float32)
    return w


def config4(seed: int = 4, Nx=1024, Ny=1024, Nz=512, steps=3000, full=(1024, 1024, 512)) -> Workload:
    """Precision workload: homogeneous air, order 2 at Courant 0.5, walls with
    R = 0.95 on every side, 2 sources and 128 receivers.  Source and receiver
    positions are drawn as fractions of the full 1024x1024x512 room and mapped
    proportionally, so a smaller grid is the crop-and-rescale of the same
    seeded set-up (same dx; convention used for the FP64 comparison)."""
    rng = np.random.default_rng(seed)
    fs = rng.uniform(0.2, 0.8, size=(2, 3))
    fr = rng.uniform(0.1, 0.9, size=(128, 3))
    w = Workload(f"precision-room-{Nx}x{Ny}x{Nz}-order2", Nx, Ny, Nz, 0.01, steps, 2, 0.5,
                 "impedance", [0.95] * 6)
    dims = np.array([Nx, Ny, Nz])
    si = np.minimum((fs * dims).astype(int), dims - 1)
    w.src = [int(k + Nz * (j + Ny * i)) for i, j, k in si]
    w.rec = (fr * (dims - 1)).astype(np.float32)
    return w


def config3(seed: int = 3, n_rooms=256, lo=48, hi=160, steps=1000):
    """Batched rooms: dimensions uniform per axis in [lo, hi], per-room
    Courant in [0.4, 0.55] and reflection coefficient in [0.6, 0.98] (all six
    sides), one differentiated-Gaussian source and 8 receivers per room."""
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(n_rooms, 3))
    courant = rng.uniform(0.4, 0.55, n_rooms).astype(np.float32)
    refl = np.repeat(rng.uniform(0.6, 0.98, (n_rooms, 1)), 6, axis=1).astype(np.float32)
    src, rec = [], np.zeros((n_rooms, 8, 3), np.float32)
    for r, (Nx, Ny, Nz) in enumerate(dims):
        i, j, k = (int(rng.integers(8, d - 8)) for d in (Nx, Ny, Nz))
        src.append(k + int(Nz) * (j + int(Ny) * i))
        rec[r] = rng.uniform(8, np.array([Nx, Ny, Nz]) - 9, size=(8, 3))
    sig = differentiated_gaussian(steps, 8.0, 48.0, 1.0)
    return dict(dims=dims, courant=courant, refl=refl, src=src, rec=rec, steps=steps,
                signals=np.repeat(sig[None, :], n_rooms, axis=0))


def config5(seed: int = 5, N=128, steps=2000):
    """Differentiation: N^3 cells, log rho and log K seeded log-normal with
    sigma 0.2, per-boundary-cell admittance sum eta, 32 receivers, targets
    from a seeded perturbed parameter set."""
    rng = np.random.default_rng(seed)
    shape = (N, N, N)
    m = np.log(RHO0) + 0.2 * rng.standard_normal(shape)
    l = np.log(RHO0 * C0 * C0) + 0.2 * rng.standard_normal(shape)
    eta = np.full(shape, 2 * (1 - 0.9) / (1 + 0.9))
    pert = [0.05 * rng.standard_normal(shape), 0.05 * rng.standard_normal(shape)]
    cmax = float(np.exp(0.5 * (l - m)).max())
    dtdx = 0.5 / cmax
    src = [int(N // 3 + N * (N // 2 + N * (N // 2)))]
    rec = rng.uniform(8, N - 9, size=(32, 3))
    sig = differentiated_gaussian(steps, 8.0, 48.0, 1.0).astype(np.float64)[None, :]
    return dict(shape=shape, m=m, l=l, eta=eta, pert=pert, dtdx=dtdx, src=src, rec=rec,
                steps=steps, signals=sig)
