"""Pin of the torch float64 CPU step used as the autograd reference for the
adjoint: its forward traces equal the C oracle (float64) on the same
heterogeneous medium with impedance walls (no GPU needed)."""
import numpy as np
import torch

from conftest import rel_l2


def test_torch_step_matches_c_oracle(oracle_mod):
    from oracle.torch_step import forward_traces
    from afdtd.adjoint import trilinear_stencils
    Nx, Ny, Nz, steps = 14, 12, 16, 60
    rng = np.random.default_rng(21)
    rho = (1.2 * np.exp(0.2 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    bulk = (1.2 * 343.0 ** 2 * np.exp(0.2 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    refl = [0.6, 0.7, 0.8, 0.85, 0.9, 0.95]
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.2, n_pml=0, order=2,
                              courant=0.45, boundary=1, refl=refl, rho=rho, bulk=bulk,
                              precision="f64")
    src = [5 + Nz * (6 + Ny * 7)]
    sig = np.zeros((1, steps), np.float32)
    sig[0, :10] = np.sin(np.arange(10) * 0.7).astype(np.float32)
    rec = rng.uniform(1, 10, size=(5, 3)).astype(np.float32)
    o.set_point_sources(src, sig)
    o.set_receivers(rec)
    o.record_receivers(steps)
    o.step(steps)
    want = o.get_receiver_traces()
    dtdx = float(np.float64(np.float32(o.dt)) / np.float64(np.float32(0.01)))
    g = [(1 - float(np.float32(R))) / (1 + float(np.float32(R))) for R in refl]  # float32 inputs
    I, J, K = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    eta = (g[0] * (I == 0) + g[1] * (I == Nx - 1) + g[2] * (J == 0) + g[3] * (J == Ny - 1)
           + g[4] * (K == 0) + g[5] * (K == Nz - 1)).astype(np.float64)
    m = torch.tensor(np.log(rho.astype(np.float64))).reshape(Nx, Ny, Nz)
    l = torch.tensor(np.log(bulk.astype(np.float64))).reshape(Nx, Ny, Nz)
    idx8, w8 = trilinear_stencils(Nx, Ny, Nz, rec)
    got = forward_traces(m, l, torch.tensor(eta), dtdx, torch.tensor(src),
                         torch.tensor(sig, dtype=torch.float64), torch.tensor(idx8).long(),
                         torch.tensor(w8), steps)
    assert np.abs(want).max() > 1e-6
    assert rel_l2(got.numpy(), want) < 1e-10
