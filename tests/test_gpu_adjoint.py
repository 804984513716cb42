"""The hand-written adjoint (fdtd_adjoint_*, torch.autograd.Function) against
autograd through the float64 CPU step and against central finite
differences of the GPU forward, in float64; the float32 adjoint against the
float64 one."""
import numpy as np
import pytest
import torch

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _setup(Nx, Ny, Nz, steps, seed, dtype="float64", ckpt=25):
    from afdtd.adjoint import AdjointSolver
    from afdtd.signals import differentiated_gaussian
    rng = np.random.default_rng(seed)
    shape = (Nx, Ny, Nz)
    m = np.log(1.2) + 0.2 * rng.standard_normal(shape)
    l = np.log(1.2 * 343.0 ** 2) + 0.2 * rng.standard_normal(shape)
    eta = rng.uniform(0.02, 0.5, shape)
    dtdx = 0.45 / (343.0 * np.exp(0.6))          # Courant 0.45 at the fastest plausible cell
    src = [int(Nz // 3 + Nz * (Ny // 2 + Ny * (Nx // 2)))]
    sig = differentiated_gaussian(steps, 4.0, 15.0, 1.0).astype(np.float64)[None, :]
    rec = rng.uniform(2, np.array(shape) - 3, size=(6, 3))
    s = AdjointSolver(Nx, Ny, Nz, dtdx, dtype=dtype, checkpoint_every=ckpt)
    s.set_sources(src, sig)
    s.set_receivers(rec)
    return s, m, l, eta, dtdx, src, sig, rec


def _dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a).ravel(), dtype=dtype, device="cuda")


def test_adjoint_matches_torch_autograd_32cube():
    from oracle.torch_step import forward_traces
    from afdtd.adjoint import simulate_traces, trilinear_stencils
    Nx = Ny = Nz = 32
    steps = 120
    s, m, l, eta, dtdx, src, sig, rec = _setup(Nx, Ny, Nz, steps, 1)
    rng = np.random.default_rng(2)
    target = 0.01 * rng.standard_normal((steps, len(rec)))
    # GPU adjoint
    mg, lg, eg = (_dev(a).requires_grad_() for a in (m, l, eta))
    tr = simulate_traces(s, mg, lg, eg, steps)
    loss = ((tr - torch.tensor(target, device="cuda")) ** 2).sum()
    loss.backward()
    # CPU float64 autograd reference
    mc, lc, ec = (torch.tensor(a, dtype=torch.float64).requires_grad_() for a in (m, l, eta))
    idx8, w8 = trilinear_stencils(Nx, Ny, Nz, rec)
    trc = forward_traces(mc, lc, ec, dtdx, torch.tensor(src), torch.tensor(sig),
                         torch.tensor(idx8).long(), torch.tensor(w8), steps)
    lossc = ((trc - torch.tensor(target)) ** 2).sum()
    lossc.backward()
    assert rel_l2(tr.detach().cpu().numpy(), trc.detach().numpy()) < 1e-12
    for g_gpu, g_cpu, name in ((mg.grad, mc.grad, "log_rho"), (lg.grad, lc.grad, "log_K"),
                               (eg.grad, ec.grad, "eta")):
        err = rel_l2(g_gpu.cpu().numpy(), g_cpu.numpy().ravel())
        print(f"{name}: rel-L2 vs autograd {err:.2e}")
        assert err < 1e-9, (name, err)


def test_adjoint_matches_central_differences():
    """64 random cells (log rho, log K) and 16 boundary cells (eta), float64."""
    from afdtd.adjoint import simulate_traces
    Nx, Ny, Nz, steps = 40, 36, 44, 200
    s, m, l, eta, dtdx, src, sig, rec = _setup(Nx, Ny, Nz, steps, 3, ckpt=40)
    rng = np.random.default_rng(4)
    target = torch.tensor(0.01 * rng.standard_normal((steps, len(rec))), device="cuda")

    def loss_of(mm, ll, ee):
        with torch.no_grad():
            tr = s.forward(_dev(mm), _dev(ll), _dev(ee), steps)
            return float(((tr - target) ** 2).sum())

    mg, lg, eg = (_dev(a).requires_grad_() for a in (m, l, eta))
    tr = simulate_traces(s, mg, lg, eg, steps)
    ((tr - target) ** 2).sum().backward()
    grads = {"m": mg.grad.cpu().numpy(), "l": lg.grad.cpu().numpy(), "eta": eg.grad.cpu().numpy()}
    N = Nx * Ny * Nz
    I, J, K = np.unravel_index(np.arange(N), (Nx, Ny, Nz))
    edge = np.flatnonzero((I == 0) | (I == Nx - 1) | (J == 0) | (J == Ny - 1) | (K == 0) | (K == Nz - 1))
    picks = [("m", c) for c in rng.choice(N, 32, replace=False)] + \
            [("l", c) for c in rng.choice(N, 32, replace=False)] + \
            [("eta", c) for c in rng.choice(edge, 16, replace=False)]
    worst = 0.0
    scale = {k: np.abs(v).max() for k, v in grads.items()}
    for name, c in picks:
        h = 1e-5
        base = {"m": m.ravel().copy(), "l": l.ravel().copy(), "eta": eta.ravel().copy()}
        plus = {k: v.copy() for k, v in base.items()}
        minus = {k: v.copy() for k, v in base.items()}
        plus[name][c] += h
        minus[name][c] -= h
        fd = (loss_of(plus["m"], plus["l"], plus["eta"]) - loss_of(minus["m"], minus["l"], minus["eta"])) / (2 * h)
        ad = grads[name][c]
        err = abs(ad - fd) / max(abs(fd), 1e-3 * scale[name])
        worst = max(worst, err)
    print(f"finite differences: worst relative error {worst:.2e} over {len(picks)} cells")
    assert worst < 1e-4


def test_float32_adjoint_close_to_float64_and_checkpoint_invariant():
    from afdtd.adjoint import simulate_traces
    Nx, Ny, Nz, steps = 36, 32, 40, 150
    outs = {}
    for dtype, tdt, ck in (("float64", torch.float64, 30), ("float64", torch.float64, 7),
                           ("float32", torch.float32, 30)):
        s, m, l, eta, *_ = _setup(Nx, Ny, Nz, steps, 5, dtype=dtype, ckpt=ck)
        rng = np.random.default_rng(6)
        target = torch.tensor(0.01 * rng.standard_normal((steps, 6)), device="cuda", dtype=tdt)
        mg, lg, eg = (_dev(a, tdt).requires_grad_() for a in (m, l, eta))
        tr = simulate_traces(s, mg, lg, eg, steps)
        ((tr - target) ** 2).sum().backward()
        outs[(dtype, ck)] = [g.grad.double().cpu().numpy() for g in (mg, lg, eg)]
    for a, b in zip(outs[("float64", 30)], outs[("float64", 7)]):
        assert np.array_equal(a, b)                      # recompute is deterministic
    for a, b, name in zip(outs[("float32", 30)], outs[("float64", 30)], ("m", "l", "eta")):
        err = rel_l2(a, b)
        print(f"float32 vs float64 adjoint, {name}: rel-L2 {err:.2e}")
        assert err < 1e-3
