"""The remaining workloads at their specified sizes.

* Differentiation: 128^3 cells, seeded log-normal log rho / log K (sigma
  0.2), 2000 steps with checkpoints every 100 steps and recompute, 32
  receivers, loss = L2 against traces of a seeded perturbed parameter set.
  The hand-written adjoint's gradient is checked against float64 central
  differences of the GPU forward on 64 random cells (alternating log rho /
  log K) and 8 boundary cells (admittance eta), within 1e-4 relative; the
  float32 adjoint against the float64 one.  (The autograd reference through
  the float64 CPU step runs at 32^3 in test_gpu_adjoint.py.)
* The reference's own scheme on the 512 x 384 x 256 grid: order 4 at the
  reference dt, PML 10 cells, homogeneous air, 4 point sources and 64
  receivers, from a random state: fields and traces equal the float32
  oracle bit for bit after 6 steps.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _same_bits(a, b):
    return int(np.count_nonzero(np.asarray(a).view(np.uint32) != np.asarray(b).view(np.uint32)))


def test_adjoint_128cube_2000_steps_vs_central_differences():
    from afdtd import workloads as W
    from afdtd.adjoint import AdjointSolver, simulate_traces
    w = W.config5()
    shape, steps = w["shape"], w["steps"]
    N = int(np.prod(shape))
    print(f"adjoint {shape} {steps} steps, checkpoints every 100", flush=True)
    res = {}
    for dtype, tdt in (("float64", torch.float64), ("float32", torch.float32)):
        s = AdjointSolver(*shape, w["dtdx"], dtype=dtype, checkpoint_every=100)
        s.set_sources(w["src"], w["signals"])
        s.set_receivers(w["rec"])

        def dev(x, tdt=tdt):
            return torch.tensor(np.ascontiguousarray(x).ravel(), dtype=tdt, device="cuda")

        with torch.no_grad():
            target = s.forward(dev(w["m"] + w["pert"][0]), dev(w["l"] + w["pert"][1]),
                               dev(w["eta"]), steps).clone()
        mg, lg, eg = (dev(x).requires_grad_() for x in (w["m"], w["l"], w["eta"]))
        tr = simulate_traces(s, mg, lg, eg, steps)
        assert tr.shape == (steps, 32)
        loss = ((tr - target) ** 2).sum()
        loss.backward()
        res[dtype] = dict(s=s, dev=dev, target=target, loss=float(loss.detach()),
                          grads=[g.grad.double().cpu().numpy() for g in (mg, lg, eg)])
        print(f"{dtype}: loss {float(loss):.6e}", flush=True)
    r = res["float64"]
    s, dev, target = r["s"], r["dev"], r["target"]
    assert r["loss"] > 0

    def loss_of(p):
        with torch.no_grad():
            tr = s.forward(dev(p[0]), dev(p[1]), dev(p[2]), steps)
            return float(((tr - target) ** 2).sum())

    rng = np.random.default_rng(55)
    cells = rng.choice(N, 64, replace=False)
    I, J, K = np.unravel_index(np.arange(N), shape)
    n = shape[0]
    edge = np.flatnonzero((I == 0) | (I == n - 1) | (J == 0) | (J == n - 1) | (K == 0) | (K == n - 1))
    picks = [(q % 2, int(c)) for q, c in enumerate(cells)] + \
            [(2, int(c)) for c in rng.choice(edge, 8, replace=False)]
    scale = [np.abs(g).max() for g in r["grads"]]
    worst, h = 0.0, 1e-5
    for which, c in picks:
        vals = []
        for sgn in (1, -1):
            p = [w["m"].ravel().copy(), w["l"].ravel().copy(), w["eta"].ravel().copy()]
            p[which][c] += sgn * h
            vals.append(loss_of(p))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = r["grads"][which][c]
        worst = max(worst, abs(ad - fd) / max(abs(fd), 1e-3 * scale[which]))
    print(f"central differences, {len(picks)} cells: worst relative error {worst:.2e}", flush=True)
    assert worst < 1e-4
    for a, b, name in zip(res["float32"]["grads"], r["grads"], ("log_rho", "log_K", "eta")):
        err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(f"float32 vs float64 gradient {name}: rel-L2 {err:.2e}", flush=True)
        assert err < 1e-3


def test_room_512x384x256_order4_pml10_matches_oracle(oracle_mod):
    from afdtd import AcousticFDTD, workloads as W
    steps = 6
    w = W.config2_order4_pml(steps=steps)
    Nx, Ny, Nz = w.Nx, w.Ny, w.Nz
    print(f"order-4 PML10 room {Nx}x{Ny}x{Nz}, {steps} steps vs oracle", flush=True)
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    sig = np.tile(np.linspace(0.5, 1.5, steps, dtype=np.float32), (len(w.src), 1))
    g = AcousticFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.21, cell_type=ct, plate_iz=Nz // 2,
                     plate_active=pa, n_pml=10)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.21, cell_type=ct,
                              plate_iz=Nz // 2, plate_active=pa, n_pml=10, precision="f32")
    assert g.dt == o.dt
    print(f"kernel {g.kernel_family}", flush=True)
    rng = np.random.default_rng(23)
    fields = (rng.standard_normal(Nx * Ny * Nz, dtype=np.float32),
              1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz, dtype=np.float32),
              1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz, dtype=np.float32),
              1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1), dtype=np.float32))
    g.set_state(*fields)
    o.set_fields(*fields)
    del fields
    for m in (g, o):
        m.set_point_sources(w.src, sig)
        m.set_receivers(w.rec)
        m.record_receivers(steps)
    assert g.step(steps) == 0 and o.step(steps) == 0
    tg, to = g.get_receiver_traces(), o.get_receiver_traces().astype(np.float32)
    assert tg.shape == (steps, 64)
    assert _same_bits(tg, to) == 0
    Po, vxo, vyo, vzo = o.get_fields()
    for name, a, b in (("P", g.get_pressure_field().ravel(), Po), ("Vx", g.get_velocity_x(), vxo),
                       ("Vy", g.get_velocity_y(), vyo), ("Vz", g.get_velocity_z(), vzo)):
        nbad = _same_bits(a, b)
        assert nbad == 0, f"{name}: {nbad} of {a.size} values differ"
    g.close()
