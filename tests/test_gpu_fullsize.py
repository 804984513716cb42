"""Checks at the workloads' full sizes.

* 512 x 384 x 256 room (heterogeneous media, impedance walls, 4 sources and
  64 receivers), default sweep (two steps per launch): fields and receiver
  traces after 12 steps from a random state equal the float32 oracle bit for
  bit (the oracle runs the full grid at about 1 G cell-updates/s on the
  host's cores).
* 1024 x 1024 x 512 room: the two-step sweep equals the single-step sweep bit
  for bit after 6 steps from a random state (the oracle would need ~30 GB of
  host memory at this size, so the two tilings check each other on the
  GPU), and scaling the state by 2 scales the result exactly (every
  operation of the scheme commutes with an exact power-of-two scale).
* 1024 x 1024 x 512 room, 16-bit storage: receiver traces of BF16, FP16 and
  bfp16 storage against FP32 storage on the full grid (600 steps: the pulse
  reaches 49 of the 128 receivers), with bounds set by each format's
  mantissa; the FP64 comparison runs on the crop (test_gpu_precision.py).
* 256 batched rooms (281.9 M cells): the two-step batch equals the
  single-step batch bit for bit (traces of every room, fields of the first
  and last room), and the first room equals the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _random_fields(rng, Nx, Ny, Nz, amp=1.0):
    P = amp * rng.standard_normal(Nx * Ny * Nz, dtype=np.float32)
    vx = amp * 1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz, dtype=np.float32)
    vy = amp * 1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz, dtype=np.float32)
    vz = amp * 1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1), dtype=np.float32)
    return P, vx, vy, vz


def _same_bits(a, b):
    return int(np.count_nonzero(np.asarray(a).view(np.uint32) != np.asarray(b).view(np.uint32)))


def test_room_512x384x256_matches_oracle(oracle_mod):
    from afdtd import AcousticFDTD, workloads as W
    steps = 12
    w = W.config2(steps=steps)
    Nx, Ny, Nz = w.Nx, w.Ny, w.Nz
    print(f"full-size room {Nx}x{Ny}x{Nz}, {steps} steps vs oracle", flush=True)
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    sig = np.tile(np.linspace(0.5, 1.5, steps, dtype=np.float32), (len(w.src), 1))
    g = AcousticFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct, plate_iz=Nz // 2,
                     plate_active=pa, n_pml=0, **w.solver_kwargs())
    assert g.kernel_family == "fused2"
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct,
                              plate_iz=Nz // 2, plate_active=pa, n_pml=0, order=2,
                              courant=w.courant, boundary=1, refl=w.refl, rho=w.rho,
                              bulk=w.bulk, precision="f32")
    assert g.dt == o.dt
    fields = _random_fields(np.random.default_rng(22), Nx, Ny, Nz)
    g.set_state(*fields)
    o.set_fields(*fields)
    del fields
    for m in (g, o):
        m.set_point_sources(w.src, sig)
        m.set_receivers(w.rec)
        m.record_receivers(steps)
    g.step(steps)
    o.step(steps)
    tg, to = g.get_receiver_traces(), o.get_receiver_traces().astype(np.float32)
    assert tg.shape == (steps, 64)
    assert _same_bits(tg, to) == 0
    Po, vxo, vyo, vzo = o.get_fields()
    for name, a, b in (("P", g.get_pressure_field().ravel(), Po), ("Vx", g.get_velocity_x(), vxo),
                       ("Vy", g.get_velocity_y(), vyo), ("Vz", g.get_velocity_z(), vzo)):
        nbad = _same_bits(a, b)
        assert nbad == 0, f"{name}: {nbad} of {a.size} values differ"
    g.close()


def test_room_1024x1024x512_two_step_equals_single_step_and_scales():
    from afdtd import AcousticFDTD, workloads as W
    steps = 6
    w = W.config4(steps=steps)
    Nx, Ny, Nz = w.Nx, w.Ny, w.Nz
    print(f"full-size room {Nx}x{Ny}x{Nz}, {steps} steps, two-step vs single-step", flush=True)
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    fields = _random_fields(np.random.default_rng(44), Nx, Ny, Nz)
    out = {}
    for kernel, scale in (("fused2", 1.0), ("fused", 1.0), ("fused2", 2.0)):
        g = AcousticFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct, plate_iz=Nz // 2,
                         plate_active=pa, n_pml=0, kernel=kernel, **w.solver_kwargs())
        assert g.kernel_family == kernel
        g.set_state(*(f * np.float32(scale) for f in fields))
        g.step(steps)
        out[(kernel, scale)] = (g.get_pressure_field().ravel(), g.get_velocity_x())
        g.close()
        del g
    del fields
    for a, b, name in zip(out[("fused2", 1.0)], out[("fused", 1.0)], ("P", "Vx")):
        nbad = _same_bits(a, b)
        assert nbad == 0, f"two-step vs single-step {name}: {nbad} values differ"
    for a, b, name in zip(out[("fused2", 2.0)], out[("fused2", 1.0)], ("P", "Vx")):
        nbad = _same_bits(a, b * np.float32(2.0))
        assert nbad == 0, f"scaled state {name}: {nbad} values differ from 2x"
    assert np.abs(out[("fused2", 1.0)][0]).max() > 0


def test_room_1024x1024x512_16bit_storage_vs_f32():
    from afdtd import AcousticFDTD, workloads as W
    from conftest import rel_l2
    steps = 600
    w = W.config4(steps=steps)
    Nx, Ny, Nz = w.Nx, w.Ny, w.Nz
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    sig = w.signals(steps)
    out = {}
    for st in ("f32", "bf16", "f16", "bfp16"):
        print(f"full-size room {Nx}x{Ny}x{Nz}, {steps} steps, storage {st}", flush=True)
        g = AcousticFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct, plate_iz=Nz // 2,
                         plate_active=pa, n_pml=0, storage=st, **w.solver_kwargs())
        g.set_point_sources(w.src, sig)
        g.set_receivers(w.rec)
        g.record_receivers(steps)
        g.step(steps)
        out[st] = (g.get_receiver_traces().astype(np.float64), float(sum(g.energy())))
        g.close()
        del g
    ref, e_ref = out["f32"]
    assert np.abs(ref).max() > 0 and e_ref > 0
    bound = {"bfp16": 1e-3, "f16": 5e-3, "bf16": 5e-2}
    for st, tol in bound.items():
        err = rel_l2(out[st][0], ref)
        de = abs(out[st][1] - e_ref) / e_ref
        print(f"{st:6s} vs f32, full size: receiver rel-L2 {err:.3e}, energy rel. diff {de:.3e}",
              flush=True)
        assert err <= tol
        assert de <= tol


def test_batch_256_rooms_two_step_equals_single_step(oracle_mod):
    from afdtd import RoomBatch, differentiated_gaussian, workloads as W
    steps = 10
    w = W.config3(steps=steps)
    dims = w["dims"]
    n = len(dims)
    sig = np.repeat(differentiated_gaussian(steps, 1.5, 3.0, 1.0)[None, :], n, axis=0)
    print(f"full-size batch {n} rooms, {int(np.prod(dims, axis=1).sum())} cells, {steps} steps",
          flush=True)
    out = {}
    for kernel in ("fused2", "fused"):
        b = RoomBatch(dims, 0.01, w["courant"], w["refl"], c=343.0, rho_air=1.2, order=2,
                      tile_y=16, seg_x=32, kernel=kernel)
        assert b.kernel_family == kernel
        b.set_sources(w["src"], sig)
        b.set_receivers(w["rec"])
        b.record(steps)
        b.step(steps)
        out[kernel] = (b.get_traces(), b.get_pressure(0), b.get_pressure(n - 1))
        del b
    for a, c, name in zip(out["fused2"], out["fused"], ("traces", "room 0", f"room {n - 1}")):
        nbad = _same_bits(a, c)
        assert nbad == 0, f"two-step vs single-step batch, {name}: {nbad} values differ"
    assert np.abs(out["fused2"][0]).max() > 0
    Nx, Ny, Nz = (int(v) for v in dims[0])
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.2, n_pml=0, order=2,
                              courant=float(w["courant"][0]), boundary=1, refl=w["refl"][0],
                              precision="f32")
    o.set_point_sources([w["src"][0]], sig[:1])
    o.set_receivers(w["rec"][0])
    o.record_receivers(steps)
    o.step(steps)
    assert _same_bits(out["fused2"][0][:, 0, :], o.get_receiver_traces().astype(np.float32)) == 0
    assert _same_bits(out["fused2"][1], o.get_pressure_field()) == 0
