"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

The FP32 kernels evaluate the reference's expressions in the same order
without FMA contraction, so fields and traces are compared BIT FOR BIT with
the float32 oracle; the looser 1e-5 relative-L2 bound of the task is checked
as well and reported in the assertion message.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5   # relative L2 bound for FP32 parity


def _pair(oracle_mod, Nx, Ny, Nz, dx=0.01, seed=0, walls=0.0, plate=False, **kw):
    from afdtd import AcousticFDTD
    rng = np.random.default_rng(seed)
    ct = np.zeros((Nx, Ny, Nz), np.uint8)
    if walls > 0:
        ct[rng.random((Nx, Ny, Nz)) < walls] = 1
    plate_iz = kw.pop("plate_iz", Nz // 2)
    pa = np.zeros((Nx, Ny), np.uint8)
    if plate:
        pa[Nx // 4: 3 * Nx // 4, Ny // 4: 3 * Ny // 4] = 1
    common = dict(c=343.0, rho_air=1.21, cell_type=ct.ravel(), plate_iz=plate_iz,
                  plate_active=pa.ravel())
    okw = dict(kw)
    gkw = dict(kw)
    if "courant" in okw and okw["courant"] is None:
        okw["courant"] = 0.0
    if "boundary" in okw:
        okw["boundary"] = {"rigid": 0, "impedance": 1}[okw["boundary"]]
    for k in ("kernel", "storage"):
        okw.pop(k, None)
    g = AcousticFDTD(Nx, Ny, Nz, dx, **common, **gkw)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, dx, **common, **okw, precision="f32")
    return g, o


def _random_state(o, g, seed, amp=1.0):
    rng = np.random.default_rng(seed + 100)
    Nx, Ny, Nz = o.Nx, o.Ny, o.Nz
    P = (amp * rng.standard_normal(Nx * Ny * Nz)).astype(np.float32)
    vx = (amp * 1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz)).astype(np.float32)
    vy = (amp * 1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz)).astype(np.float32)
    vz = (amp * 1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1))).astype(np.float32)
    ct = o.get_cell_type()
    P[(ct == 1) | (ct == 2)] = 0.0
    o.set_fields(P, vx, vy, vz)
    g.set_state(P, vx, vy, vz)


def _assert_fields_equal(g, o, what=""):
    Po, vxo, vyo, vzo = o.get_fields()
    pairs = [("P", g.get_pressure_field().ravel(), Po), ("Vx", g.get_velocity_x(), vxo),
             ("Vy", g.get_velocity_y(), vyo), ("Vz", g.get_velocity_z(), vzo)]
    for name, a, b in pairs:
        err = rel_l2(a, b)
        assert err <= TOL, f"{what} {name}: rel-L2 {err:.3e} > {TOL}"
        nbad = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
        assert nbad == 0, f"{what} {name}: {nbad} values differ bitwise (rel-L2 {err:.3e})"


@pytest.mark.parametrize("kernel", ["generic", "auto"])
@pytest.mark.parametrize("n_pml", [0, 2, 6])
def test_order4_rigid_pml_bitwise(oracle_mod, kernel, n_pml):
    g, o = _pair(oracle_mod, 21, 19, 70, n_pml=n_pml, kernel=kernel)
    assert g.dt == o.dt
    _random_state(o, g, seed=1)
    for chunk in (1, 7, 32):
        assert g.step(chunk) == 0 and o.step(chunk) == 0
        _assert_fields_equal(g, o, f"order4 pml={n_pml} after {g.step_count} steps")
    assert g.step_count == o.step_count == 40


def test_walls_plate_cells_bitwise(oracle_mod):
    g, o = _pair(oracle_mod, 18, 22, 30, walls=0.08, n_pml=3, seed=3)
    _random_state(o, g, seed=3)
    g.step(25); o.step(25)
    _assert_fields_equal(g, o, "walls")


def test_face_fractions_bitwise(oracle_mod):
    g, o = _pair(oracle_mod, 16, 17, 33, walls=0.05, n_pml=0, seed=4)
    rng = np.random.default_rng(44)
    nvx, nvy, nvz = g.get_velocity_dims()
    fr = [np.where(rng.random(n) < 0.1, 0.0, rng.uniform(0.2, 1.0, n)).astype(np.float32)
          for n in (nvx, nvy, nvz)]
    g.set_face_fractions(*fr)
    o.set_face_fractions(*fr)
    _random_state(o, g, seed=4)
    g.step(20); o.step(20)
    _assert_fields_equal(g, o, "fracs")


def test_plate_coupling_bitwise(oracle_mod):
    g, o = _pair(oracle_mod, 20, 24, 16, plate=True, n_pml=2, seed=5, plate_iz=8)
    Nz, Ny = 16, 24
    src = np.array([(9 + 1) + Nz * (10 + Ny * 9), (9 + 1) + Nz * (12 + Ny * 11)], np.int32)
    w = np.array([1.0, 0.5], np.float32)
    g.set_bridge_sources(src, w)
    o.set_bridge_sources(src, w)
    for s in range(30):
        d = np.array([np.sin(0.3 * s), np.cos(0.2 * s)], np.float32) * 2e-3
        g.inject_bridge_drive(d, 1.0)
        o.inject_bridge_drive(d, 1.0)
        g.step(1); o.step(1)
    _assert_fields_equal(g, o, "plate")
    wg = g.get_plate_displacement()
    wo = o.get_plate_displacement()
    assert np.array_equal(wg.view(np.uint32), wo.astype(np.float32).view(np.uint32))
    assert np.abs(wo).max() > 0


def test_order2_impedance_hetero_bitwise(oracle_mod):
    Nx, Ny, Nz = 20, 18, 26
    rng = np.random.default_rng(7)
    rho = (1.2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    bulk = (1.2 * 343.0 ** 2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    refl = [0.6, 0.7, 0.8, 0.9, 0.95, 0.98]
    g, o = _pair(oracle_mod, Nx, Ny, Nz, order=2, courant=0.55, boundary="impedance",
                 refl=refl, rho=rho, bulk=bulk, n_pml=0)
    assert g.dt == o.dt
    _random_state(o, g, seed=7)
    g.step(40); o.step(40)
    _assert_fields_equal(g, o, "order2 impedance hetero")


def test_sources_receivers_traces(oracle_mod):
    from afdtd import differentiated_gaussian
    g, o = _pair(oracle_mod, 32, 30, 34, n_pml=4, seed=8)
    rng = np.random.default_rng(8)
    src = [int(k + 34 * (j + 30 * i)) for i, j, k in rng.integers(6, 24, size=(2, 3))]
    sig = np.stack([differentiated_gaussian(300, 8.0, 48.0, 1.0),
                    differentiated_gaussian(300, 6.0, 30.0, 0.5)])
    rec = rng.uniform(4, 26, size=(16, 3)).astype(np.float32)
    rec[:4] = np.round(rec[:4])
    for m in (g, o):
        m.set_point_sources(src, sig)
        m.set_receivers(rec)
        m.record_receivers(300)
    g.step(300); o.step(300)
    tg, to = g.get_receiver_traces(), o.get_receiver_traces()
    assert tg.shape == to.shape == (300, 16)
    assert rel_l2(tg, to) <= TOL
    assert np.array_equal(tg.view(np.uint32), to.astype(np.float32).view(np.uint32))
    assert np.abs(to).max() > 1e-3
    _assert_fields_equal(g, o, "sources")


def test_samplers_match_oracle(oracle_mod):
    g, o = _pair(oracle_mod, 15, 13, 17, n_pml=0, seed=9)
    _random_state(o, g, seed=9)
    g.step(5); o.step(5)
    rng = np.random.default_rng(9)
    xyz = rng.uniform(-1, 18, size=(40, 3)).astype(np.float32)
    a, b = g.sample_pressure(xyz), o.sample_pressure(xyz)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    for u, v in zip(g.sample_velocity(xyz), o.sample_velocity(xyz)):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    i8, w8 = g.precompute_pressure_samplers(xyz)
    oi8, ow8 = o.precompute_pressure_samplers(xyz)
    assert np.array_equal(i8, oi8) and np.array_equal(w8, ow8)
    P = o.get_pressure_field().ravel()
    want = np.zeros(len(xyz), np.float32)
    for r in range(len(xyz)):   # reference summation order (acoustic_fdtd.cpp:1672)
        acc = w8[8 * r] * P[i8[8 * r]]
        for t in range(1, 8):
            acc = np.float32(acc + w8[8 * r + t] * P[i8[8 * r + t]])
        want[r] = acc
    got = g.sample_pressure_precomputed(i8, w8)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    vs = g.precompute_velocity_samplers(xyz)
    ovs = o.precompute_velocity_samplers(xyz)
    for u, v in zip(vs, ovs):
        assert np.array_equal(u, v)
    gv = g.sample_velocity_precomputed(*vs)
    for u, v in zip(gv, o.sample_velocity(xyz)):
        assert np.array_equal(u.view(np.uint32), v.view(np.uint32))
    nrm = rng.standard_normal((40, 3)).astype(np.float32)
    Pg, vng = g.get_surface_emission(xyz, nrm)
    ox, oy, oz = o.sample_velocity(xyz)
    vn = ox * nrm[:, 0] + oy * nrm[:, 1] + oz * nrm[:, 2]
    assert np.array_equal(Pg.view(np.uint32), b.view(np.uint32))
    assert np.array_equal(vng.view(np.uint32), vn.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("n_pml", [0, 8])
def test_config1_parity_64cube_2000_steps(oracle_mod, n_pml):
    """64^3, reference dt (Courant 0.4446), order 4, rigid / PML(8), a
    differentiated-Gaussian source (sigma 8 steps) and 16 seeded receivers,
    2000 steps: receiver traces and final fields match the oracle."""
    from afdtd import differentiated_gaussian
    N = 64
    g, o = _pair(oracle_mod, N, N, N, dx=0.01, n_pml=n_pml)
    rng = np.random.default_rng(2024)
    s = rng.integers(12, 52, size=3)
    src = [int(s[2] + N * (s[1] + N * s[0]))]
    sig = differentiated_gaussian(2000, 8.0, 48.0, 1.0)[None, :]
    rec = rng.integers(12, 52, size=(16, 3)).astype(np.float32)
    for m in (g, o):
        m.set_point_sources(src, sig)
        m.set_receivers(rec)
        m.record_receivers(2000)
    assert g.step(2000) == 0 and o.step(2000) == 0
    tg, to = g.get_receiver_traces(), o.get_receiver_traces()
    err = rel_l2(tg, to)
    print(f"config1 n_pml={n_pml}: trace rel-L2 {err:.3e}")
    assert err <= TOL
    assert np.array_equal(tg.view(np.uint32), to.astype(np.float32).view(np.uint32))
    _assert_fields_equal(g, o, f"config1 pml={n_pml}")


def test_unstable_source_reports_at_step_256(oracle_mod):
    g, o = _pair(oracle_mod, 20, 20, 20, n_pml=0)
    sig = np.full((1, 400), 1e6, np.float32)    # mean pressure passes 2e3 Pa
    src = [10 + 20 * (10 + 20 * 10)]
    for m in (g, o):
        m.set_point_sources(src, sig)
    rc_o = o.step(400)
    with pytest.raises(RuntimeError, match="FDTD_ERR_UNSTABLE"):
        g.step(400)
    assert rc_o == -3
    assert g.step_count == o.step_count == 256
    _assert_fields_equal(g, o, "unstable")


def test_reset_and_errors(oracle_mod):
    import ctypes as C
    from afdtd import lib
    g, o = _pair(oracle_mod, 10, 11, 12, n_pml=0)
    _random_state(o, g, seed=11)
    g.step(3)
    g.reset()
    assert g.step_count == 0
    assert not g.get_pressure_field().any()
    assert g.get_velocity_dims() == (9 * 11 * 12, 10 * 10 * 12, 10 * 11 * 11)
    buf = np.zeros(5, np.float32)
    L = lib()
    assert L.fdtd_get_pressure_field(g._h, buf.ctypes.data_as(C.c_void_p), 5) == -2
    assert L.fdtd_get_pressure_field(None, buf.ctypes.data_as(C.c_void_p), 5) == -1
    assert L.fdtd_inject_bridge(g._h, 1.0, 0.0, 1.0) == -3
    with pytest.raises(RuntimeError, match="disabled"):
        g.inject_bridge(1.0, 0.0)


@pytest.mark.parametrize("order,hetero,imped,n_pml", [(4, False, False, 0), (4, True, False, 5),
                                                      (2, True, True, 0), (2, False, False, 3)])
@pytest.mark.parametrize("tile_y,seg_x", [(16, 64), (8, 5), (32, 13)])
def test_fused_sweep_bitwise(oracle_mod, order, hetero, imped, n_pml, tile_y, seg_x):
    """The fused plane-marching sweep equals the oracle bit for bit on ragged
    grids (Ny, Nz not multiples of the tile, Nx not a multiple of the segment)."""
    Nx, Ny, Nz = 23, 37, 70
    kw = dict(order=order, n_pml=n_pml, kernel="fused", tile_y=tile_y, seg_x=seg_x)
    if order == 2:
        kw["courant"] = 0.5
    if imped:
        kw.update(boundary="impedance", refl=[0.6, 0.7, 0.8, 0.85, 0.9, 0.95])
    if hetero:
        rng = np.random.default_rng(12)
        kw["rho"] = (1.2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
        kw["bulk"] = (1.2 * 343.0 ** 2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    okw = {k: v for k, v in kw.items() if k not in ("tile_y", "seg_x")}
    g, o = _pair(oracle_mod, Nx, Ny, Nz, **okw)
    g.close()
    from afdtd import AcousticFDTD
    ct = np.zeros(Nx * Ny * Nz, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.21, cell_type=ct, plate_iz=Nz // 2,
                     plate_active=pa, **kw)
    assert g.kernel_family == "fused"
    _random_state(o, g, seed=13)
    for chunk in (1, 2, 9):
        g.step(chunk); o.step(chunk)
        _assert_fields_equal(g, o, f"fused o{order} het={hetero} imp={imped} ty={tile_y} seg={seg_x} n={g.step_count}")


def test_fused_unstable_restores_buffers(oracle_mod):
    """A failing stability check stops the fused sweep at step 256 with the
    same fields as the oracle (velocity double-buffer parity restored)."""
    g, o = _pair(oracle_mod, 20, 20, 20, n_pml=0, kernel="fused")
    assert g.kernel_family == "fused"
    sig = np.full((1, 401), 1e6, np.float32)
    src = [10 + 20 * (10 + 20 * 10)]
    for m in (g, o):
        m.set_point_sources(src, sig)
    assert o.step(401) == -3
    with pytest.raises(RuntimeError, match="FDTD_ERR_UNSTABLE"):
        g.step(401)
    assert g.step_count == o.step_count == 256
    _assert_fields_equal(g, o, "fused unstable")


@pytest.mark.parametrize("hetero", [False, True])
def test_energy_matches_oracle(oracle_mod, hetero):
    """fdtd_energy (FP64 device reduction) against the oracle's energy terms
    on identical fields: potential sum p^2/(2K) dV and kinetic sum
    rho_f v^2/2 dV.  The kinetic terms use the same float inputs (the fields
    are bitwise equal), so they agree to summation order; the potential term
    divides by K recovered from kappa = K dt/dx on the device (K itself in
    the oracle), a float rounding of 1e-7."""
    Nx, Ny, Nz = 19, 23, 37
    kw = dict(order=2, courant=0.5)
    if hetero:
        rng = np.random.default_rng(12)
        kw["rho"] = (1.2 * np.exp(0.2 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
        kw["bulk"] = (1.2 * 343.0 ** 2 * np.exp(0.2 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    g, o = _pair(oracle_mod, Nx, Ny, Nz, walls=0.04, n_pml=3, seed=12, **kw)
    _random_state(o, g, seed=12)
    g.step(9); o.step(9)
    _assert_fields_equal(g, o, "energy run")
    eg = g.energy()
    eo = o.energy()                 # [P^n P^(n-1)/2K, kinetic, P^2/2K]
    assert eo[2] > 0 and eo[1] > 0
    assert abs(eg[0] - eo[2]) <= 1e-6 * eo[2], (eg[0], eo[2])
    assert abs(eg[1] - eo[1]) <= 1e-12 * eo[1], (eg[1], eo[1])


def test_plate_samplers_match_oracle(oracle_mod):
    """fdtd_sample_plate_weighted and fdtd_sample_plate_displacement_batch on a
    plate driven through the bridge, against the oracle's restatement, bit
    for bit (out-of-range weighted indices are skipped by both)."""
    g, o = _pair(oracle_mod, 20, 24, 16, plate=True, n_pml=2, seed=6, plate_iz=8)
    Nz, Ny = 16, 24
    src = np.array([(9 + 1) + Nz * (10 + Ny * 9), (9 + 1) + Nz * (12 + Ny * 11)], np.int32)
    w = np.array([1.0, 0.5], np.float32)
    for m in (g, o):
        m.set_bridge_sources(src, w)
    for s in range(25):
        d = np.array([np.sin(0.3 * s), np.cos(0.2 * s)], np.float32) * 2e-3
        for m in (g, o):
            m.inject_bridge_drive(d, 1.0)
            m.step(1)
    rng = np.random.default_rng(61)
    idx = rng.integers(-5, 20 * 24 + 5, 60).astype(np.int32)
    wgt = rng.standard_normal(60).astype(np.float32)
    a, b = np.float32(g.sample_plate_weighted(idx, wgt)), np.float32(o.sample_plate_weighted(idx, wgt))
    assert a != 0 and a.view(np.uint32) == b.view(np.uint32), (a, b)
    idx4 = rng.integers(0, 20 * 24, (33, 4)).astype(np.int32)
    wgt4 = rng.uniform(0, 1, (33, 4)).astype(np.float32)
    a = g.sample_plate_displacement_batch(idx4, wgt4)
    b = o.sample_plate_displacement_batch(idx4, wgt4).astype(np.float32)
    assert np.abs(b).max() > 0
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
