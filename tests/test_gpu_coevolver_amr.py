"""The AMR-backed instrument co-evolver (coevolver_create_amr,
acoustic_coevolver.cpp:1387-1746, and the pressure-backend branches of the
per-sample loop, :1934-2219) on the device against the co-evolver oracle
built on the AMR oracle, bit for bit: output rings, raw string and neck
state, modal amplitudes, the AMR body's cell pressures and the plate, across
plucks, forces, the block-drive form, the async worker, a PML border, the
order-8 gradient, reset (which leaves the AMR body as it is) and a divergent
drive."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _instrument(nx=20, ny=24, nz=16, kp=5, dx=0.01, n_segs=80, n_strings=6, neck=True,
                n_pml=0, border_mode=0, border_R=0.0, gradient_order=2, omni=False,
                neck_nodes=True):
    from afdtd.instrument import amr_box_desc, guitar_defs
    org = (-0.10, -0.20, 0.0)
    body_h = org[2] + kp * dx
    bridges = [(-0.02, -0.13), (0.0, -0.13), (0.02, -0.13)]
    D = amr_box_desc(nx, ny, nz, kp, dx, org, bridges,
                     neck_positions=[(0.0, 0.5)] if neck_nodes else [], n_pml=n_pml,
                     border_mode=border_mode, border_R=border_R, gradient_order=gradient_order)
    S, P, M = guitar_defs(body_h, bridges, n_strings=n_strings, n_segs=n_segs, neck=neck,
                          mic_ring=4)
    P = P + [{"type": 2, "pos": np.zeros(3, np.float32), "axis": np.zeros(3, np.float32),
              "pole_sigma": 0.0, "sensitivity": 0.7, "string_mask": 0b110110}]
    if omni:
        M = M + [{"pos": np.array([0.03, -0.05, body_h + 0.04], np.float32), "gain": 2.0}]
    return S, P, M, D


def _pair(oracle_co, **kw):
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, D = _instrument(**kw)
    g = AcousticCoEvolver.from_amr_desc(D, S, P, M, 44100.0, modal_stride=8)
    o = oracle_co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0, modal_stride=8)
    return g, o, (S, P, M, D)


def _check(g, o, n_out, tag):
    assert g.samples_written == o.samples_written, tag
    for p in range(g.n_pickups):
        assert np.array_equal(_bits(g.get_pickup_output(p, n_out)),
                              _bits(o.get_pickup_output(p, n_out))), f"{tag}: pickup {p}"
    for m in range(g.n_mics):
        assert np.array_equal(_bits(g.get_mic_output(m, n_out)),
                              _bits(o.get_mic_output(m, n_out))), f"{tag}: mic {m}"
    for si in range(g.n_strings):
        a, b = g.string_state(si), o.string_state(si)
        for k in b:
            assert np.array_equal(_bits(np.atleast_1d(a[k])), _bits(np.atleast_1d(b[k]))), \
                f"{tag}: string {si} {k}"
        ra, ia = g.get_modal_amplitudes(si)
        rb, ib = o.modal(si)
        assert np.array_equal(_bits(ra), _bits(rb)) and np.array_equal(_bits(ia), _bits(ib)), \
            f"{tag}: modal {si}"
    assert np.array_equal(_bits(g.neck_state()), _bits(o.neck_state())), f"{tag}: neck"
    assert np.array_equal(_bits(g.get_amr_pressure()), _bits(o.fields()[0])), f"{tag}: P"
    assert np.array_equal(_bits(g.get_plate_displacement().ravel()), _bits(o.plate())), \
        f"{tag}: plate"


@pytest.fixture(scope="module")
def oracle_co():
    from oracle import coevo
    coevo.build()
    return coevo


def test_amr_coevolver_plucks_bitwise(oracle_co):
    g, o, _ = _pair(oracle_co, omni=True)
    print(f"amr coevolver: K={g.substeps} strings={g.n_strings} pickups={g.n_pickups} "
          f"mics={g.n_mics}", flush=True)
    assert g.substeps == o.substeps and g.dt_fdtd == o.dt_fdtd
    for x in (g, o):
        x.pluck_string(0, 0.2, 0.003, 32)
        x.pluck_string(4, 0.7, -0.002, 0)
        x.schedule_pluck(37, 2, 0.3, 0.002, 16)
    assert g.step(60) == 0 and o.step(60) == 0
    _check(g, o, 60, "block 1")
    assert g.step(140) == 0 and o.step(140) == 0
    _check(g, o, 200, "block 2")
    assert np.abs(o.get_mic_output(0, 200)).max() > 0 and np.abs(o.neck_state()).max() > 0
    assert np.abs(o.plate()).max() > 0


@pytest.mark.parametrize("kw", [dict(n_pml=2, border_mode=1, border_R=0.5),
                                dict(gradient_order=8, neck_nodes=False)])
def test_amr_coevolver_border_and_order8(oracle_co, kw):
    g, o, _ = _pair(oracle_co, n_segs=60, **kw)
    print(f"amr coevolver: {kw}", flush=True)
    for x in (g, o):
        x.pluck_string(1, 0.3, 0.003, 16)
        x.inject_string_force(3, (0.05 * np.sin(np.arange(40) * 0.3)).astype(np.float32), 0.4)
    assert g.step(90) == 0 and o.step(90) == 0
    _check(g, o, 90, str(kw))
    if not kw.get("neck_nodes", True):
        assert not np.any(o.neck_state())                  # no neck plate nodes: neck off


def test_amr_coevolver_drive_async_reset(oracle_co):
    g, o, (S, *_r) = _pair(oracle_co, n_segs=60)
    rng = np.random.default_rng(5)
    blocks = [(0.2 * rng.standard_normal(40)).astype(np.float32) for _ in S]
    out_g = g.step_block_with_drive(blocks, 0.15, 40)
    rc, out_o = o.step_block_with_drive(blocks, 0.15, 40)
    assert rc == 0 and np.array_equal(_bits(out_g), _bits(out_o))
    _check(g, o, 40, "block with drive")
    for x in (g, o):
        x.schedule_pluck(50, 0, 0.2, 0.002, 10)
    g.step_async(70)
    assert g.wait() == 0
    assert o.step_worker(70) == 0
    _check(g, o, 110, "async")
    P_before = g.get_amr_pressure()
    g.reset()
    o.reset()
    _check(g, o, 16, "reset")
    assert np.array_equal(_bits(g.get_amr_pressure()), _bits(P_before))   # body kept
    for x in (g, o):
        x.pluck_string(5, 0.4, 0.003, 12)
    assert g.step(30) == 0 and o.step(30) == 0
    _check(g, o, 30, "after reset")


def test_amr_coevolver_divergence_stops_at_same_sample(oracle_co):
    g, o, _ = _pair(oracle_co, n_segs=60, n_strings=2)
    f = np.linspace(0.0, 80.0, 400).astype(np.float32)
    for x in (g, o):
        x.pluck_string(0, 0.5, 0.001, 4)
        x.inject_string_force(1, f, 0.5)
    assert o.step(400) == -3
    with pytest.raises(RuntimeError, match="CE_ERR_UNSTABLE"):
        g.step(400)
    print(f"amr coevolver: divergence after {o.samples_written} samples", flush=True)
    assert 0 < o.samples_written < 400
    n = o.samples_written
    assert g.samples_written == n
    for m in range(g.n_mics):
        assert np.array_equal(_bits(g.get_mic_output(m, n)), _bits(o.get_mic_output(m, n)))
    assert np.array_equal(_bits(g.get_amr_pressure()), _bits(o.fields()[0]))
    assert np.array_equal(_bits(g.get_plate_displacement().ravel()), _bits(o.plate()))


def test_amr_coevolver_api_edges():
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, D = _instrument(n_segs=40, n_strings=2)
    bad = dict(D, n_bridge_plate=0)
    with pytest.raises(RuntimeError, match=r"\[DESC_INVALID_REQUIRED_FIELD\]: bridge plate "
                                           r"mapping is missing or empty \(stage: "
                                           r"pressure_backend_create\)"):
        AcousticCoEvolver.from_amr_desc(bad, S, P, M, 44100.0)
    g = AcousticCoEvolver.from_amr_desc(D, S, P, M, 44100.0)
    assert g.get_plate_dims() == (D["plate_Nx"], D["plate_Ny"], D["plate_Nx"] * D["plate_Ny"])
    cc = g.get_amr_cell_centers()
    assert cc.shape == (D["n_cells"], 3)
    assert np.allclose(cc, np.asarray(D["cell_centers"], np.float32))
    g.pluck_string(0, 0.5, 0.002, 8)
    g.step(20)
    field = g.get_pressure_field()            # scattered onto the max(4 min_dx, 5 mm) grid
    viz = max(4 * D["min_dx"], 0.005)
    shape = tuple(max(1, min(int(round((D["bounds_max"][q] - D["bounds_min"][q]) / viz)) + 1, 256))
                  for q in range(3))
    assert field.shape == shape and np.all(np.isfinite(field))
    xyz = np.array([[0.0, 0.0, 0.1]], np.float32)
    with pytest.raises(RuntimeError, match="sample_pressure error: -4"):
        g.sample_pressure(xyz)
    with pytest.raises(RuntimeError, match="sample_velocity error: -1"):
        g.sample_velocity(xyz)
    with pytest.raises(RuntimeError, match="coevolver_set_face_fractions error: -1"):
        g.set_face_fractions(np.ones(4), np.ones(4), np.ones(4))
