"""The instrument co-evolver on the device against the co-evolver oracle,
bit for bit: six strings (both polarisations, stiffness, geometric tension,
per-string sub-steps), the shared neck mode, per-string bridge impedances,
bridge and neck loads on a voxelised two-bout body with its plate, saddle
feedback, modal tracker, magnetic / piezo pickups and polar mics, per audio
sample (acoustic_coevolver.cpp:1910-2226).  Checked: every output ring, the
raw string and neck state, the modal amplitudes, the body's pressure field
and plate, and the step and sample counters, across plucks, scheduled
plucks, injected force buffers, the block-with-drive form, the async worker,
the damping scale, reset, and a divergent drive that must stop at the same
sample."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _instrument(dx=0.008, pad=6, n_segs=120, n_strings=6, neck=True, extra_pickups=False,
                omni=False, n_pml=4):
    from afdtd.geometry import two_bout_outline
    from afdtd.instrument import body_dict, guitar_defs
    from oracle import geometry as og
    poly = two_bout_outline(96)
    body_h = 0.1
    ct, pa, g = og.voxelise(poly, body_h, dx, pad)
    info = dict(Nx=g["Nx"], Ny=g["Ny"], Nz=g["Nz"], dx=dx, gx_min=g["gx_min"],
                gy_min=g["gy_min"], gz_min=g["gz_min"], plate_iz=g["plate_iz"])
    bridges = [(-0.02, -0.13), (0.0, -0.13), (0.02, -0.13)]
    S, P, M = guitar_defs(body_h, bridges, n_strings=n_strings, n_segs=n_segs, neck=neck)
    if extra_pickups:
        y = -0.13
        P = P + [{"type": 1, "pos": np.array([0.0, y + 0.05, body_h + 0.01], np.float32),
                  "axis": np.array([0.0, 0.0, 1.0], np.float32), "pole_sigma": 0.006,
                  "coil_spacing": 0.018, "string_mask": 0b101011},
                 {"type": 2, "pos": np.zeros(3, np.float32), "axis": np.zeros(3, np.float32),
                  "pole_sigma": 0.0, "sensitivity": 0.7, "string_mask": 0b110110},
                 {"type": 0, "pos": np.array([0.01, y + 0.1, body_h + 0.012], np.float32),
                  "axis": np.array([0.2, 0.1, 1.0], np.float32), "pole_sigma": 0.004,
                  "string_mask": 0b000111}]
    if omni:
        M = M + [{"pos": np.array([0.05, 0.02, body_h + 0.03], np.float32), "gain": 2.0}]
    B = body_dict(ct, pa, info, bridges, body_h, n_pml=n_pml)
    return S, P, M, B


def _pair(oracle_co, **kw):
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, B = _instrument(**kw)
    g = AcousticCoEvolver(S, P, M, B, 44100.0, modal_stride=8)
    o = oracle_co.OracleCoEvolver(S, P, M, B, 44100.0, modal_stride=8)
    return g, o, (S, P, M, B)


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
    P = o.fields()[0]
    assert np.array_equal(_bits(g.get_pressure_field().ravel()), _bits(P)), f"{tag}: P"
    assert np.array_equal(_bits(g.get_plate_displacement().ravel()), _bits(o.plate())), \
        f"{tag}: plate"


@pytest.fixture(scope="module")
def oracle_co():
    from oracle import coevo
    coevo.build()
    return coevo


def test_coevolver_plucks_bitwise(oracle_co):
    g, o, _ = _pair(oracle_co, extra_pickups=True, omni=True)
    print(f"coevolver: K={g.substeps} strings={g.n_strings} pickups={g.n_pickups} "
          f"mics={g.n_mics} sub={[g.string_substeps(i) for i in range(g.n_strings)]}", flush=True)
    assert g.substeps == o.substeps and g.dt_fdtd == o.dt_fdtd
    for x in (g, o):
        x.pluck_string(0, 0.2, 0.003, 32)
        x.pluck_string(4, 0.7, -0.002, 0)
        x.schedule_pluck(37, 2, 0.3, 0.002, 16)
        x.schedule_pluck(90, 5, 0.5, 0.004, 8)
    assert g.step(60) == 0 and o.step(60) == 0
    _check(g, o, 60, "block 1")
    assert g.step(75) == 0 and o.step(75) == 0
    _check(g, o, 135, "block 2")
    mic = o.get_mic_output(0, 135)
    assert np.abs(mic).max() > 0 and np.abs(o.get_pickup_output(0, 135)).max() > 0
    assert np.abs(o.neck_state()).max() > 0


def test_coevolver_forces_damping_reset(oracle_co):
    g, o, _ = _pair(oracle_co, n_segs=80)
    rng = np.random.default_rng(7)
    f = (0.05 * rng.standard_normal(50)).astype(np.float32)
    for x in (g, o):
        x.inject_string_force(1, f, 0.3)
        x.set_damping_scale(20.0)
    assert g.step(30) == 0 and o.step(30) == 0
    _check(g, o, 30, "forced, damped")
    for x in (g, o):
        x.set_damping_scale(1.0)
        x.pluck_string(3, 0.4, 0.003, 12)
    assert g.step(40) == 0 and o.step(40) == 0
    _check(g, o, 70, "undamped")
    g.reset()
    o.reset()
    _check(g, o, 16, "reset")
    for x in (g, o):
        x.clear_pluck_schedule()
        x.schedule_pluck(3, 0, 0.25, 0.002, 10)
    assert g.step(25) == 0 and o.step(25) == 0
    _check(g, o, 25, "after reset")


def test_coevolver_block_with_drive_bitwise(oracle_co):
    g, o, (S, *_r) = _pair(oracle_co, n_segs=100)
    rng = np.random.default_rng(11)
    n = 48
    blocks = [(0.2 * rng.standard_normal(n)).astype(np.float32) for _ in S]
    blocks[2][:] = 0.0
    for x in (g, o):
        x.schedule_pluck(5, 1, 0.3, 0.002, 8)          # fires, its ramp is not applied here
    out_g = g.step_block_with_drive(blocks, 0.15, n)
    rc, out_o = o.step_block_with_drive(blocks, 0.15, n)
    assert rc == 0
    assert np.array_equal(_bits(out_g), _bits(out_o))
    _check(g, o, n, "block with drive")
    out_g = g.step_block_with_drive(None, 0.15, 16)
    rc, out_o = o.step_block_with_drive(None, 0.15, 16)
    assert rc == 0 and np.array_equal(_bits(out_g), _bits(out_o))
    with pytest.raises(RuntimeError, match="n_strings mismatch"):
        g.step_block_with_drive(blocks[:2], 0.15, 8)


def test_coevolver_async_worker_bitwise(oracle_co):
    g, o, _ = _pair(oracle_co, n_segs=80)
    for x in (g, o):
        x.pluck_string(2, 0.3, 0.003, 20)
    assert g.step(10) == 0 and o.step(10) == 0
    for x in (g, o):
        x.schedule_pluck(40, 0, 0.2, 0.002, 10)
    g.step_async(150)
    assert g.wait() == 0
    assert o.step_worker(150) == 0
    p = g.get_progress()
    assert p["status_str"] == "done" and p["samples_done"] == 150 and p["samples_total"] == 150
    _check(g, o, 160, "async")
    assert g.is_running() == 0


def test_coevolver_divergence_stops_at_same_sample(oracle_co):
    g, o, _ = _pair(oracle_co, n_segs=60, n_strings=2)
    f = np.linspace(0.0, 80.0, 400).astype(np.float32)  # ramps past the small-slope limit
    for x in (g, o):
        x.pluck_string(0, 0.5, 0.001, 4)
        x.inject_string_force(1, f, 0.5)
    rc_o = o.step(400)
    assert rc_o == -3
    with pytest.raises(RuntimeError, match="CE_ERR_UNSTABLE"):
        g.step(400)
    print(f"coevolver: divergence after {o.samples_written} samples", flush=True)
    assert 0 < o.samples_written < 400
    assert g.samples_written == o.samples_written
    n = o.samples_written
    for m in range(g.n_mics):
        assert np.array_equal(_bits(g.get_mic_output(m, n)), _bits(o.get_mic_output(m, n)))
    assert np.array_equal(_bits(g.get_pickup_output(0, n)), _bits(o.get_pickup_output(0, n)))
    assert np.array_equal(_bits(g.get_pressure_field().ravel()), _bits(o.fields()[0]))


def test_coevolver_api_edges():
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, B = _instrument(n_segs=40, n_strings=2)
    with pytest.raises(RuntimeError, match="coevolver_create"):
        AcousticCoEvolver(S, [], M, B, 44100.0)          # empty pickup list -> NULL defs
    g = AcousticCoEvolver(S, P, M, B, 44100.0)
    assert (g.n_strings, g.n_pickups, g.n_mics) == (2, 1, len(M))
    assert g.get_plate_dims() == (B["Nx"], B["Ny"], B["Nx"] * B["Ny"])
    with pytest.raises(RuntimeError, match="coevolver_pluck_string error: -5"):
        g.pluck_string(7, 0.5, 0.001)
    with pytest.raises(RuntimeError, match="not an AMR coevolver"):
        g.get_amr_cell_centers()
    with pytest.raises(RuntimeError, match="get_mic_output error: -2"):
        g.get_mic_output(0, 5000)
    g.pluck_string(0, 0.5, 0.002, 8)
    g.step(12)
    v = g.get_string_velocity_n(0, 40)
    d = g.get_string_displacement_n(0, 40)
    pos = g.get_string_position_n(0, 41)
    assert v.shape == (40, 3) and d.shape == (40, 3) and pos.shape == (41, 3)
    assert np.abs(d).max() > 0
    base = np.asarray(S[0]["path_xyz"], np.float64)[:40]
    assert np.allclose(pos[:40].astype(np.float64) - base, d, rtol=0, atol=2e-8)
    with pytest.raises(RuntimeError, match="out_len mismatch"):
        g.get_string_velocity(0)
    xyz = np.array([[0.0, 0.0, 0.12], [0.02, -0.1, 0.05]], np.float32)
    assert g.sample_pressure(xyz).shape == (2,)
    vx, vy, vz = g.sample_velocity(xyz)
    Pe, vn = g.get_surface_emission(xyz, np.array([[0, 0, 1], [0, 1, 0]], np.float32))
    assert Pe.shape == (2,) and vn.shape == (2,)
