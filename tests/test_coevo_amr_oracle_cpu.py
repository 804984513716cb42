"""Pins of the AMR-backed co-evolver oracle (orc_coevo_create_amr in
oracle/coevo_oracle.c, reference coevolver_create_amr,
acoustic_coevolver.cpp:1387-1746, and the pressure-backend branches of the
per-sample loop, :1934-2219).  The string side is the uniform co-evolver's
(pinned in test_coevo_oracle_cpu.py) and the AMR body is the AMR oracle
(pinned in test_amr_oracle_cpu.py); these pin what joins them:

  * timing: dt_fdtd is the AMR step 0.77 min_dx / (c sqrt 3) in float32 and
    K = ceil(dt_audio / dt_fdtd);
  * set-up tables against closed forms: the descriptor's bridge and neck
    plate lists are taken as given; each string's bridge kernel is the
    normalised Gaussian of sigma 2 plate_dx over the bridge nodes' plate
    positions (origin + (i, j) plate_dx); the saddle taps interpolate any
    linear function of the plate coordinates exactly at the saddle;
  * without neck plate nodes the neck mode stays off (:1510-1512);
  * an unexcited instrument stays exactly silent (rings, body, plate);
  * the mic rings are the backend's stencil samples of the body's end-of-
    sample state, mixed as (p a + b rho c v.axis) gain, checked through an
    independent AMR oracle set to the same state;
  * reset clears the strings and rings and leaves the AMR body as it is
    (the reference resets only a uniform body, :2702).
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def co():
    from oracle import coevo
    coevo.build()
    return coevo


def _instrument(n_strings=3, n_segs=40, neck=True, neck_nodes=True, omni=True):
    from afdtd.instrument import amr_box_desc, guitar_defs
    nx, ny, nz, kp, dx = 12, 14, 12, 4, 0.01
    org = (-0.06, -0.18, 0.0)
    body_h = org[2] + kp * dx
    bridges = [(-0.01, -0.15), (0.0, -0.15), (0.01, -0.15)]
    D = amr_box_desc(nx, ny, nz, kp, dx, org, bridges,
                     neck_positions=[(0.0, 0.3)] if neck_nodes else [])
    S, P, M = guitar_defs(body_h, bridges, n_strings=n_strings, n_segs=n_segs, neck=neck,
                          mic_ring=2, mic_aperture_r=0.01)
    if omni:
        M = M + [{"pos": np.array([0.02, -0.1, body_h + 0.03], np.float32), "gain": 2.0}]
    return S, P, M, D


def test_amr_timing(co):
    S, P, M, D = _instrument()
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 48000.0)
    dt = np.float32(0.77 * D["min_dx"] / (D["c"] * math.sqrt(3.0)))
    assert o.dt_fdtd == dt
    assert o.substeps == math.ceil(np.float32(1.0) / np.float32(48000.0) / dt)


def test_amr_setup_tables_closed_forms(co):
    S, P, M, D = _instrument()
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0)
    t = o.tables()
    assert np.array_equal(t["bidx"], D["bridge_plate_idx"])
    assert np.array_equal(t["bwgt"], D["bridge_plate_wgt"])
    assert np.array_equal(t["nidx"], D["neck_plate_idx"])
    ny, pdx = D["plate_Ny"], float(D["plate_dx"])
    org = np.asarray(D["plate_origin"], np.float64)
    i, j = D["bridge_plate_idx"] // ny, D["bridge_plate_idx"] % ny
    sigma = 2.0 * pdx
    for si, s in enumerate(S):
        sx, sy = (float(v) for v in s["path_xyz"][-1, :2])
        g = np.exp(-((sx - (org[0] + i * pdx)) ** 2 + (sy - (org[1] + j * pdx)) ** 2)
                   / (2 * sigma * sigma))
        assert np.allclose(t["sbk"][si], g / g.sum(), rtol=2e-6, atol=1e-9)
        # saddle taps: bilinear interpolation reproduces a linear field exactly
        gi, gj = (sx - org[0]) / pdx, (sy - org[1]) / pdx
        f = lambda n: 3.0 + 0.5 * (n // ny) - 0.25 * (n % ny)
        val = float(np.dot(t["swgt"][si].astype(np.float64), f(t["sidx"][si])))
        assert abs(val - (3.0 + 0.5 * gi - 0.25 * gj)) < 1e-5


def test_amr_neck_off_without_neck_nodes(co):
    S, P, M, D = _instrument(neck=True, neck_nodes=False)
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0)
    o.pluck_string(0, 0.3, 0.003, 8)
    assert o.step(40) == 0
    assert not np.any(o.neck_state())
    assert np.abs(o.string_state(0)["u0"]).max() > 0
    S2, P2, M2, D2 = _instrument(neck=True, neck_nodes=True)
    o2 = co.OracleCoEvolver.from_amr(D2, S2, P2, M2, 44100.0)
    o2.pluck_string(0, 0.3, 0.003, 8)
    assert o2.step(40) == 0
    assert np.abs(o2.neck_state()).max() > 0


def test_amr_idle_silence(co):
    S, P, M, D = _instrument()
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0)
    assert o.step(60) == 0
    assert not np.any(o.fields()[0]) and not np.any(o.plate())
    for m in range(len(M)):
        assert not np.any(o.get_mic_output(m, 60))


def test_amr_mics_are_backend_stencil_samples(co):
    from oracle.amr_oracle import OracleAMR
    S, P, M, D = _instrument()
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0)
    o.pluck_string(1, 0.25, 0.004, 6)
    assert o.step(90) == 0
    p, v = o.fields()
    assert np.abs(p).max() > 0
    a = OracleAMR(D["cell_centers"], D["cell_volumes"], D["open_volume_frac"], D["cell_types"],
                  D["face_cell_neg"], D["face_cell_pos"], D["face_area"], D["face_open_frac"],
                  D["face_distance"], D["c"], D["rho_air"], D["min_dx"])
    a.set_state(p, v)
    pos = np.array([m["pos"] for m in M], np.float32)
    axis = np.array([m.get("axis", (0.0, 0.0, 1.0)) for m in M], np.float32)
    a.setup_mic_samplers(pos, axis, True)
    mp, vx, vy, vz = (x.astype(np.float64) for x in a.sample_mics())
    rhoc = float(np.float32(D["rho_air"])) * float(np.float32(D["c"]))
    for m, d in enumerate(M):
        pa, pb, gain = d.get("polar_a", 1.0), d.get("polar_b", 0.0), d.get("gain", 1.0)
        ax = np.asarray(d.get("axis", (0.0, 0.0, 1.0)), np.float64)
        want = gain * (pa * mp[m] + pb * rhoc * (vx[m] * ax[0] + vy[m] * ax[1] + vz[m] * ax[2]))
        got = float(o.get_mic_output(m, 1)[0])
        assert abs(got - want) <= 1e-5 * max(abs(want), 1e-12), (m, got, want)


def test_amr_reset_keeps_body(co):
    S, P, M, D = _instrument()
    o = co.OracleCoEvolver.from_amr(D, S, P, M, 44100.0)
    o.pluck_string(0, 0.3, 0.003, 8)
    assert o.step(50) == 0
    p0, w0 = o.fields()[0].copy(), o.plate().copy()
    assert np.abs(p0).max() > 0
    o.reset()
    assert np.array_equal(o.fields()[0], p0) and np.array_equal(o.plate(), w0)
    assert not np.any(o.string_state(0)["u0"]) and o.samples_written == 0
    assert not np.any(o.get_mic_output(0, 50)) and not np.any(o.neck_state())
