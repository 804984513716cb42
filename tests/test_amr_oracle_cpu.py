"""Pins of the AMR (CSR, gradient order 2) CPU oracle, and the AMR C-ABI
checks that need no GPU.

The reference's AMR stepper (csrc/kernels/acoustic_amr.cpp) needs Eigen and
cannot be built here, and its tests hold no golden vectors for it.  The
oracle is pinned instead by:
  * the hand-derived one-step answer on the 8-cell line of
    test_amr_throughput.py:26-64 (pulse in cell 0);
  * exact flux conservation (sum of p*V*open) on a random irregular topology;
  * reduction to the independently pinned uniform Yee oracle on a box grid;
  * the border condition's special cases (R = 1 is the identity, graded
    absorption reflects less than a partial panel, which reflects less than
    a rigid end);
  * brute-force nearest-cell injection (ties, skipped wall cells) and
    uniform-grid scatter.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from oracle.amr_oracle import OracleAMR
from oracle.oracle import OracleFDTD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def line_grid(nc=8, dx=0.05):
    """The 8-cell row of test_amr_throughput.py:26-64."""
    from afdtd.amr import AcousticAMRGrid
    hx = dx / 2
    nf = nc - 1
    cc = np.zeros((nc, 3))
    cc[:, 0] = np.arange(nc) * dx
    return AcousticAMRGrid(
        base_dx=dx, min_dx=dx, max_refinement_level=0,
        bounds_min=np.array([0.0, -hx, -hx]), bounds_max=np.array([nc * dx, hx, hx]),
        cell_centers=cc, cell_half_sizes=np.full(nc, hx), cell_levels=np.zeros(nc, np.int16),
        cell_types=np.zeros(nc, np.uint8), importance=np.zeros(nc, np.uint8),
        cell_volumes=np.full(nc, dx ** 3), open_volume_fraction=np.ones(nc),
        face_cell_neg=np.arange(nf, dtype=np.int32), face_cell_pos=np.arange(1, nc, dtype=np.int32),
        face_axis=np.zeros(nf, np.uint8), face_area=np.full(nf, dx ** 2),
        face_open_fraction=np.ones(nf), face_distance=np.full(nf, dx))


def random_topology(seed, nc=400, nf=1400, walls=True):
    """Irregular cells and faces with random volumes, open fractions, areas,
    distances; a few wall/plate cells; no self-faces."""
    rng = np.random.default_rng(seed)
    cc = rng.uniform(0.0, 1.0, (nc, 3))
    vol = rng.uniform(2e-6, 4e-6, nc)     # stable at the order-2 dt for this degree
    ov = rng.uniform(0.2, 1.0, nc)
    ct = np.zeros(nc, np.uint8)
    if walls:
        ct[rng.choice(nc, nc // 20, replace=False)] = rng.integers(1, 3, nc // 20)
        ov[rng.choice(nc, 5, replace=False)] = 0.0
    a = rng.integers(0, nc, nf)
    b = (a + rng.integers(1, nc, nf)) % nc
    area = rng.uniform(0.5e-4, 1.5e-4, nf)
    fo = rng.uniform(0.0, 1.0, nf)
    dist = rng.uniform(0.008, 0.012, nf)
    return dict(cell_centers=cc, cell_volumes=vol, open_volume_fraction=ov, cell_types=ct,
                face_cell_neg=a.astype(np.int32), face_cell_pos=b.astype(np.int32),
                face_area=area, face_open_fraction=fo, face_distance=dist,
                c=343.0, rho_air=1.21, min_dx=0.01, gradient_order=2)


def test_line_one_step_known_answer():
    g = line_grid()
    o = OracleAMR.from_grid(g)
    dt = 0.77 * 0.05 / (343.0 * math.sqrt(3.0))
    assert o.dt == pytest.approx(dt, rel=1e-15)
    o.inject_pressure_nearest(g.cell_centers[0], 1.0)
    o.step(1)
    p, v = o.get_pressure().astype(np.float64), o.get_velocity().astype(np.float64)
    # v0 = -(dt/rho) * (0 - 1)/dx ; p0 = 1 - C^2 ; p1 = C^2 with C = c dt/dx = 0.77/sqrt(3)
    C2 = 0.77 ** 2 / 3.0
    assert v[0] == pytest.approx(dt / (1.21 * 0.05), rel=1e-6)
    assert np.all(v[1:] == 0.0)
    assert p[0] == pytest.approx(1.0 - C2, rel=1e-6)
    assert p[1] == pytest.approx(C2, rel=1e-5)
    assert np.all(p[2:] == 0.0)
    assert o.step_count == 1


def test_flux_conservation_irregular():
    t = random_topology(3, walls=False)
    o = OracleAMR(**t)
    rng = np.random.default_rng(4)
    o.set_state(rng.standard_normal(400).astype(np.float32) * 0.1)
    w = t["cell_volumes"] * np.maximum(t["open_volume_fraction"], 1e-12)
    m0 = float(np.sum(o.get_pressure() * w))
    scale = float(np.sum(np.abs(o.get_pressure()) * w))
    o.step(200)
    m1 = float(np.sum(o.get_pressure() * w))
    assert abs(m1 - m0) < 1e-5 * scale
    assert np.abs(o.get_pressure()).max() > 1e-4     # the field did move


def test_uniform_box_reduces_to_yee_oracle():
    from afdtd.amr import AcousticAMRGrid
    Nx, Ny, Nz, dx = 12, 10, 14, 0.01
    g = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=Nx * dx, y_min=0, y_max=Ny * dx,
                                          z_min=0, z_max=Nz * dx, dx=dx)
    a = OracleAMR.from_grid(g)
    y = OracleFDTD(Nx, Ny, Nz, dx, n_pml=0, order=2, courant=0.77 / math.sqrt(3.0),
                   boundary=0, precision="f64")
    x, yy, z = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    p0 = np.exp(-((x - 5.3) ** 2 + (yy - 4.1) ** 2 + (z - 6.7) ** 2) / 4.0).ravel()
    p0[17] += 0.5          # an off-centre spike breaks every symmetry
    a.set_state(p0.astype(np.float32))
    y.set_fields(P=p0.astype(np.float32).astype(np.float64))
    a.step(150)
    y.step(150)
    pa = a.get_pressure().astype(np.float64)
    py = y.get_pressure_field().ravel()
    assert np.linalg.norm(pa - py) <= 1e-5 * np.linalg.norm(py)
    # faces: x faces then y faces then z faces, each in neg-cell order
    va = a.get_velocity().astype(np.float64)
    vy_ = np.concatenate([y.get_velocity_x(), y.get_velocity_y(), y.get_velocity_z()])
    assert np.linalg.norm(va - vy_) <= 1e-5 * np.linalg.norm(vy_)


def _column(n=400, n_pml=20, mode=None, R=0.0):
    """1-D column of cells along x, PML-typed at both x ends, with the border
    box far away in y and z so only the x distance grades the absorber.
    Returns the peak of the pulse that comes back from the +x end."""
    from afdtd.amr import AcousticAMRGrid
    dx = 0.01
    g = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=n * dx, y_min=0, y_max=dx, z_min=0,
                                          z_max=dx, dx=dx)
    xs = np.arange(n)
    ct = np.zeros(n, np.uint8)
    if mode is not None:
        ct[(xs < n_pml) | (xs >= n - n_pml)] = 3
    o = OracleAMR(g.cell_centers, g.cell_volumes, g.open_volume_fraction, ct, g.face_cell_neg,
                  g.face_cell_pos, g.face_area, g.face_open_fraction, g.face_distance,
                  343.0, 1.21, dx)
    if mode is not None:
        o.set_border_condition(mode, 3.0 if mode == 0 else 2.0, R, 0.0, n_pml,
                               [0.0, -1.0, -1.0], [n * dx, 1.0, 1.0])
    # right-going pulse: v = p/(rho c), v taken half a step earlier
    C_ = 343.0 * o.dt / dx
    p0 = np.exp(-((xs - 150) / 8.0) ** 2)
    v0 = np.exp(-((xs[:-1] + 0.5 - 150 + 0.5 * C_) / 8.0) ** 2) / (1.21 * 343.0)
    o.set_state(p0.astype(np.float32), v0.astype(np.float32))
    o.step(int(round((n - 150 + n - 200) / C_)))      # out to x = n and back to x = 200
    return float(np.abs(o.get_pressure()[150:250]).max())


def test_border_condition_orders_reflection():
    """Absorption behaves as graded damping should: a rigid end returns the
    pulse, an absorber returns a small fraction that shrinks with depth, and a
    room panel returns more the larger its R.  (At equal open volumes the
    reference's face sigma, 2*(vp*a_n + vn*a_p)/(vn + vp), is the SUM of the two
    cell values, twice the cells' own damping; with the steep anechoic profile
    that leaves about 12 % at n_pml = 20.)"""
    rigid = _column(mode=None)
    anech20, anech40 = _column(mode=0), _column(mode=0, n_pml=40)
    panel5, panel9 = _column(mode=1, R=0.5), _column(mode=1, R=0.9)
    assert rigid > 0.95
    assert anech20 < 0.15 * rigid
    assert anech40 < 0.7 * anech20
    assert panel5 < panel9 < 0.6 * rigid


def test_room_panel_full_reflection_is_identity():
    t = random_topology(5, walls=False)
    t["cell_types"][::3] = 3
    a, b = OracleAMR(**t), OracleAMR(**t)
    b.set_border_condition(1, 2.0, 1.0, 0.0, 4, [0, 0, 0], [1, 1, 1])
    pd, ps, vd = b.get_coeffs()
    assert np.all(pd == 1.0) and np.all(ps == 1.0) and np.all(vd == 1.0)
    s = np.random.default_rng(6).standard_normal(400).astype(np.float32)
    for o in (a, b):
        o.set_state(s)
        o.step(40)
    assert np.array_equal(a.get_pressure().view(np.uint32), b.get_pressure().view(np.uint32))


def test_border_grades_only_pml_cells():
    t = random_topology(7, walls=False)
    t["cell_types"][:200] = 3
    o = OracleAMR(**t)
    o.set_border_condition(0, 3.0, 0.0, 0.0, 30, [0, 0, 0], [1, 1, 1])
    pd, ps, vd = o.get_coeffs()
    assert np.all(pd[200:] == 1.0) and np.all(ps[200:] == 1.0)
    assert np.all(pd[:200] <= 1.0) and np.all(pd[:200] > 0.0) and np.any(pd[:200] < 1.0)
    # P_src = (1 - P_damp)/s with s = -log(P_damp): the exact integrating
    # factor; float cancellation in 1 - exp(-s) limits it for small s
    assert np.all((ps > 0) & (ps <= 1.0 + 1e-3))
    s = -np.log(pd[:200].astype(np.float64))
    big = s > 1e-2
    assert big.sum() > 20
    assert np.allclose(ps[:200][big], (1 - np.exp(-s[big])) / s[big], rtol=1e-4)
    neg, pos = t["face_cell_neg"], t["face_cell_pos"]
    interior = (neg >= 200) & (pos >= 200)
    assert np.all(vd[interior] == 1.0) and np.all(vd[~interior] <= 1.0)


def _lround(x):
    return np.where(x >= 0, np.floor(x + 0.5), -np.floor(-x + 0.5)).astype(np.int64)


def test_inject_nearest_brute_force_and_ties():
    t = random_topology(8)
    o = OracleAMR(**t)
    cc, ct, ov = t["cell_centers"], t["cell_types"], t["open_volume_fraction"]
    cand = np.where(((ct == 0) | (ct == 3)) & (ov > 0))[0]
    rng = np.random.default_rng(9)
    for q in rng.uniform(0, 1, (20, 3)):
        d2 = ((cc[cand] - q) ** 2).sum(1)
        want = cand[np.argmin(d2)]
        o.reset()
        o.inject_pressure_nearest(q, 2.5)
        p = o.get_pressure()
        assert p[want] == np.float32(2.5) and np.count_nonzero(p) == 1
    # exact tie between two cells: the lower index wins
    g = line_grid()
    o2 = OracleAMR.from_grid(g)
    o2.inject_pressure_nearest([0.025, 0.0, 0.0], 1.0)
    p = o2.get_pressure()
    assert p[0] == 1.0 and p[1] == 0.0
    with pytest.raises(RuntimeError):
        o2.inject_pressure_nearest([0.0, 0.0, 0.0], float("nan"))


def test_scatter_uniform_brute_force():
    t = random_topology(10)
    o = OracleAMR(**t)
    o.set_state(np.arange(400, dtype=np.float32) + 1.0)
    Nx, Ny, Nz = 7, 5, 6
    bmin, bmax = np.array([0.0, 0.1, 0.0]), np.array([1.0, 0.9, 1.1])
    got = o.scatter_pressure_uniform(Nx, Ny, Nz, bmin, bmax)
    want = np.zeros((Nx, Ny, Nz), np.float32)
    r = (np.array([Nx, Ny, Nz]) - 1) / (bmax - bmin)
    for ci, q in enumerate(t["cell_centers"]):
        ix, iy, iz = _lround((q - bmin) * r)
        if 0 <= ix < Nx and 0 <= iy < Ny and 0 <= iz < Nz:
            want[ix, iy, iz] = ci + 1.0
    assert np.array_equal(got, want)


def test_stability_guard_and_validation():
    g = line_grid()
    o = OracleAMR.from_grid(g)
    o.inject_pressure_nearest(g.cell_centers[3], 5e3)
    assert o.step_rc(5) == -3 and o.step_count == 0      # |p| > 2e3 at step 0
    o.reset()
    o.step(3)
    s = np.zeros(8, np.float32)
    s[2] = np.inf
    o.set_state(s)
    assert o.step_rc(5) == -3 and o.step_count == 3      # non-finite: every step
    t = random_topology(11)
    t["face_cell_pos"] = t["face_cell_neg"].copy()       # self-face
    with pytest.raises(RuntimeError):
        OracleAMR(**t)
    with pytest.raises(RuntimeError):
        OracleAMR.from_grid(g, gradient_order=4)         # only orders 2 and 8


# ── C ABI without a GPU ─────────────────────────────────────────────────

def _amr_declared():
    txt = open(os.path.join(ROOT, "include", "afdtd_amr.h")).read()
    return sorted(set(re.findall(r"SK_API\s+[\w\s\*]*?\b(amr_\w+)\s*\(", txt)))


REFERENCE_AMR_ORDER2 = (
    "amr_create", "amr_get_last_create_error_stage", "amr_get_last_create_error_message",
    "amr_get_create_progress_active", "amr_get_create_progress_done_faces",
    "amr_get_create_progress_total_faces", "amr_destroy", "amr_step", "amr_reset",
    "amr_inject_pressure_nearest", "amr_get_pressure", "amr_get_cell_centers",
    "amr_get_velocity", "amr_get_dt", "amr_get_step_count", "amr_scatter_pressure_uniform",
    "amr_set_border_condition")


def test_amr_header_and_exports():
    from afdtd import LIB_PATH
    names = _amr_declared()
    assert all(n in names for n in REFERENCE_AMR_ORDER2)
    L = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_amr_abi_argument_checks_without_device():
    from afdtd.amr import lib, last_create_error
    L = lib()
    assert L.amr_step(None, 1) == -1
    assert L.amr_reset(None) == -1
    assert L.amr_get_pressure(None, None, 0) == -1
    assert L.amr_get_dt(None) == 0.0 and L.amr_get_step_count(None) == 0
    L.amr_destroy(None)
    t = random_topology(12)
    t["face_distance"][5] = 0.0
    from afdtd.amr import _p
    k = [np.ascontiguousarray(t[n], d) for n, d in (
        ("cell_centers", np.float64), ("cell_volumes", np.float64),
        ("open_volume_fraction", np.float64), ("cell_types", np.uint8),
        ("face_cell_neg", np.int32), ("face_cell_pos", np.int32), ("face_area", np.float64),
        ("face_open_fraction", np.float64), ("face_distance", np.float64))]
    h = L.amr_create(400, *[_p(a) for a in k[:4]], k[4].size, *[_p(a) for a in k[4:]],
                     343.0, 1.21, 0.01, 2)
    assert not h and last_create_error()[0] == "validate_inputs"
    k[8][5] = 0.01
    h = L.amr_create(400, *[_p(a) for a in k[:4]], k[4].size, *[_p(a) for a in k[4:]],
                     343.0, 1.21, 0.01, 4)
    assert not h and last_create_error()[1] == "gradient_order must be 2 or 8"


# ── Plate coupling and mic samplers (acoustic_amr.cpp:939-1072, :1412-1792) ──

def plated_box(nx, ny, nz, kp, active=None, dx=0.01):
    """Box grid with a plate layer at z index kp: plate node (i, j) is cell
    (i, j, kp), typed PLATE when active; its above face joins it to (i, j,
    kp+1) and its below face joins (i, j, kp-1) to it.  Returns the grid and
    the keyword arguments of setup_plate."""
    from afdtd.amr import AcousticAMRGrid
    g = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=nx * dx, y_min=0, y_max=ny * dx,
                                          z_min=0, z_max=nz * dx, dx=dx)
    act = np.ones((nx, ny), np.uint8) if active is None else np.asarray(active, np.uint8)
    ct = g.cell_types.copy().reshape(nx, ny, nz)
    ct[:, :, kp][act > 0] = 2
    flat = np.flatnonzero(act.ravel()).astype(np.int32)
    i, j = flat // ny, flat % ny
    F0 = (nx - 1) * ny * nz + nx * (ny - 1) * nz          # first z face
    zf = lambda ii, jj, kk: F0 + (ii * ny + jj) * (nz - 1) + kk
    cell = lambda ii, jj, kk: (ii * ny + jj) * nz + kk
    n = flat.size
    # plate node spacing 5 dx keeps the explicit plate update stable at the air dt
    kw = dict(plate_Nx=nx, plate_Ny=ny, plate_dx=5 * dx, plate_origin=[0, 0, kp * dx],
              plate_active=act.ravel(), plate_rho_h=7.0 * 0.003, plate_D=0.45,
              plate_alpha_M=2.0, plate_beta_K=1e-5, plate_active_flat_idx=flat,
              face_above_starts=np.arange(n + 1, dtype=np.int32), face_above_idx=zf(i, j, kp),
              face_above_wgt=np.ones(n, np.float32),
              face_below_starts=np.arange(n + 1, dtype=np.int32),
              face_below_idx=zf(i, j, kp - 1), face_below_wgt=np.ones(n, np.float32),
              cell_above=cell(i, j, kp + 1), cell_below=cell(i, j, kp - 1))
    grid = AcousticAMRGrid(**{**g.__dict__, "cell_types": ct.ravel()})
    return grid, kw


def _plate_consts(o, kw):
    dt = np.float64(np.float32(o.dt))
    rh, D, aM, bK = (np.float64(np.float32(kw[k])) for k in
                     ("plate_rho_h", "plate_D", "plate_alpha_M", "plate_beta_K"))
    dx4 = 1.0 / np.float64(np.float32(kw["plate_dx"])) ** 4
    return dt, rh, D, aM, bK, dx4


def test_plate_acoustic_load_and_velocity_bc_known_answer():
    """1-cell column, plate at kp: pressure pa above and pb below load the
    plate by (pb - pa) dx^2; the next step drives the plate faces with
    +-(w - w_prev)/dt."""
    nz, kp, pa, pb = 12, 5, 0.3, 1.1
    grid, kw = plated_box(1, 1, nz, kp)
    o = OracleAMR.from_grid(grid)
    o.setup_plate(**kw)
    p0 = np.zeros(nz, np.float32)
    p0[kp + 1:] = pa
    p0[:kp] = pb
    o.set_state(p0)
    o.step(1)
    dt, rh, D, aM, bK, dx4 = _plate_consts(o, kw)
    dxp = np.float64(np.float32(kw["plate_dx"]))
    w1 = dt * dt / rh * ((pb - pa) * dxp * dxp) / (1 + 0.5 * aM * dt)
    assert o.get_plate_displacement(1)[0] == pytest.approx(w1, rel=1e-5)
    assert np.allclose(o.get_pressure()[kp + 1:], pa) and np.allclose(o.get_pressure()[:kp], pb)
    o.step(1)
    v = o.get_velocity()
    assert v[kw["face_above_idx"][0]] == pytest.approx(w1 / dt, rel=1e-5)
    assert v[kw["face_below_idx"][0]] == pytest.approx(-w1 / dt, rel=1e-5)


def test_plate_isolated_recurrence_and_stencil_weights():
    """Plate decoupled from the air (empty face lists, the same fallback cell
    on both sides): a point force gives the node recurrence with the
    biharmonic centre weight 20, and after two steps the neighbours carry
    the 13-point weights -8 : 2 : 1."""
    grid, kw = plated_box(5, 5, 6, 2)
    n = kw["plate_active_flat_idx"].size
    kw.update(face_above_starts=np.zeros(n + 1, np.int32), face_above_idx=np.zeros(0, np.int32),
              face_above_wgt=np.zeros(0, np.float32), face_below_starts=np.zeros(n + 1, np.int32),
              face_below_idx=np.zeros(0, np.int32), face_below_wgt=np.zeros(0, np.float32),
              cell_above=np.zeros(n, np.int32), cell_below=np.zeros(n, np.int32))
    o = OracleAMR.from_grid(grid)
    o.setup_plate(**kw)
    o.inject_plate_force([12], [1.0], 2.0)                # centre node (2, 2)
    o.step(1)
    w1 = o.get_plate_displacement(25).reshape(5, 5)
    dt, rh, D, aM, bK, dx4 = _plate_consts(o, kw)
    df, db = 1 + 0.5 * aM * dt, 1 - 0.5 * aM * dt
    assert w1[2, 2] == pytest.approx(dt * dt / rh * 2.0 / df, rel=1e-6)
    assert np.count_nonzero(w1) == 1
    o.step(1)
    w2 = o.get_plate_displacement(25).reshape(5, 5).astype(np.float64)
    c0 = D * (1 + bK / dt)
    assert w2[2, 3] / w2[3, 3] == pytest.approx(-8 / 2, rel=1e-5)
    assert w2[2, 4] / w2[3, 3] == pytest.approx(1 / 2, rel=1e-5)
    assert w2[2, 3] == pytest.approx(dt * dt / rh * (c0 * 8 * dx4 * w1[2, 2]) / df, rel=1e-4)
    # the centre node alone: free recurrence in float64 over 150 steps
    g1, kw1 = plated_box(1, 1, 6, 2)
    kw1.update(face_above_starts=[0, 0], face_above_idx=np.zeros(0, np.int32),
               face_above_wgt=np.zeros(0, np.float32), face_below_starts=[0, 0],
               face_below_idx=np.zeros(0, np.int32), face_below_wgt=np.zeros(0, np.float32),
               cell_above=[0], cell_below=[0])
    o1 = OracleAMR.from_grid(g1)
    o1.setup_plate(**kw1)
    o1.inject_plate_force([0], [1.0], 2.0)
    ws = []
    for _ in range(150):
        o1.step(1)
        ws.append(o1.get_plate_displacement(1)[0])
    dt, rh, D, aM, bK, dx4 = _plate_consts(o1, kw1)
    c0, cp = D * (1 + bK / dt), D * (bK / dt)
    w, wp, ref = 0.0, 0.0, []
    f = 2.0
    for _ in range(150):
        rhs = f - c0 * 20 * dx4 * w + cp * 20 * dx4 * wp
        w, wp = (db * (2 * w - wp) + dt * dt / rh * rhs) / df, w
        ref.append(w)
        f = 0.0
    ref = np.array(ref)
    assert np.linalg.norm(np.array(ws) - ref) <= 1e-4 * np.linalg.norm(ref)


def test_plate_injection_order_and_samplers():
    grid, kw = plated_box(4, 3, 6, 2)
    o = OracleAMR.from_grid(grid)
    o.setup_plate(**kw)
    o.set_bridge_plate_sources([1, 5, 5], [0.5, 0.25, 0.25])
    o.inject_bridge_drive([2.0, 4.0, -1.0], 3.0)
    with pytest.raises(RuntimeError, match="rc=-2"):     # stops at the bad node
        o.inject_plate_force([7, 99, 8], [1.0, 1.0, 1.0], 0.5)
    o.step(1)
    w = o.get_plate_displacement(12).astype(np.float64)
    dt, rh, D, aM, bK, dx4 = _plate_consts(o, kw)
    s = dt * dt / rh / (1 + 0.5 * aM * dt)
    # loads: node 1: 3.0, node 5: 3.0 - 0.75 = 2.25, node 7: 0.5, node 8: none
    assert w[1] / s == pytest.approx(3.0, rel=1e-4)
    assert w[5] / s == pytest.approx(2.25, rel=1e-4)
    assert w[7] / s == pytest.approx(0.5, rel=1e-4)
    assert w[8] == 0.0
    idx4 = np.array([[1, 5, 7, 0], [2, 3, 4, 6]], np.int32)
    wgt4 = np.array([[0.1, 0.2, 0.3, 0.4], [0.25, 0.25, 0.25, 0.25]], np.float32)
    got = o.sample_plate_displacement_batch(idx4, wgt4)
    wf = o.get_plate_displacement(12)
    assert np.allclose(got, (wgt4 * wf[idx4]).sum(1), rtol=1e-6, atol=0)
    assert o.sample_plate_weighted([1, 5], [0.5, 0.5]) == pytest.approx(
        0.5 * wf[1] + 0.5 * wf[5], rel=1e-6)


def mic_grid():
    """10x9x8 box with a few wall cells and half-open faces."""
    from afdtd.amr import AcousticAMRGrid
    g = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=0.1, y_min=0, y_max=0.09, z_min=0,
                                          z_max=0.08, dx=0.01)
    rng = np.random.default_rng(15)
    ct = g.cell_types.copy()
    ct[rng.choice(g.n_cells, 20, replace=False)] = 1
    fo = g.face_open_fraction.copy()
    fo[rng.choice(g.n_faces, 300, replace=False)] = 0.4
    return dict(cell_centers=g.cell_centers, cell_volumes=g.cell_volumes,
                open_volume_fraction=g.open_volume_fraction, cell_types=ct,
                face_cell_neg=g.face_cell_neg, face_cell_pos=g.face_cell_pos,
                face_area=g.face_area, face_open_fraction=fo, face_distance=g.face_distance,
                c=343.0, rho_air=1.21, min_dx=0.01, gradient_order=2)


# off the lattice, so no cell or face centre sits at the search radius
# (ties there would decide by float rounding); the third lies outside the
# grid and is clamped onto its envelope
MIC_POS = np.array([[0.0513, 0.0447, 0.0391], [0.0132, 0.071, 0.0277], [0.2, -0.05, 0.0333]],
                   np.float32)


def test_mic_samplers_brute_force():
    t = mic_grid()
    o = OracleAMR(**t)
    rng = np.random.default_rng(14)
    p = rng.standard_normal(o.n_cells).astype(np.float32)
    v = rng.standard_normal(o.n_faces).astype(np.float32)
    o.set_state(p, v)
    pos = MIC_POS
    o.setup_mic_samplers(pos, np.tile([0, 0, 1], (3, 1)), True)
    op, vx, vy, vz = o.sample_mics()
    cc, ct, ov = t["cell_centers"], t["cell_types"], t["open_volume_fraction"]
    ac = np.flatnonzero(((ct == 0) | (ct == 3)) & (ov > 0))
    R = 2.0 * np.cbrt(t["cell_volumes"][ac].mean())
    lo, hi = cc[ac].min(0), cc[ac].max(0)
    for m, q in enumerate(np.clip(pos.astype(np.float64), lo, hi)):
        r = np.linalg.norm(cc[ac] - q, axis=1)
        sel = r < R
        w = 1.0 / r[sel]
        assert op[m] == pytest.approx(np.sum(w * p[ac[sel]]) / w.sum(), rel=1e-4, abs=1e-6)
        a, b = t["face_cell_neg"], t["face_cell_pos"]
        fc = 0.5 * (cc[a] + cc[b])
        nrm = cc[b] - cc[a]
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        aw = t["face_area"] * t["face_open_fraction"]
        fs = (np.linalg.norm(fc - q, axis=1) < R) & (t["face_open_fraction"] >= 0.5)
        if fs.any():
            want = (nrm[fs] * (aw[fs] * v[fs])[:, None]).sum(0) / aw[fs].sum()
            assert np.allclose([vx[m], vy[m], vz[m]], want, rtol=1e-4, atol=1e-6)


# ── Gradient order 8: the Fornberg stencil (acoustic_amr.cpp:313-679) ─────
#
# The reference's fornberg_d1 (acoustic_amr.cpp:398-426) updates the new
# node's weights with (x[n] - xi) where Fornberg's recurrence uses
# (x[n-1] - xi).  For two nodes the result is still the exact difference;
# for three or more the weights are not first-derivative weights (on the
# symmetric 8-point stencil they annihilate x as well as 1).  The oracle and
# the GPU path reproduce the reference's weights as written; their values
# beyond two nodes are "parity unpinned": only the properties below pin them.


def test_fornberg_reference_recurrence_properties():
    from oracle.amr_oracle import fornberg_d1
    rng = np.random.default_rng(16)
    for _ in range(5):                      # two nodes: the exact difference
        x = np.sort(rng.uniform(-4.0, 4.0, 2)).astype(np.float32)
        w = fornberg_d1(x)
        assert w[1] == pytest.approx(1.0 / (float(x[1]) - float(x[0])), rel=1e-6)
        assert w[0] == -w[1]
    for n in (3, 5, 8):                     # constants are annihilated
        x = np.sort(rng.uniform(-4.0, 4.0, n)).astype(np.float32)
        w = fornberg_d1(x).astype(np.float64)
        assert abs(w.sum()) < 1e-5 * np.abs(w).sum()
    xs = np.float32([-3.5, -2.5, -1.5, -0.5, 0.5, 1.5, 2.5, 3.5])
    w = fornberg_d1(xs).astype(np.float64)
    assert np.allclose(w, -w[::-1], rtol=1e-5, atol=1e-9)   # odd on symmetric nodes
    # the defect itself: not the 8th-order weights, and blind to a linear field
    assert not np.allclose(w[4:], C8, rtol=0.1)
    assert abs(np.sum(w * xs)) < 1e-6


# 8th-order staggered first-derivative coefficients (closed form)
C8 = np.array([1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168])


def _box(nx, ny, nz, dx=0.01, rot=0.0):
    from afdtd.amr import AcousticAMRGrid
    g = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=nx * dx, y_min=0, y_max=ny * dx,
                                          z_min=0, z_max=nz * dx, dx=dx)
    if rot:
        c, s_ = np.cos(rot), np.sin(rot)
        cc = g.cell_centers.copy()
        cc[:, 0], cc[:, 1] = c * g.cell_centers[:, 0] - s_ * g.cell_centers[:, 1], \
            s_ * g.cell_centers[:, 0] + c * g.cell_centers[:, 1]
        g = type(g)(**{**g.__dict__, "cell_centers": cc})
    return g


@pytest.mark.parametrize("rot", [0.0, np.pi / 4])
def test_order8_stencil_closed_form_on_box(rot):
    """Interior faces of a box grid (also rotated by 45 degrees, which takes
    the CSR-search walk instead of the axis tables) take the 4 + 4 cells
    along the face normal, nearest first on each side, weighted by the
    reference recurrence over their offsets; faces with fewer than 4 cells
    on a side fall back to -1/d, +1/d."""
    from oracle.amr_oracle import fornberg_d1
    nx, ny, nz, dx = 11, 10, 9, 0.01
    g = _box(nx, ny, nz, dx, rot)
    o = OracleAMR.from_grid(g, gradient_order=8)
    assert o.dt == pytest.approx(0.10 * dx / (343.0 * np.sqrt(3.0)), rel=1e-14)
    sc, sw = o.get_stencil()
    cid = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    negs = [cid[:-1].ravel(), cid[:, :-1].ravel(), cid[:, :, :-1].ravel()]
    strides = (ny * nz, nz, 1)
    dims = (nx, ny, nz)
    f = 0
    n_full = n_fb = 0
    for ax in range(3):
        for c0 in negs[ax]:
            ijk = np.unravel_index(c0, dims)
            lo, hi = ijk[ax], dims[ax] - 2 - ijk[ax]      # cells beyond each side
            if lo >= 3 and hi >= 3:
                cells = c0 + strides[ax] * np.array([0, -1, -2, -3, 1, 2, 3, 4])
                assert np.array_equal(sc[f], cells)
                # weights of the reference recurrence over the node offsets
                # along the normal, farthest neg node first
                want = fornberg_d1(np.float32([-3.5, -2.5, -1.5, -0.5, 0.5, 1.5, 2.5, 3.5])
                                   * np.float32(dx))
                want = np.concatenate([want[3::-1], want[4:]])
                assert np.allclose(sw[f], want, rtol=2e-3, atol=1e-3 * np.abs(want).max())
                n_full += 1
            else:
                assert sc[f, 0] == c0 and sc[f, 4] == c0 + strides[ax]
                assert sw[f, 0] == -np.float32(1 / np.float32(dx)) or \
                    sw[f, 0] == -(np.float32(1.0) / np.float32(dx))
                assert sw[f, 4] == -sw[f, 0] and not np.any(sw[f, [1, 2, 3, 5, 6, 7]])
                n_fb += 1
            f += 1
    assert n_full > 100 and n_fb > 100


def test_order8_stops_at_walls_and_conserves_flux():
    """A wall cell ends a stencil walk (two-point fallback next to it), and
    the order-8 step keeps sum(p*V) (the divergence is the same telescoping
    sum)."""
    g = _box(14, 6, 6)
    ct = g.cell_types.copy()
    wall = (7 * 6 + 3) * 6 + 3                             # cell (7, 3, 3)
    ct[wall] = 1
    g2 = type(g)(**{**g.__dict__, "cell_types": ct})
    o = OracleAMR.from_grid(g2, gradient_order=8)
    sc, sw = o.get_stencil()
    xf = np.flatnonzero((g.face_cell_neg == (5 * 6 + 3) * 6 + 3) & (g.face_axis == 0))[0]
    assert not np.any(sw[xf, [1, 2, 3, 5, 6, 7]])         # face (5|6): walk hits the wall at 7
    rng = np.random.default_rng(17)
    o2 = OracleAMR.from_grid(g, gradient_order=8)
    o2.set_state(rng.standard_normal(g.n_cells).astype(np.float32))
    m0 = float(np.sum(o2.get_pressure().astype(np.float64)))
    sc0 = float(np.sum(np.abs(o2.get_pressure()).astype(np.float64)))
    o2.step(300)
    assert abs(float(np.sum(o2.get_pressure().astype(np.float64))) - m0) < 1e-5 * sc0


@pytest.mark.parametrize("mode, order, R", [(0, 3.0, 0.0), (1, 2.0, 0.6)])
def test_border_sigma_closed_form_three_cells(mode, order, R):
    """amr_set_border_condition (acoustic_amr.cpp:1318-1409) on three cells
    in a row along x, two PML-typed at distances 0.02 and 0.06 from the
    x- bound (all other bounds far), one air cell, n_pml = 10 at min_dx =
    0.01 (depth 0.1).  By hand: sigma_i = s_max (1 - d_i/0.1)^order with
    s_max = 3.45/(dt n_pml) (times 1 - R for a room panel); P_damp =
    exp(-sigma dt), P_src = (1 - exp(-sigma dt))/(sigma dt); each face's
    sigma is 2 (o_p sigma_n + o_n sigma_p)/(o_n + o_p) over the cells'
    open fractions o (so at equal fractions the SUM of the two cell
    values), V_damp = exp(-sigma_f dt)."""
    cc = np.array([[0.02, 0.5, 0.5], [0.06, 0.5, 0.5], [0.10, 0.5, 0.5]])
    ovf = np.array([0.3, 0.9, 0.6])
    ct = np.array([3, 3, 0], np.uint8)
    o = OracleAMR(cc, np.full(3, 1e-6), ovf, ct, np.array([0, 1], np.int32),
                  np.array([1, 2], np.int32), np.full(2, 1e-4), np.ones(2), np.full(2, 0.04),
                  343.0, 1.21, 0.01)
    assert o.set_border_condition(mode, order, R, 0.0, 10, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]) == 0
    pd, ps, vd = (x.astype(np.float64) for x in o.get_coeffs())
    dt = float(np.float32(o.dt))
    smax = 3.45 / (dt * 10) * (1.0 - R)
    sig = np.array([smax * 0.8 ** order, smax * 0.4 ** order, 0.0])
    np.testing.assert_allclose(pd, np.exp(-sig * dt), rtol=2e-6)
    s = sig[:2] * dt
    np.testing.assert_allclose(ps[:2], (1 - np.exp(-s)) / s, rtol=1e-4)
    assert ps[2] == 1.0 and pd[2] == 1.0
    sf = np.array([2 * (ovf[1] * sig[0] + ovf[0] * sig[1]) / (ovf[0] + ovf[1]),
                   2 * (ovf[2] * sig[1] + ovf[1] * sig[2]) / (ovf[1] + ovf[2])])
    np.testing.assert_allclose(vd, np.exp(-sf * dt), rtol=2e-6)
