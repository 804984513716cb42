"""GPU parity of the AMR (CSR, gradient order 2) stepper: the HIP kernels
behind the amr_* C ABI equal the CPU oracle bit for bit, on the reference's
8-cell line, on irregular topologies with wall cells and partial openings,
on uniform boxes with each border mode, through the divergence guard, and
at 3.9 M and 16.8 M cells."""
import numpy as np
import pytest

from oracle.amr_oracle import OracleAMR
from test_amr_oracle_cpu import line_grid, random_topology

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same(g, o):
    assert g.step_count == o.step_count
    assert np.array_equal(_bits(g.get_pressure()), _bits(o.get_pressure()))
    assert np.array_equal(_bits(g.get_velocity()), _bits(o.get_velocity()))


def _pair_grid(grid, ct=None):
    from afdtd.amr import AcousticAMR
    args = (grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction,
            grid.cell_types if ct is None else ct, grid.face_cell_neg, grid.face_cell_pos,
            grid.face_area, grid.face_open_fraction, grid.face_distance, 343.0, 1.21,
            grid.min_dx, 2)
    return AcousticAMR(*args), OracleAMR(*args)


def test_line_pulse_bitwise():
    print("amr: 8-cell line", flush=True)
    g, o = _pair_grid(line_grid())
    assert g.dt == o.dt
    for m in (g, o):
        m.inject_pressure_nearest(line_grid().cell_centers[0], 1.0)
        m.step(1)
    _same(g, o)
    g.step(299)
    o.step(299)
    _same(g, o)


@pytest.mark.parametrize("seed,walls", [(21, True), (22, False)])
def test_irregular_topology_bitwise(seed, walls):
    from afdtd.amr import AcousticAMR
    print(f"amr: irregular seed={seed} walls={walls}", flush=True)
    t = random_topology(seed, walls=walls)
    g, o = AcousticAMR(**t), OracleAMR(**t)
    assert g.n_csr_entries == o.n_csr_entries == 2 * g.n_faces
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(g.n_cells).astype(np.float32)
    v = (rng.standard_normal(g.n_faces) * 1e-3).astype(np.float32)
    for m in (g, o):
        m.set_state(p, v)
        m.step(60)
    _same(g, o)


@pytest.mark.parametrize("mode,R", [(0, 0.0), (1, 0.6), (2, 0.0)])
def test_uniform_box_border_bitwise(mode, R):
    from afdtd.amr import AcousticAMRGrid
    print(f"amr: box 20x16x24 border mode={mode}", flush=True)
    dx = 0.01
    grid = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=20 * dx, y_min=0, y_max=16 * dx,
                                             z_min=0, z_max=24 * dx, dx=dx).with_pml_shell(4)
    g, o = _pair_grid(grid)
    for m in (g, o):
        m.set_border_condition(mode, 3.0, R, 0.0, 4, grid.bounds_min, grid.bounds_max)
        m.inject_pressure_nearest([0.083, 0.071, 0.118], 1.0)
        m.inject_pressure_nearest([0.121, 0.052, 0.061], -0.5)
        m.step(250)
    _same(g, o)
    assert np.abs(g.get_pressure()).max() > 0
    assert np.array_equal(g.get_cell_centers(), grid.cell_centers.astype(np.float32))


def test_injection_ties_and_scatter():
    from afdtd.amr import AcousticAMR
    print("amr: nearest-cell injection and scatter", flush=True)
    t = random_topology(23)
    g, o = AcousticAMR(**t), OracleAMR(**t)
    rng = np.random.default_rng(24)
    for q in rng.uniform(-0.1, 1.1, (12, 3)):
        val = float(rng.standard_normal())
        g.inject_pressure_nearest(q, val)
        o.inject_pressure_nearest(q, val)
    assert np.array_equal(_bits(g.get_pressure()), _bits(o.get_pressure()))
    bmin, bmax = [0.0, 0.1, -0.05], [1.0, 0.95, 1.05]
    assert np.array_equal(g.scatter_pressure_uniform(9, 7, 11, bmin, bmax),
                          o.scatter_pressure_uniform(9, 7, 11, bmin, bmax))
    gl, ol = _pair_grid(line_grid())
    for m in (gl, ol):
        m.inject_pressure_nearest([0.025, 0.0, 0.0], 1.0)    # exact tie: lower index
    assert np.array_equal(_bits(gl.get_pressure()), _bits(ol.get_pressure()))
    assert gl.get_pressure()[0] == 1.0
    with pytest.raises(RuntimeError, match="rc=-3"):
        gl.inject_pressure_nearest([0.0, 0.0, 0.0], float("inf"))


def test_divergence_guard_matches_oracle():
    print("amr: divergence guard", flush=True)
    grid = line_grid()
    g, o = _pair_grid(grid)
    for m in (g, o):
        m.inject_pressure_nearest(grid.cell_centers[3], 5e3)
        assert m.step_rc(5) == -3
    _same(g, o)                            # failed at step 0: one update, step_count 0
    with pytest.raises(RuntimeError, match="amr_step failed: rc=-3"):
        g.step(1)
    s = np.zeros(8, np.float32)
    s[2] = np.inf
    for m in (g, o):
        m.reset()
        m.step(3)
        m.set_state(s)
        assert m.step_rc(5) == -3
        assert m.step_count == 3
    pg, po = g.get_pressure(), o.get_pressure()
    assert np.array_equal(np.isnan(pg), np.isnan(po))
    fin = np.isfinite(po)
    assert np.array_equal(_bits(pg[fin]), _bits(po[fin]))


def test_reset_and_reuse():
    print("amr: reset", flush=True)
    g, o = _pair_grid(line_grid(16))
    for m in (g, o):
        m.inject_pressure_nearest([0.2, 0, 0], 1.0)
        m.step(40)
        m.reset()
        assert m.step_count == 0 and not np.any(m.get_pressure())
        m.inject_pressure_nearest([0.3, 0, 0], 0.7)
        m.step(33)
    _same(g, o)


def test_full_size_boxes():
    """3.9 M cells with an anechoic shell: bitwise over 20 steps.  16.8 M
    cells (the benchmark box): bitwise over 6 steps from a random state, then
    flux conservation (sum of p*V) over 100 more steps, rigid walls."""
    from afdtd.amr import AcousticAMRGrid
    dx = 0.01
    print("amr: 160x128x192 anechoic", flush=True)
    grid = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=160 * dx, y_min=0, y_max=128 * dx,
                                             z_min=0, z_max=192 * dx,
                                             dx=dx).with_pml_shell(8)
    g, o = _pair_grid(grid)
    for m in (g, o):
        m.set_border_condition(0, 3.0, 0.0, 0.0, 8, grid.bounds_min, grid.bounds_max)
        for q, a in (([0.4, 0.5, 0.9], 1.0), ([1.1, 0.7, 1.3], -2.0), ([0.8, 0.64, 0.96], 0.5)):
            m.inject_pressure_nearest(q, a)
        m.step(20)
    _same(g, o)
    del g, o, grid
    print("amr: 256^3 rigid", flush=True)
    grid = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=256 * dx, y_min=0, y_max=256 * dx,
                                             z_min=0, z_max=256 * dx, dx=dx)
    g, o = _pair_grid(grid)
    rng = np.random.default_rng(31)
    p = rng.standard_normal(grid.n_cells).astype(np.float32)
    for m in (g, o):
        m.set_state(p)
        m.step(6)
    _same(g, o)
    del o
    w = float(dx ** 3)
    m0 = float(np.sum(g.get_pressure().astype(np.float64))) * w
    scale = float(np.sum(np.abs(p).astype(np.float64))) * w
    g.step(100)
    m1 = float(np.sum(g.get_pressure().astype(np.float64))) * w
    assert abs(m1 - m0) <= 1e-6 * scale


def test_plate_coupled_run_bitwise():
    """Plate layer with a ragged active mask, bridge drive and a distributed
    force injected between steps, single steps and a 50-step call."""
    from afdtd.amr import AcousticAMR
    from test_amr_oracle_cpu import plated_box
    print("amr: plate-coupled box 12x10x16", flush=True)
    rng = np.random.default_rng(41)
    act = (rng.uniform(size=(12, 10)) < 0.8).astype(np.uint8)
    grid, kw = plated_box(12, 10, 16, 6, active=act)
    g, o = _pair_grid(grid)
    flat = kw["plate_active_flat_idx"]
    for m in (g, o):
        m.setup_plate(**kw)
        m.set_bridge_plate_sources(flat[[3, 4, 9]], [0.5, 0.3, 0.2])
        m.inject_pressure_nearest([0.061, 0.052, 0.113], 1.0)
    drives = rng.standard_normal((80, 3)).astype(np.float32)
    for s in range(80):
        for m in (g, o):
            m.inject_bridge_drive(drives[s], 0.01)
            if s % 7 == 0:
                m.inject_plate_force(flat[[1, 2, 1]], [0.5, 0.25, 0.25], 0.02)
            m.step(1)
    _same(g, o)
    n = 12 * 10
    assert np.array_equal(_bits(g.get_plate_displacement(n)), _bits(o.get_plate_displacement(n)))
    assert np.abs(g.get_plate_displacement(n)).max() > 0
    for m in (g, o):
        m.step(50)
    _same(g, o)
    assert np.array_equal(_bits(g.get_plate_displacement(n)), _bits(o.get_plate_displacement(n)))
    idx4 = flat[rng.integers(0, flat.size, (9, 4))]
    wgt4 = rng.uniform(size=(9, 4)).astype(np.float32)
    assert np.array_equal(_bits(g.sample_plate_displacement_batch(idx4, wgt4)),
                          _bits(o.sample_plate_displacement_batch(idx4, wgt4)))
    assert g.sample_plate_weighted(flat[:5], [0.2] * 5) == o.sample_plate_weighted(flat[:5],
                                                                                  [0.2] * 5)
    for m in (g, o):
        m.reset()
        assert not np.any(m.get_plate_displacement(n))


def test_mic_samplers_bitwise():
    from afdtd.amr import AcousticAMR
    from test_amr_oracle_cpu import MIC_POS, mic_grid
    print("amr: mic samplers", flush=True)
    t = mic_grid()
    g, o = AcousticAMR(**t), OracleAMR(**t)
    rng = np.random.default_rng(42)
    p = rng.standard_normal(g.n_cells).astype(np.float32)
    v = rng.standard_normal(g.n_faces).astype(np.float32)
    axis = np.tile([0, 0, 1], (3, 1))
    for m in (g, o):
        m.set_state(p, v)
        m.setup_mic_samplers(MIC_POS, axis, True)
    for a, b in zip(g.sample_mics(), o.sample_mics()):
        assert np.array_equal(_bits(a), _bits(b))
    for m in (g, o):
        m.step(20)
        m.setup_mic_samplers(MIC_POS[:2], axis[:2], False)   # a second set-up appends
    for a, b in zip(g.sample_mics(), o.sample_mics()):
        assert np.array_equal(_bits(a), _bits(b))


@pytest.mark.parametrize("rot", [0.0, np.pi / 4])
def test_order8_bitwise(rot):
    """Gradient order 8: the per-face 8-cell stencil built on the host (axis
    tables on the box, the CSR-search walk on the rotated box) and the
    order-8 face kernel equal the oracle bit for bit."""
    from test_amr_oracle_cpu import _box
    print(f"amr: order 8 rot={rot:.3f}", flush=True)
    grid = _box(20, 16, 24, rot=rot)
    if not rot:
        grid = grid.with_pml_shell(4)
    args = (grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction, grid.cell_types,
            grid.face_cell_neg, grid.face_cell_pos, grid.face_area, grid.face_open_fraction,
            grid.face_distance, 343.0, 1.21, grid.min_dx, 8)
    from afdtd.amr import AcousticAMR, lib
    g, o = AcousticAMR(*args), OracleAMR(*args)
    assert g.dt == o.dt
    assert lib().amr_get_create_progress_total_faces() == grid.n_faces
    rng = np.random.default_rng(43)
    p = rng.standard_normal(grid.n_cells).astype(np.float32)
    for m in (g, o):
        if not rot:
            m.set_border_condition(0, 3.0, 0.0, 0.0, 4, grid.bounds_min, grid.bounds_max)
        m.set_state(p)
        m.step(120)
    _same(g, o)


@pytest.mark.parametrize("layout", ["direct", "compact"])
def test_record_layouts_bitwise(layout):
    """Both record layouts of the step kernels (full per-face/per-cell
    records, or class ids into tables of distinct records) give the oracle's
    results bit for bit: a bordered box (order 2), an irregular topology
    with walls, and an order-8 box."""
    from afdtd.amr import AcousticAMR, AcousticAMRGrid
    from test_amr_oracle_cpu import _box
    print(f"amr: layout {layout}", flush=True)
    dx = 0.01
    box = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=18 * dx, y_min=0, y_max=14 * dx,
                                            z_min=0, z_max=22 * dx, dx=dx).with_pml_shell(3)
    g, o = _pair_grid(box)
    g.record_layout = layout
    assert g.record_layout == layout
    for m in (g, o):
        m.set_border_condition(1, 2.0, 0.7, 0.0, 3, box.bounds_min, box.bounds_max)
        m.inject_pressure_nearest([0.071, 0.063, 0.104], 1.0)
        m.step(150)
    assert g.record_layout == layout              # kept across the coefficient rebuild
    _same(g, o)
    t = random_topology(44)
    g, o = AcousticAMR(**t), OracleAMR(**t)
    g.record_layout = layout
    p = np.random.default_rng(45).standard_normal(g.n_cells).astype(np.float32)
    for m in (g, o):
        m.set_state(p)
        m.step(40)
    _same(g, o)
    b8 = _box(16, 12, 18)
    args = (b8.cell_centers, b8.cell_volumes, b8.open_volume_fraction, b8.cell_types,
            b8.face_cell_neg, b8.face_cell_pos, b8.face_area, b8.face_open_fraction,
            b8.face_distance, 343.0, 1.21, b8.min_dx, 8)
    g, o = AcousticAMR(*args), OracleAMR(*args)
    g.record_layout = layout
    p = np.random.default_rng(46).standard_normal(g.n_cells).astype(np.float32)
    for m in (g, o):
        m.set_state(p)
        m.step(60)
    _same(g, o)


def test_record_mic_samples_bitwise():
    """The acoustic half of the co-evolver's per-sample loop on the device:
    sub-steps, then each mic's polar mix (p*a + b*rho_c*(v . axis)) * gain
    per sample, equals the oracle's step / sample_mics / mix loop bit for
    bit; a divergence stops the recording at the failing sample."""
    from afdtd.amr import AcousticAMR
    from test_amr_oracle_cpu import MIC_POS, line_grid, mic_grid
    print("amr: recorded mic samples", flush=True)
    t = mic_grid()
    g, o = AcousticAMR(**t), OracleAMR(**t)
    rng = np.random.default_rng(47)
    p = rng.standard_normal(g.n_cells).astype(np.float32)
    axis = np.array([[0, 0, 1], [0.6, 0.8, 0], [1, 0, 0]], np.float32)
    pa = np.float32([1.0, 0.5, 0.0])
    pb = np.float32([0.0, 0.5, 1.0])
    gain = np.float32([1.0, 2.0, 0.25])
    rho_c = np.float32(np.float32(1.21) * np.float32(343.0))
    for m in (g, o):
        m.set_state(p)
        m.setup_mic_samplers(MIC_POS, axis, True)
    out, rc, done = g.record_mic_samples(25, 4, pa, pb, axis, gain, rho_c)
    assert rc == 0 and done == 25 and g.step_count == 100
    want = np.zeros((25, 3), np.float32)
    for j in range(25):
        o.step(4)
        sp, vx, vy, vz = o.sample_mics()
        sig = sp * pa
        vn = vx * axis[:, 0] + vy * axis[:, 1] + vz * axis[:, 2]
        sig = np.where(pb != 0, sig + (pb * rho_c) * vn, sig).astype(np.float32)
        want[j] = sig * gain
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    _same(g, o)
    gl, _ = _pair_grid(line_grid())
    gl.setup_mic_samplers([[0.1, 0, 0]], [[1, 0, 0]], True)
    gl.step(250)                                   # next stability test at step 256
    gl.inject_pressure_nearest([0.15, 0, 0], 5e4)
    out, rc, done = gl.record_mic_samples(5, 4, [1.0], [0.0], [[1, 0, 0]], [1.0], rho_c)
    assert rc == -3 and done == 1 and gl.step_count == 256   # failed at step 256
    assert out[0, 0] != 0 and not np.any(out[1:])


def test_record_mic_samples_graph_replays():
    """Blocks of the same shape replay the captured launch sequence: three
    blocks in a row with eager work between them (an injection, a border
    change and a second mic set-up that reallocates the mic buffers, both
    forcing a new capture) each equal the oracle bit for bit."""
    from afdtd.amr import AcousticAMR, AcousticAMRGrid
    print("amr: recorded mic samples, graph replays", flush=True)
    dx = 0.01
    box = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=14 * dx, y_min=0, y_max=12 * dx,
                                            z_min=0, z_max=16 * dx, dx=dx).with_pml_shell(2)
    g, o = _pair_grid(box)
    mics = np.array([[0.031, 0.052, 0.077], [0.093, 0.061, 0.112]], np.float32)
    axis = np.array([[0, 0, 1], [1, 0, 0]], np.float32)
    pa, pb, gain = np.float32([1.0, 0.3]), np.float32([0.0, 0.7]), np.float32([1.0, 3.0])
    rho_c = np.float32(np.float32(1.21) * np.float32(343.0))
    for m in (g, o):
        m.setup_mic_samplers(mics, axis, True)
        m.inject_pressure_nearest([0.07, 0.06, 0.08], 1.0)

    def oracle_block(n, k):
        rows = []
        for _ in range(n):
            o.step(k)
            sp, vx, vy, vz = o.sample_mics()
            sig = sp * pa
            vn = vx * axis[:, 0] + vy * axis[:, 1] + vz * axis[:, 2]
            sig = np.where(pb != 0, sig + (pb * rho_c) * vn, sig).astype(np.float32)
            rows.append(sig * gain)
        return np.stack(rows)

    for blk in range(6):
        if blk == 1:
            for m in (g, o):
                m.inject_pressure_nearest([0.05, 0.04, 0.1], -0.5)
        if blk == 2:
            for m in (g, o):
                m.set_border_condition(0, 3.0, 0.0, 0.0, 2, box.bounds_min, box.bounds_max)
        if blk == 4:
            # one mic now; its stencil keeps the old entries and appends the new ones
            extra = np.array([[0.062, 0.043, 0.094]], np.float32)
            for m in (g, o):
                m.setup_mic_samplers(extra, [[0, 1, 0]], True)
            axis = np.float32([[0, 1, 0]])
            pa, pb, gain = pa[:1], np.float32([0.5]), gain[:1]
        out, rc, done = g.record_mic_samples(7, 3, pa, pb, axis, gain, rho_c)
        assert rc == 0 and done == 7
        want = oracle_block(7, 3)
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), blk
        _same(g, o)


def test_record_mic_samples_with_plate_graph_slots():
    """The co-evolver's acoustic half with a plate attached: blocks of 5
    samples x 4 sub-steps (20 steps, not a multiple of the plate's three
    displacement buffers) with a bridge drive and a plate force injected
    between blocks.  Consecutive blocks start at plate rotations 0, 2, 1, 0,
    ... and replay the graph captured for their rotation; every block's mic
    output, the fields and the plate displacement equal the oracle's step /
    sample_mics / mix loop bit for bit."""
    from test_amr_oracle_cpu import plated_box
    print("amr: recorded mic samples with a plate, graph slots", flush=True)
    rng = np.random.default_rng(48)
    act = (rng.uniform(size=(12, 10)) < 0.85).astype(np.uint8)
    grid, kw = plated_box(12, 10, 16, 6, active=act)
    g, o = _pair_grid(grid)
    flat = kw["plate_active_flat_idx"]
    mics = np.array([[0.031, 0.052, 0.097], [0.083, 0.041, 0.032]], np.float32)
    axis = np.array([[0, 0, 1], [0.6, 0, 0.8]], np.float32)
    pa, pb, gain = np.float32([1.0, 0.4]), np.float32([0.0, 0.6]), np.float32([1.0, 2.5])
    rho_c = np.float32(np.float32(1.21) * np.float32(343.0))
    for m in (g, o):
        m.setup_plate(**kw)
        m.set_bridge_plate_sources(flat[[3, 4, 9]], [0.5, 0.3, 0.2])
        m.setup_mic_samplers(mics, axis, True)
        m.inject_pressure_nearest([0.061, 0.052, 0.113], 1.0)
    n = 12 * 10
    for blk in range(7):
        d = rng.standard_normal(3).astype(np.float32)
        for m in (g, o):
            m.inject_bridge_drive(d, 0.01)
            if blk % 3 == 1:
                m.inject_plate_force(flat[[1, 2]], [0.6, 0.4], 0.02)
        out, rc, done = g.record_mic_samples(5, 4, pa, pb, axis, gain, rho_c)
        assert rc == 0 and done == 5
        want = np.zeros((5, 2), np.float32)
        for j in range(5):
            o.step(4)
            sp, vx, vy, vz = o.sample_mics()
            sig = sp * pa
            vn = vx * axis[:, 0] + vy * axis[:, 1] + vz * axis[:, 2]
            sig = np.where(pb != 0, sig + (pb * rho_c) * vn, sig).astype(np.float32)
            want[j] = sig * gain
        assert np.array_equal(out.view(np.uint32), want.view(np.uint32)), blk
        _same(g, o)
        assert np.array_equal(_bits(g.get_plate_displacement(n)),
                              _bits(o.get_plate_displacement(n))), blk
    assert np.abs(o.get_plate_displacement(n)).max() > 0


def test_irregular_topology_unstaged_csr_rows_bitwise():
    """A topology whose 512-cell tiles hold more CSR entries than the cell
    kernel stages in LDS (1200 cells, 7000 faces: mean degree 11.7), so both
    the staged and the unstaged row walks run, against the oracle."""
    from afdtd.amr import AcousticAMR
    print("amr: irregular topology, unstaged CSR rows", flush=True)
    t = random_topology(23, nc=1200, nf=7000, walls=True)
    t["cell_volumes"] = t["cell_volumes"] * 3.0      # keep the order-2 step stable at this degree
    g, o = AcousticAMR(**t), OracleAMR(**t)
    assert g.n_csr_entries == o.n_csr_entries == 14000
    rng = np.random.default_rng(23)
    p = rng.standard_normal(g.n_cells).astype(np.float32)
    v = (rng.standard_normal(g.n_faces) * 1e-3).astype(np.float32)
    for m in (g, o):
        m.set_state(p, v)
        m.step(40)
    _same(g, o)
    assert np.abs(g.get_pressure()).max() > 0
