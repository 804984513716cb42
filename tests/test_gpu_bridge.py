"""The reference bridge's call sequence (acoustic_fdtd_bridge.py:648-784) on
this engine against the same sequence on the oracle, bit for bit: voxelise a
two-bout body at the factory's defaults (dx 4 mm, 12 cells of padding, PML
10), cut a soundhole, build the solver, set the cut-cell face fractions and
the Gaussian bridge sources, then per audio sample drive the bridge cells
and step steps_per_sample steps; read the pressure field, the volume slices,
the soundboard plane, the plate and world-space receivers, and reset.  Run
on the generic kernels and on the default choice, the single-step sweep with
cell geometry (walls, face fractions and an active plate)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("kernel", ["generic", "auto"])
def test_fdtd_coevolver_call_sequence_matches_oracle(oracle_mod, kernel):
    from afdtd.bridge import FDTDCoEvolver
    from afdtd.geometry import apply_soundhole_cutout, two_bout_outline
    from oracle import geometry as og
    poly = two_bout_outline(96)
    body_h, dx, pad, n_pml = 0.1, 0.004, 12, 10
    bridges = [(-0.02, -0.13), (0.0, -0.13), (0.02, -0.13)]
    hole = (0.0, 0.06, 0.04)
    ev = FDTDCoEvolver.from_outline(poly, body_h, bridges, soundhole=hole, dx=dx,
                                    pad_cells=pad, n_pml=n_pml, kernel=kernel)
    info = ev.info
    Nx, Ny, Nz = info["Nx"], info["Ny"], info["Nz"]
    print(f"bridge sequence: {Nx}x{Ny}x{Nz}, {len(ev.bridge_cells)} bridge cells, "
          f"{ev.steps_per_sample} steps per sample, kernel {ev.fdtd.kernel_family}", flush=True)
    assert ev.fdtd.kernel_family == ("generic" if kernel == "generic" else "fused")
    # the same sequence on the oracle
    ct, pa, g = og.voxelise(poly, body_h, dx, pad)
    apply_soundhole_cutout(ct, pa, g, hole)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, dx, c=343.0, rho_air=1.21, cell_type=ct.ravel(),
                              plate_iz=g["plate_iz"], plate_active=pa.ravel(),
                              plate_mass_density=7.0, plate_stiffness_D=0.45,
                              plate_alpha_M=2.0, plate_beta_K=1e-5, n_pml=n_pml,
                              precision="f32")
    o.set_face_fractions(*og.face_fractions(ct, pa, poly, g))
    from afdtd.geometry import build_bridge_kernel
    idx, wgt = build_bridge_kernel(bridges, info)
    assert np.array_equal(idx, ev.bridge_cells)
    o.set_bridge_sources(idx, wgt)
    assert ev.dt == o.dt
    assert np.array_equal(o.get_cell_type().reshape(Nx, Ny, Nz)[:, :, g["plate_iz"]] == 2,
                          pa.astype(bool))
    rng = np.random.default_rng(31)
    samples = 24
    drives = (1e-3 * rng.standard_normal((samples, len(idx)))).astype(np.float32)
    for blk in (drives[:10], drives[10:]):
        ev.drive_bridge_block(blk)
        for row in blk:
            o.inject_bridge_drive(row, 1.0)
            assert o.step(ev.steps_per_sample) == 0
        assert ev.fdtd.step_count == o.step_count
        P = o.get_pressure_field()
        assert np.array_equal(_bits(ev.get_pressure_field()), _bits(P))
        assert np.array_equal(_bits(ev.get_plate_displacement()),
                              _bits(o.get_plate_displacement()))
    assert np.abs(o.get_plate_displacement()).max() > 0 and np.abs(P).max() > 0
    for axis, k in ((0, -1), (1, -1), (2, -1), (1, 7), (2, 10 ** 6)):
        kk = P.shape[axis] // 2 if k < 0 else min(k, P.shape[axis] - 1)
        assert np.array_equal(_bits(ev.get_volume_slice(axis, k)), _bits(np.take(P, kk, axis)))
    assert np.array_equal(_bits(ev.get_soundboard_pressure()), _bits(P[:, :, g["plate_iz"]]))
    world = np.array([[0.0, -0.05, 0.05], [0.03, 0.12, 0.08], [-0.1, -0.2, 0.02],
                      [0.0, 0.06, 0.11]])
    gc = np.empty_like(world)
    for a, key in enumerate(("gx_min", "gy_min", "gz_min")):
        gc[:, a] = (world[:, a] - info[key]) / info["dx"]
    assert np.array_equal(_bits(ev.sample_at_receivers(world)),
                          _bits(o.sample_pressure(gc.astype(np.float32))))
    with pytest.raises(RuntimeError, match="disabled"):
        ev.push_driver_block(np.zeros(4, np.float32))
    ev.reset()
    assert not np.any(ev.get_pressure_field()) and ev.fdtd.step_count == 0
