"""The two-step sweep (kernel "fused2": stage A n -> n+1 on the tile plus a
one-cell ring, stage B n+1 -> n+2 on the tile, one HBM pass per two steps)
against the float32 oracle bit for bit, and, with 16-bit storage, against
the single-step fused sweep bit for bit.

Covered: ragged grids (Ny, Nz not multiples of the 16 x 64 tile, Nx not a
multiple of the x-segment, tiny segments), heterogeneous media, impedance
walls, PML, odd step counts (a single step finishes the call), point sources
at tile edges and corners (applied by every tile whose ring holds the cell),
receiver taps of the intermediate step, the stability check at step 256 and
the failure restore of a pair.
"""
import numpy as np
import pytest

from conftest import rel_l2
from test_gpu_parity import _assert_fields_equal, _pair, _random_state

pytestmark = pytest.mark.gpu


def _media(Nx, Ny, Nz, seed):
    rng = np.random.default_rng(seed)
    rho = (1.2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    bulk = (1.2 * 343.0 ** 2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    return rho, bulk


def _make(oracle_mod, Nx, Ny, Nz, hetero, imped, n_pml, seg_x, kernel="fused2", storage="f32",
          seed=12, tile_y=0):
    kw = dict(order=2, courant=0.5, n_pml=n_pml)
    if imped:
        kw.update(boundary="impedance", refl=[0.6, 0.7, 0.8, 0.85, 0.9, 0.95])
    if hetero:
        kw["rho"], kw["bulk"] = _media(Nx, Ny, Nz, seed)
    g, o = _pair(oracle_mod, Nx, Ny, Nz, **kw)
    g.close()
    from afdtd import AcousticFDTD
    ct = np.zeros(Nx * Ny * Nz, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.21, cell_type=ct, plate_iz=Nz // 2,
                     plate_active=pa, kernel=kernel, storage=storage, seg_x=seg_x,
                     tile_y=tile_y, **kw)
    assert g.kernel_family == kernel
    return g, o


@pytest.mark.parametrize("hetero,imped,n_pml", [(False, False, 0), (True, True, 0),
                                                (False, True, 0), (True, False, 3)])
@pytest.mark.parametrize("seg_x,tile_y", [(64, 16), (5, 16), (1, 16), (64, 8), (3, 8)])
def test_fused2_bitwise(oracle_mod, hetero, imped, n_pml, seg_x, tile_y):
    """tile_y 16: two extended rows per thread; 8: one row per thread."""
    Nx, Ny, Nz = 23, 37, 70
    print(f"fused2 bitwise: het={hetero} imp={imped} pml={n_pml} seg={seg_x} ty={tile_y}", flush=True)
    g, o = _make(oracle_mod, Nx, Ny, Nz, hetero, imped, n_pml, seg_x, tile_y=tile_y)
    _random_state(o, g, seed=13)
    for chunk in (1, 2, 9, 30):
        g.step(chunk); o.step(chunk)
        _assert_fields_equal(g, o, f"fused2 het={hetero} imp={imped} pml={n_pml} "
                                   f"seg={seg_x} n={g.step_count}")


def test_fused2_multi_tile_grid(oracle_mod):
    """Several tiles in y and z (3 x 3) and x-segments, odd total."""
    Nx, Ny, Nz = 19, 40, 150
    print("fused2 multi-tile 19x40x150", flush=True)
    g, o = _make(oracle_mod, Nx, Ny, Nz, True, True, 0, 7)
    _random_state(o, g, seed=21)
    g.step(13); o.step(13)
    _assert_fields_equal(g, o, "fused2 multi-tile")


@pytest.mark.parametrize("tile_y", [16, 8])
def test_fused2_sources_receivers_traces(oracle_mod, tile_y):
    """Sources on tile edges / corners and in the interior, receivers at
    fractional and cell-centre positions; 300 steps cross the stability check
    at 256, so one single step is scheduled before it."""
    from afdtd import differentiated_gaussian
    Nx, Ny, Nz = 30, 34, 130
    print("fused2 sources/receivers 30x34x130", flush=True)
    g, o = _make(oracle_mod, Nx, Ny, Nz, True, True, 0, 11, tile_y=tile_y)
    cells = [(5, 15, 63), (6, 16, 64), (7, 15, 64), (10, 16, 63), (12, 0, 0), (20, 33, 129),
             (15, 17, 65), (9, 31, 127), (5, 15, 63)]          # last: same cell twice
    src = [int(k + Nz * (j + Ny * i)) for i, j, k in cells]
    sig = np.stack([differentiated_gaussian(300, 6.0 + s, 20.0 + 3 * s, 0.5 + 0.1 * s)
                    for s in range(len(src))])
    rng = np.random.default_rng(5)
    rec = rng.uniform(1, [Nx - 2, Ny - 2, Nz - 2], size=(24, 3)).astype(np.float32)
    rec[:6] = np.array([[5, 15, 63], [6, 16, 64], [7, 15.5, 63.5], [3, 0, 0],
                        [29, 33, 129], [10, 16.2, 64.7]], np.float32)
    for m in (g, o):
        m.set_point_sources(src, sig)
        m.set_receivers(rec)
        m.record_receivers(300)
    for chunk in (1, 200, 99):
        g.step(chunk); o.step(chunk)
    tg, to = g.get_receiver_traces(), o.get_receiver_traces()
    assert tg.shape == to.shape == (300, 24)
    assert rel_l2(tg, to) <= 1e-5
    bad = np.argwhere(tg.view(np.uint32) != to.astype(np.float32).view(np.uint32))
    assert bad.size == 0, f"trace mismatch at (step, receiver) {bad[:5].tolist()}"
    assert np.abs(to).max() > 1e-4
    _assert_fields_equal(g, o, "fused2 sources")


@pytest.mark.parametrize("storage", ["bf16", "f16"])
def test_fused2_16bit_equals_single_step(oracle_mod, storage):
    """With 16-bit storage the two-step sweep rounds P^{n+1} and V^{n+1/2}
    where the single-step sweep stores and reloads them: equal bit for bit."""
    from afdtd import differentiated_gaussian
    Nx, Ny, Nz = 21, 35, 70
    print(f"fused2 {storage} vs fused", flush=True)
    outs = []
    for kernel in ("fused", "fused2"):
        g, o = _make(oracle_mod, Nx, Ny, Nz, True, True, 0, 6, kernel=kernel, storage=storage)
        src = [int(33 + Nz * (17 + Ny * 10)), int(64 + Nz * (16 + Ny * 5))]
        sig = np.stack([differentiated_gaussian(120, 6.0, 20.0, 1.0)] * 2)
        rec = np.array([[10, 17, 33], [5.5, 16.2, 63.7], [12, 30, 5]], np.float32)
        g.set_point_sources(src, sig)
        g.set_receivers(rec)
        g.record_receivers(120)
        g.step(1); g.step(119)
        outs.append((g.get_pressure_field().ravel(), g.get_velocity_x(), g.get_velocity_y(),
                     g.get_velocity_z(), g.get_receiver_traces()))
        g.close()
    for name, a, b in zip(("P", "Vx", "Vy", "Vz", "traces"), outs[0], outs[1]):
        nbad = int(np.sum(a.view(np.uint32) != b.view(np.uint32)))
        assert nbad == 0, f"{storage} {name}: {nbad} values differ between fused and fused2"
    assert np.abs(outs[0][4]).max() > 1e-4


def test_fused2_unstable_restores_buffers(oracle_mod):
    """A failing stability check stops the two-step sweep at step 256 with
    the oracle's fields (pair / single launch parity restored)."""
    g, o = _make(oracle_mod, 20, 20, 20, False, False, 0, 32)
    sig = np.full((1, 401), 1e6, np.float32)
    src = [10 + 20 * (10 + 20 * 10)]
    for m in (g, o):
        m.set_point_sources(src, sig)
    assert o.step(401) == -3
    with pytest.raises(RuntimeError, match="FDTD_ERR_UNSTABLE"):
        g.step(401)
    assert g.step_count == o.step_count == 256
    _assert_fields_equal(g, o, "fused2 unstable")
