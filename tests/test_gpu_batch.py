"""Batched rooms: every room of one graph-captured batch sweep equals the
oracle run of that room alone, bit for bit (FP32), with single steps and
with the two-step sweep (two steps per launch, odd step counts ending in a
single step); both sweeps agree bit for bit under 16-bit storage too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rooms(n, seed, lo=20, hi=45):
    rng = np.random.default_rng(seed)
    dims = rng.integers(lo, hi + 1, size=(n, 3))
    courant = rng.uniform(0.4, 0.55, n).astype(np.float32)
    refl = np.repeat(rng.uniform(0.6, 0.98, (n, 1)), 6, axis=1).astype(np.float32)
    return rng, dims, courant, refl


@pytest.mark.parametrize("graph,kernel,tile_y", [(True, "auto", 8), (False, "auto", 16),
                                                  (True, "fused", 8)])
def test_batch_rooms_match_oracle(oracle_mod, graph, kernel, tile_y):
    from afdtd import RoomBatch, differentiated_gaussian
    n, steps = 5, 60
    rng, dims, courant, refl = _rooms(n, 31)
    print(f"batch vs oracle: graph={graph} kernel={kernel} tile_y={tile_y}", flush=True)
    b = RoomBatch(dims, 0.01, courant, refl, c=343.0, rho_air=1.2, order=2, tile_y=tile_y,
                  seg_x=16, kernel=kernel)
    assert b.kernel_family == ("fused" if kernel == "fused" else "fused2")
    src = []
    rec = np.zeros((n, 8, 3), np.float32)
    for r, (Nx, Ny, Nz) in enumerate(dims):
        i, j, k = (rng.integers(4, d - 4) for d in (Nx, Ny, Nz))
        src.append(int(k + Nz * (j + Ny * i)))
        rec[r] = rng.uniform(3, np.array([Nx, Ny, Nz]) - 4, size=(8, 3))
    sig = np.stack([differentiated_gaussian(steps, 6.0, 20.0, 1.0)] * n)
    b.set_sources(src, sig)
    b.set_receivers(rec)
    b.record(steps)
    b.step(7, graph=graph)
    b.step(steps // 2 - 7, graph=graph)
    b.step(steps - steps // 2, graph=graph)
    tr = b.get_traces()
    assert tr.shape == (steps, n, 8)
    for r, (Nx, Ny, Nz) in enumerate(dims):
        o = oracle_mod.OracleFDTD(int(Nx), int(Ny), int(Nz), 0.01, c=343.0, rho_air=1.2,
                                  n_pml=0, order=2, courant=float(courant[r]), boundary=1,
                                  refl=refl[r], precision="f32")
        o.set_point_sources([src[r]], sig[r:r + 1])
        o.set_receivers(rec[r])
        o.record_receivers(steps)
        o.step(steps)
        to = o.get_receiver_traces().astype(np.float32)
        assert np.array_equal(tr[:, r, :].view(np.uint32), to.view(np.uint32)), f"room {r}"
        Pg = b.get_pressure(r)
        Po = o.get_pressure_field()
        assert np.array_equal(Pg.view(np.uint32), Po.view(np.uint32)), f"room {r} field"
        assert np.abs(to).max() > 0


def test_batch_autotune_keeps_results():
    """The autotuner changes the tile table only: after tuning (which must
    leave the batch at step 0 with zero fields) the traces equal those of an
    untuned batch bit for bit; tuning after a step is refused."""
    from afdtd import RoomBatch, differentiated_gaussian
    n, steps = 6, 40
    rng, dims, courant, refl = _rooms(n, 7, lo=18, hi=70)
    src = []
    rec = np.zeros((n, 4, 3), np.float32)
    for r, (Nx, Ny, Nz) in enumerate(dims):
        i, j, k = (rng.integers(3, d - 3) for d in (Nx, Ny, Nz))
        src.append(int(k + Nz * (j + Ny * i)))
        rec[r] = rng.uniform(2, np.array([Nx, Ny, Nz]) - 3, size=(4, 3))
    sig = np.stack([differentiated_gaussian(steps, 5.0, 15.0, 1.0)] * n)
    traces = []
    for tune in (False, True):
        b = RoomBatch(dims, 0.01, courant, refl, c=343.0, rho_air=1.2, order=2)
        b.set_sources(src, sig)
        if tune:
            choice = b.autotune(trial_steps=3)
            print("autotune choice", choice, flush=True)
            assert choice["tile_y"] in (8, 16, 32) and choice["seg_x"] in (16, 32, 64, 128)
            assert choice["kernel"] in ("fused", "fused2")
            assert b.step_count == 0
        b.set_receivers(rec)
        b.record(steps)
        b.step(steps)
        traces.append(b.get_traces())
        if tune:
            with pytest.raises(RuntimeError):
                b.autotune(trial_steps=2)
        del b
    assert np.abs(traces[0]).max() > 0
    assert np.array_equal(traces[0].view(np.uint32), traces[1].view(np.uint32))


@pytest.mark.parametrize("storage", ["f32", "bf16", "f16"])
def test_batch_two_step_equals_single_step(storage):
    """The two-step sweep equals two single steps bit for bit in every room,
    receivers and sources included (the first step's in the sweep)."""
    from afdtd import RoomBatch, differentiated_gaussian
    n, steps = 7, 45
    rng, dims, courant, refl = _rooms(n, 19, lo=18, hi=90)
    src = []
    rec = np.zeros((n, 5, 3), np.float32)
    for r, (Nx, Ny, Nz) in enumerate(dims):
        i, j, k = (rng.integers(2, d - 2) for d in (Nx, Ny, Nz))
        src.append(int(k + Nz * (j + Ny * i)))
        rec[r] = rng.uniform(1, np.array([Nx, Ny, Nz]) - 2, size=(5, 3))
    src[2] = -1                                                  # a room without a source
    sig = np.stack([differentiated_gaussian(steps, 4.0, 12.0, 1.0)] * n)
    out = {}
    for kernel in ("fused", "fused2"):
        print(f"batch {storage} {kernel}", flush=True)
        b = RoomBatch(dims, 0.01, courant, refl, c=343.0, rho_air=1.2, order=2, storage=storage,
                      tile_y=16, seg_x=32, kernel=kernel)
        b.set_sources(src, sig)
        b.set_receivers(rec)
        b.record(steps)
        for chunk in (1, 2, 13, 29):
            b.step(chunk)
        assert b.step_count == steps
        out[kernel] = (b.get_traces(), [b.get_pressure(r) for r in range(n)])
        del b
    ta, pa = out["fused"]
    tb, pb = out["fused2"]
    assert np.abs(ta).max() > 0
    assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
    for r in range(n):
        assert np.array_equal(pa[r].view(np.uint32), pb[r].view(np.uint32)), f"room {r}"

