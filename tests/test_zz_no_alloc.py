"""The allocation-free sampler entry point through the C ABI:
fdtd_sample_pressure_precomputed_noalloc equals
fdtd_sample_pressure_precomputed bit for bit on two calls in a row, the
second with fewer receivers (the scratch array keeps the first call's
contents, none of which may leak into the second), on two handles of the
same GPU, and rejects more receivers than its maximum and out-of-range
indices."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


def test_noalloc_sampler_matches_allocating_entry_point():
    from afdtd import AcousticFDTD, differentiated_gaussian, lib
    L = lib()
    Nx, Ny, Nz = 40, 36, 70
    print(f"no-alloc sampler: {Nx}x{Ny}x{Nz}", flush=True)
    rng = np.random.default_rng(77)
    handles = []
    for h in range(2):
        g = AcousticFDTD(Nx, Ny, Nz, 0.01, cell_type=np.zeros(Nx * Ny * Nz, np.uint8),
                         plate_iz=Nz // 2, plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=4)
        g.set_point_sources([30 + Nz * (17 + Ny * 20)], differentiated_gaussian(60)[None, :])
        g.step(60 + 7 * h)
        handles.append(g)
    for g in handles:
        for n in (300, 41):
            xyz = rng.uniform([0, 0, 0], [Nx - 1, Ny - 1, Nz - 1], (n, 3)).astype(np.float32)
            idx8, wgt8 = g.precompute_pressure_samplers(xyz)
            idx8 = np.ascontiguousarray(idx8, np.int32)
            wgt8 = np.ascontiguousarray(wgt8, np.float32)
            ref = np.zeros(n, np.float32)
            got = np.full(n + 5, np.nan, np.float32)
            assert L.fdtd_sample_pressure_precomputed(g._h, n, _vp(idx8), _vp(wgt8), _vp(ref)) == 0
            assert L.fdtd_sample_pressure_precomputed_noalloc(g._h, n, _vp(idx8), _vp(wgt8),
                                                              _vp(got)) == 0
            assert np.abs(ref).max() > 0
            assert np.array_equal(got[:n].view(np.uint32), ref.view(np.uint32)), n
            assert np.all(np.isnan(got[n:]))            # nothing written past n outputs
    g = handles[0]
    n = 65537
    idx8 = np.zeros(8 * n, np.int32)
    wgt8 = np.zeros(8 * n, np.float32)
    out = np.zeros(n, np.float32)
    assert L.fdtd_sample_pressure_precomputed_noalloc(g._h, n, _vp(idx8), _vp(wgt8), _vp(out)) == -2
    bad = np.zeros(8, np.int32); bad[3] = Nx * Ny * Nz
    assert L.fdtd_sample_pressure_precomputed_noalloc(g._h, 1, _vp(bad), _vp(wgt8[:8]),
                                                      _vp(out[:1])) == -2
    assert L.fdtd_sample_pressure_precomputed_noalloc(g._h, 0, _vp(idx8), _vp(wgt8), _vp(out)) == 0
    for g in handles:
        g.close()
