"""Storage precision against the FP64 oracle (precision workload, crop and
rescale of the 1024 x 1024 x 512 room: same dx, source and receiver positions
mapped proportionally), and the block floating-point format (bfp16) against a
numpy model of it, bit for bit.

bfp16: int16 mantissas with one power-of-two exponent per 64-cell z-row
segment; the exponent is the smallest e with max|v| * 2^-e < 2^15 (from the
FP32 exponent field of max|v|: e = max(biased - 141, -128)) and the mantissa
rint(v * 2^-e) clamped to +-32767.

f16s (FP16 with a per-row power-of-two scale): the same segments and
exponent rule, the value v * 2^-e rounded to nearest even in FP16.

The FP64 comparison runs on the 256 x 256 x 128 crop for 3000 steps, the
precision workload's own accuracy set-up (about a minute of FP64 oracle time
on the host's cores), for every storage type on the single-step sweep and,
where it runs them, the two-step sweep.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def bfp_model(v, Nz, fmt="bfp16"):
    """Encode + decode a float32 field laid out (..., Nz) as bfp16 or f16s."""
    v = np.asarray(v, np.float32).reshape(-1, Nz)
    out = np.empty_like(v)
    for s0 in range(0, Nz, 64):
        seg = v[:, s0:s0 + 64]
        m = np.abs(seg).max(axis=1).astype(np.float32)
        biased = (m.view(np.uint32) >> 23) & 0xFF
        e = np.maximum(biased.astype(np.int64) - 141, -128)
        x = np.ldexp(seg.astype(np.float32), -e[:, None]).astype(np.float32)
        if fmt == "bfp16":
            q = np.clip(np.rint(x), -32767, 32767).astype(np.float32)
        else:
            q = x.astype(np.float16).astype(np.float32)
        out[:, s0:s0 + 64] = np.ldexp(q, e[:, None]).astype(np.float32)
    return out.ravel()


def test_bfp_model_properties():
    """The model itself: error at most 2^-16 of the segment maximum (plus the
    clamp), exact for values that fit, zero rows stay zero."""
    rng = np.random.default_rng(0)
    v = (rng.standard_normal((40, 70)) * np.exp(rng.uniform(-30, 10, (40, 1)))).astype(np.float32)
    v[3] = 0.0
    d = bfp_model(v, 70).reshape(40, 70)
    for s0 in (0, 64):
        seg, dd = v[:, s0:s0 + 64], d[:, s0:s0 + 64]
        m = np.abs(seg).max(axis=1, keepdims=True)
        assert np.all(np.abs(dd - seg) <= m * 2.0 ** -15 + 1e-45)
    assert np.all(d[3] == 0)
    ints = np.arange(-300, 300, dtype=np.float32)
    assert np.array_equal(bfp_model(np.tile(ints[:64], (2, 1)), 64), np.tile(ints[:64], 2))
    # f16s: relative error at most 2^-11 per value (values >= 2^-28 of the
    # row maximum), exact for values with at most 11 significant bits
    f = bfp_model(v, 70, "f16s").reshape(40, 70)
    for s0 in (0, 64):
        seg, ff = v[:, s0:s0 + 64], f[:, s0:s0 + 64]
        m = np.abs(seg).max(axis=1, keepdims=True)
        big = np.abs(seg) >= m * 2.0 ** -28
        assert np.all((np.abs(ff - seg) <= np.abs(seg) * 2.0 ** -11)[big])
    assert np.all(f[3] == 0)
    odd = (np.arange(64, dtype=np.float32) * 2 + 1) * 2.0 ** -20
    assert np.array_equal(bfp_model(odd, 64, "f16s"), odd)


@pytest.mark.parametrize("fmt", ["bfp16", "f16s"])
def test_bfp16_roundtrip_matches_model(fmt):
    """set_state -> get_* through the GPU encoder / decoder equals the model
    bit for bit, on a grid with a partial last z-segment (Nz = 70)."""
    from afdtd import AcousticFDTD
    Nx, Ny, Nz = 9, 11, 70
    rng = np.random.default_rng(1)
    scale = np.exp(rng.uniform(-20, 8, (Nx * Ny, 1)))

    def field(n_lines, nz):
        return (rng.standard_normal((n_lines, nz)) * scale[:n_lines]).astype(np.float32)

    P = field(Nx * Ny, Nz)
    P[5] = 0.0
    vx = field((Nx - 1) * Ny, Nz)
    vy = np.stack([field(Ny - 1, Nz) for _ in range(Nx)])
    vz = field(Nx * Ny, Nz - 1)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, cell_type=np.zeros(Nx * Ny * Nz, np.uint8), plate_iz=3,
                     plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, order=2, courant=0.5,
                     storage=fmt)
    assert g.kernel_family == "fused"
    g.set_state(P.ravel(), vx.ravel(), vy.ravel(), vz.ravel())
    got = g.get_pressure_field().ravel()
    assert np.array_equal(got.view(np.uint32), bfp_model(P, Nz, fmt).view(np.uint32))
    assert np.array_equal(g.get_velocity_x().view(np.uint32),
                          bfp_model(vx, Nz, fmt).view(np.uint32))
    # Vy: the padded layout holds rows j <= Ny-2 of each x-plane (row Ny-1 is
    # zero and shares no segment with them)
    assert np.array_equal(g.get_velocity_y().view(np.uint32),
                          bfp_model(vy.reshape(-1, Nz), Nz, fmt).view(np.uint32))
    # Vz: the padded layout stores Nz-1 faces in rows of Nz (last slot zero)
    vz_pad = np.zeros((Nx * Ny, Nz), np.float32)
    vz_pad[:, :Nz - 1] = vz
    exp_vz = bfp_model(vz_pad, Nz, fmt).reshape(-1, Nz)[:, :Nz - 1].ravel()
    assert np.array_equal(g.get_velocity_z().view(np.uint32), exp_vz.view(np.uint32))
    g.close()


@pytest.mark.parametrize("fmt", ["bfp16", "f16s"])
def test_bfp16_rejects_generic_only_features(fmt):
    from afdtd import AcousticFDTD
    Nx = Ny = Nz = 12
    ct = np.zeros(Nx * Ny * Nz, np.uint8)
    ct[100] = 1                                           # a wall cell
    with pytest.raises(RuntimeError):
        AcousticFDTD(Nx, Ny, Nz, 0.01, cell_type=ct, plate_iz=3,
                     plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, storage=fmt)
    with pytest.raises(RuntimeError):                     # no two-step sweep for block storage
        AcousticFDTD(Nx, Ny, Nz, 0.01, cell_type=np.zeros(Nx * Ny * Nz, np.uint8), plate_iz=3,
                     plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, order=2, courant=0.5,
                     storage=fmt, kernel="fused2")


# Receiver rel-L2 bound against FP64 and final-energy bound per storage and
# sweep on the 256 x 256 x 128 crop after 3000 steps.  The error is set by
# the stored mantissa: FP32 (24 bits) ~6e-6, bfp16 (15 bits below a shared
# row exponent) ~7e-4, FP16 and f16s (11 bits) ~5e-3, BF16 (8 bits)
# ~3.4e-2; each bound leaves about 1.5x headroom over the measured value and
# is below the next wider format's error, so a storage that silently fell
# back to a narrower one fails.  The two-step sweep rounds its intermediate
# step to the storage type exactly as two single steps do, so its errors are
# the single-step ones.
BOUNDS = {("f32", "fused"): (1e-5, 1e-5), ("f32", "fused2"): (1e-5, 1e-5),
          ("bfp16", "fused"): (1e-3, 1e-4), ("f16s", "fused"): (8e-3, 3e-3),
          ("f16", "fused"): (8e-3, 3e-3), ("f16", "fused2"): (8e-3, 3e-3),
          ("bf16", "fused"): (5e-2, 4e-2), ("bf16", "fused2"): (5e-2, 4e-2)}


def test_storage_accuracy_vs_fp64(oracle_mod):
    """Receiver traces and final energy of every storage type against the
    FP64 oracle on the precision crop (256 x 256 x 128, 3000 steps), each
    within the bound its mantissa width allows (BOUNDS)."""
    from afdtd import AcousticFDTD, workloads as W
    Nx, Ny, Nz, steps = 256, 256, 128, 3000
    w = W.config4(Nx=Nx, Ny=Ny, Nz=Nz, steps=steps)
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    sig = w.signals(steps)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct,
                              plate_iz=Nz // 2, plate_active=pa, n_pml=0, order=2, courant=0.5,
                              boundary=1, refl=w.refl, precision="f64")
    o.set_point_sources(w.src, sig)
    o.set_receivers(w.rec)
    o.record_receivers(steps)
    o.step(steps)
    ref = o.get_receiver_traces().astype(np.float64)
    eo = o.energy()               # [P^n P^(n-1)/2K, kinetic, P^2/2K]
    e_ref = eo[1] + eo[2]
    del o
    errs = {}
    for st, kern in BOUNDS:
        g = AcousticFDTD(Nx, Ny, Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct, plate_iz=Nz // 2,
                         plate_active=pa, n_pml=0, storage=st, kernel=kern, **w.solver_kwargs())
        assert g.kernel_family == kern
        g.set_point_sources(w.src, sig)
        g.set_receivers(w.rec)
        g.record_receivers(steps)
        g.step(steps)
        tr = g.get_receiver_traces().astype(np.float64)
        e = sum(g.energy())
        errs[st, kern] = (rel_l2(tr, ref), abs(e - e_ref) / e_ref)
        g.close()
        print(f"{st:6s} {kern:6s} {Nx}x{Ny}x{Nz} {steps} steps: receiver rel-L2 vs FP64 "
              f"{errs[st, kern][0]:.3e}, final energy rel. diff {errs[st, kern][1]:.3e}",
              flush=True)
    for key, (bt, be) in BOUNDS.items():
        assert errs[key][0] <= bt, (key, errs[key])
        assert errs[key][1] <= be, (key, errs[key])


@pytest.mark.parametrize("storage", ["f16", "bf16", "f32"])
@pytest.mark.parametrize("hetero", [False, True])
def test_fused_16bit_close_to_generic(storage, hetero):
    """The fused sweep with 16-bit storage stays within rounding of the
    generic sweep (not bit for bit: the fused divergence uses the new
    velocities before they are rounded to the storage type, the generic one
    reads them back).  Catches a storage scale applied to the wrong field."""
    from afdtd import AcousticFDTD
    Nx, Ny, Nz = 20, 24, 70
    rng = np.random.default_rng(3)
    kw = dict(order=2, courant=0.5, boundary="impedance", refl=[0.9] * 6)
    if hetero:
        kw["rho"] = (1.2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
        kw["bulk"] = (1.2 * 343.0 ** 2 * np.exp(0.1 * rng.standard_normal(Nx * Ny * Nz))).astype(np.float32)
    P0 = rng.standard_normal(Nx * Ny * Nz).astype(np.float32)
    out = {}
    for kern in ("generic", "fused"):
        g = AcousticFDTD(Nx, Ny, Nz, 0.01, cell_type=np.zeros(Nx * Ny * Nz, np.uint8), plate_iz=5,
                         plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, storage=storage,
                         kernel=kern, **kw)
        g.set_state(P0)
        g.step(40)
        out[kern] = (g.get_pressure_field().ravel(), g.get_velocity_z())
        g.close()
    tol = {"f32": 1e-6, "f16": 5e-3, "bf16": 4e-2}[storage]
    for a, b in zip(out["generic"], out["fused"]):
        assert np.all(np.isfinite(b))
        err = rel_l2(b, a)
        print(f"{storage} het={hetero}: fused vs generic rel-L2 {err:.2e}", flush=True)
        assert err <= tol
