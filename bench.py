"""Benchmark: cell-updates/s and HBM-roofline fraction of the FDTD step on one
MI355X, with the CPU oracle timed beside it on the host cores.

A "step" is one full leapfrog step (velocity + pressure half steps, sources,
receiver gather) over the whole grid.  Default workload: the single large
room, 512x384x256 cells, order-2 stencil at Courant 0.55, heterogeneous
density and bulk modulus, locally reacting walls with per-side reflection
coefficients, 4 sources and 64 receivers (synthetic, seeded).

The default line also carries `forms`: the batched, precision and adjoint
forms, the reference's own order-4 scheme and the voxelised body, each
measured in the same run at a reduced step count (--no-forms skips them).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "spectral-analyzer_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def make_workload(name, steps):
    from afdtd import workloads as W
    if name == "room":
        return W.config2(steps=steps)
    if name == "room-order4-pml":
        return W.config2_order4_pml(steps=steps)
    if name == "c1":
        return W.config1(parity=False)
    if name == "c1-parity":
        return W.config1(parity=True)
    if name == "precision":
        return W.config4(steps=steps)
    if name == "precision-crop":
        return W.config4(Nx=256, Ny=256, Nz=128, steps=steps)
    raise SystemExit(f"unknown config {name}")


def make_solver(w, storage="f32", kernel="auto", tile_y=0, seg_x=0):
    from afdtd import AcousticFDTD
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(w.Nx * w.Ny, np.uint8)
    return AcousticFDTD(w.Nx, w.Ny, w.Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct,
                        plate_iz=w.Nz // 2, plate_active=pa, n_pml=w.n_pml,
                        storage=storage, kernel=kernel, tile_y=tile_y, seg_x=seg_x,
                        **w.solver_kwargs())


def copy_calibration(nbytes=2 << 30, reps=20):
    """Measured copy ceiling: the library's 16-byte-per-lane device copy of a
    known size (also used to scale the HBM byte counters of profile runs)."""
    import ctypes as C
    from afdtd import lib
    ms = C.c_float()
    rc = lib().fdtd_copy_calibration(nbytes, reps, C.byref(ms))
    if rc:
        raise SystemExit(f"fdtd_copy_calibration failed: {rc}")
    out = {"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms_per_copy": ms.value,
           "GBps_read_plus_write": 2 * nbytes / (ms.value * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)
    return out


def stream_ceiling(w, storage, reps=10):
    """Measured streaming ceiling of the sweep's byte mix: the library's
    no-stencil kernel reading and writing FP32 arrays with the same bytes per
    cell as the algorithmic count (f32: 4 or 6 arrays read, 4 written; 16-bit
    fields: half as many field arrays, rho and K unchanged)."""
    import ctypes as C
    from afdtd import lib
    if storage == "f32":
        n_read, n_write = (6, 4) if w.hetero else (4, 4)
    else:
        n_read, n_write = (4, 2) if w.hetero else (2, 2)
    n = min(w.cells, 64 << 20)
    ms = C.c_float()
    rc = lib().fdtd_stream_calibration(n, n_read, n_write, reps, C.byref(ms))
    if rc:
        raise SystemExit(f"fdtd_stream_calibration failed: {rc}")
    return (n_read + n_write) * 4 * n / (ms.value * 1e-3) / 1e9


def cpu_baseline(w, budget_s=15.0):
    """Time the oracle (float32, OpenMP over x slabs) on the same grid for a
    bounded number of steps."""
    from oracle.oracle import OracleFDTD
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    ct = np.zeros(w.cells, np.uint8)
    pa = np.zeros(w.Nx * w.Ny, np.uint8)
    o = OracleFDTD(w.Nx, w.Ny, w.Nz, w.dx, c=343.0, rho_air=1.2, cell_type=ct,
                   plate_iz=w.Nz // 2, plate_active=pa, n_pml=w.n_pml, order=w.order,
                   courant=w.courant or 0.0, boundary=1 if w.boundary == "impedance" else 0,
                   refl=w.refl, rho=w.rho, bulk=w.bulk, precision="f32")
    o.set_point_sources(w.src, w.signals(64))
    o.step(1)                                   # thread pool start-up, first touch
    t0 = time.perf_counter()
    o.step(2)
    per = (time.perf_counter() - t0) / 2
    n = int(max(2, min(2000, budget_s / max(per, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    el = time.perf_counter() - t0
    del o
    return {"value": w.cells * n / el, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} steps of the same {w.Nx}x{w.Ny}x{w.Nz} grid "
                      f"({el:.1f} s, float32 oracle, OpenMP over x slabs)"}


def measured_traffic(workload, kernel, storage):
    """HBM bytes per launch of the stencil kernel for this workload, as
    measured with rocprofv3 FETCH_SIZE/WRITE_SIZE scaled by the copy
    calibration (the `traffic_bytes_per_launch:` line of the matching summary
    in profiles/), or None when no measurement of this configuration exists."""
    import glob
    suffix = {"config2-room-512x384x256-hetero-order2": "room_512x384x256_hetero_order2",
              "room-512x384x256-order4-pml10": "room_512x384x256_order4_pml10"}.get(workload)
    if suffix is None or kernel not in ("fused", "fused2"):
        return None
    key = f"{kernel}_{suffix}" + ("" if storage == "f32" else f"_{storage}")
    for path in glob.glob(os.path.join(ROOT, "profiles", key + ".txt")):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


def bench_batch(a):
    """Config 3: one graph-captured batched sweep over 256 rooms of uneven
    size against the same rooms stepped one handle at a time."""
    from afdtd import RoomBatch, AcousticFDTD, workloads as W
    w = W.config3(steps=a.steps)
    dims, steps = w["dims"], w["steps"]
    cells = int(np.prod(dims, axis=1).sum())
    b = RoomBatch(dims, 0.01, w["courant"], w["refl"], c=343.0, rho_air=1.2, order=2,
                  tile_y=a.tile_y, seg_x=a.seg_x)
    tuned = None
    if not (a.tile_y or a.seg_x):
        tuned = b.autotune(trial_steps=10)            # before the first step
    b.set_sources(w["src"], w["signals"])
    b.set_receivers(w["rec"])
    b.record(steps + 10)
    b.step(10)                                       # warm-up (graph path)
    t0 = time.perf_counter()
    ms_dev = b.step(steps, graph=True)
    wall_b = time.perf_counter() - t0
    tr_b = b.get_traces()
    del b
    # the same rooms, one plain handle each (per-room launches)
    t_rooms, check = 0.0, []
    for r, (Nx, Ny, Nz) in enumerate(dims):
        Nx, Ny, Nz = int(Nx), int(Ny), int(Nz)
        s = AcousticFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.2,
                         cell_type=np.zeros(Nx * Ny * Nz, np.uint8), plate_iz=Nz // 2,
                         plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, order=2,
                         courant=float(w["courant"][r]), boundary="impedance", refl=w["refl"][r],
                         tile_y=a.tile_y, seg_x=a.seg_x)
        s.set_point_sources([w["src"][r]], w["signals"][r:r + 1])
        s.set_receivers(w["rec"][r])
        s.record_receivers(steps + 10)
        s.step(10)
        t0 = time.perf_counter()
        s.step(steps)
        t_rooms += time.perf_counter() - t0
        if r < 8:
            check.append(np.array_equal(s.get_receiver_traces().view(np.uint32),
                                        tr_b[:, r, :].view(np.uint32)))
        del s
    line = {"metric": "cell-updates/s (batched rooms, 1 GPU)", "value": cells * steps / wall_b,
            "unit": "cell-updates/s", "n_gpus": 1, "steps": steps, "warmup": 10,
            "ms_per_step": wall_b * 1e3 / steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded rooms)",
            "config": {"workload": "config3-256-rooms-48to160-order2", "rooms": int(len(dims)),
                       "cells": cells, "graph_device_ms": ms_dev, "autotuned": tuned},
            "per_room_launches": {"value": cells * steps / t_rooms, "unit": "cell-updates/s",
                                  "seconds": t_rooms},
            "speedup_vs_per_room": t_rooms / wall_b,
            "traces_bitwise_equal_first8": bool(all(check))}
    print(json.dumps(line), flush=True)
    return 0


def bench_geometry(a):
    """Geometry input formats: voxelise a two-bout body (256-vertex outline,
    rib height 0.1 m) at dx = 1 mm with 64 cells of padding and compute its
    cut-cell face fractions, from and into device memory; the oracle's numpy
    restatement of the reference's host code is the CPU baseline."""
    import ctypes as C
    import torch
    from afdtd import lib
    from afdtd.geometry import body_grid, two_bout_outline
    poly = two_bout_outline(256)
    body_h, dx, pad = 0.1, 0.001, 64
    g = body_grid(poly, body_h, dx, pad)
    Nx, Ny, Nz = g["Nx"], g["Ny"], g["Nz"]
    N = Nx * Ny * Nz
    L = lib()
    dpoly = torch.tensor(poly, device="cuda")
    dct = torch.empty(N, dtype=torch.uint8, device="cuda")
    dpa = torch.empty(Nx * Ny, dtype=torch.uint8, device="cuda")
    fr = [torch.empty(n, dtype=torch.float32, device="cuda")
          for n in ((Nx - 1) * Ny * Nz, Nx * (Ny - 1) * Nz, Nx * Ny * (Nz - 1))]
    torch.cuda.synchronize()
    p = lambda t: C.c_void_p(t.data_ptr())

    def once():
        m1, m2 = C.c_float(0.0), C.c_float(0.0)
        rc = L.fdtd_voxelise_body_device(p(dpoly), len(poly), Nx, Ny, Nz, g["gx_min"],
                                         g["gy_min"], g["gz_min"], dx, 0.0, body_h, g["iz_back"],
                                         g["plate_iz"], p(dct), p(dpa), None, C.byref(m1))
        rc = rc or L.fdtd_cut_cell_fractions_device(
            p(dct), p(dpa), p(dpoly), len(poly), Nx, Ny, Nz, g["gx_min"], g["gy_min"],
            g["gz_min"], dx, g["plate_iz"], g["iz_back"], p(fr[0]), p(fr[1]), p(fr[2]), None,
            C.byref(m2))
        if rc:
            raise RuntimeError(f"geometry entry points failed (rc={rc})")
        return m1.value, m2.value

    steps = a.steps if a.steps != 4000 else 50
    for _ in range(min(a.warmup, 5)):
        once()
    t0 = time.perf_counter()
    ms = [once() for _ in range(steps)]
    wall = time.perf_counter() - t0
    vox_ms = sum(m[0] for m in ms) / steps
    frac_ms = sum(m[1] for m in ms) / steps
    # algorithmic bytes per cell: cell type written (1), three face fractions
    # written (12), the cell type read once per face array (3)
    bpc = 16
    kern_s = (vox_ms + frac_ms) * 1e-3
    achieved = N * bpc / kern_s / 1e9
    line = {"metric": "voxelised cells/s (cell types + cut-cell face fractions, 1 GPU)",
            "value": N / kern_s, "unit": "cells/s", "n_gpus": 1, "steps": steps,
            "warmup": min(a.warmup, 5), "ms_per_step": vox_ms + frac_ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8/f32",
            "data": "synthetic (two-bout body outline)",
            "config": {"workload": f"geometry-two-bout-{Nx}x{Ny}x{Nz}-dx1mm", "cells": N,
                       "outline_vertices": len(poly), "plate_nodes": int(dpa.sum().item())},
            "voxelise_ms": vox_ms, "fractions_ms": frac_ms,
            "wall_ms_per_call_pair": wall * 1e3 / steps,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "bytes_per_cell": bpc}}
    if not a.no_cpu:
        from oracle import geometry as OG
        t0 = time.perf_counter()
        ct, pa, og = OG.voxelise(poly, body_h, dx, pad)
        OG.face_fractions(ct, pa, poly, og)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": N / dt, "unit": "cells/s", "cores": 1, "kind": "port",
                                "sample": f"the same grid once ({N} cells, {dt:.1f} s, numpy "
                                          f"restatement of the reference's host code)"}
        line["equal_to_oracle"] = bool(
            np.array_equal(dct.cpu().numpy(), ct.ravel())
            and all(np.array_equal(f.cpu().numpy(), o)
                    for f, o in zip(fr, OG.face_fractions(ct, pa, poly, og))))
    print(json.dumps(line), flush=True)
    return 0


def bench_adjoint(a):
    """Config 5: 128^3, 2000 steps, checkpoints every 100: forward + adjoint
    (FP32) timing; float64 central differences on 64 random cells."""
    import torch
    from afdtd import workloads as W
    from afdtd.adjoint import AdjointSolver, simulate_traces
    w = W.config5(steps=a.steps)
    shape, steps = w["shape"], w["steps"]
    N = int(np.prod(shape))
    res = {}
    dtypes = (("float32", torch.float32),) if a.no_fd else (("float32", torch.float32),
                                                             ("float64", torch.float64))
    for dtype, tdt in dtypes:
        s = AdjointSolver(*shape, w["dtdx"], dtype=dtype, checkpoint_every=100)
        s.set_sources(w["src"], w["signals"])
        s.set_receivers(w["rec"])
        dev = lambda x: torch.tensor(np.ascontiguousarray(x).ravel(), dtype=tdt, device="cuda")
        with torch.no_grad():
            target = s.forward(dev(w["m"] + w["pert"][0]), dev(w["l"] + w["pert"][1]), dev(w["eta"]), steps).clone()
        mg, lg, eg = (dev(x).requires_grad_() for x in (w["m"], w["l"], w["eta"]))
        # untimed pass: allocates the checkpoint and segment stores
        ((simulate_traces(s, mg, lg, eg, steps) - target) ** 2).sum().backward()
        mg.grad = lg.grad = eg.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = simulate_traces(s, mg, lg, eg, steps)
        loss = ((tr - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res[dtype] = dict(solver=s, grads=[g.grad.double().cpu().numpy() for g in (mg, lg, eg)],
                          fwd=t1 - t0, bwd=t2 - t1, dev=dev, target=target)
    bwd = res["float32"]["bwd"]
    line = {"metric": "adjoint reverse-step cell-updates/s (1 GPU)",
            "value": N * steps / bwd, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
            "higher_is_better": True, "dtype": "fp32", "data": "synthetic (seeded media)",
            "config": {"workload": "config5-128cube-2000steps-ckpt100", "cells": N},
            "forward_s_fp32": res["float32"]["fwd"], "backward_s_fp32": bwd}
    if a.no_fd:
        print(json.dumps(line), flush=True)
        return 0
    # central differences (float64) on 64 random cells
    r64 = res["float64"]
    s, dev, target = r64["solver"], r64["dev"], r64["target"]
    rng = np.random.default_rng(55)
    cells = rng.choice(N, 64, replace=False)
    worst, h = 0.0, 1e-5
    gm = r64["grads"][0]
    for q, c in enumerate(cells):
        which = q % 2                                   # alternate log rho / log K
        g = r64["grads"][which]
        vals = []
        for sgn in (1, -1):
            p = [w["m"].ravel().copy(), w["l"].ravel().copy()]
            p[which][c] += sgn * h
            with torch.no_grad():
                tr = s.forward(dev(p[0]), dev(p[1]), dev(w["eta"]), steps)
                vals.append(float(((tr - target) ** 2).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        worst = max(worst, abs(g[c] - fd) / max(abs(fd), 1e-3 * np.abs(g).max()))
    g32, g64 = res["float32"]["grads"], res["float64"]["grads"]
    rel32 = [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(g32, g64)]
    line["fd_worst_rel_err_fp64_64cells"] = worst
    line["fp32_vs_fp64_grad_rel_l2"] = rel32
    print(json.dumps(line), flush=True)
    return 0


def bench_amr(a):
    """AMR (topology-driven CSR) stepper, gradient order 2: a 256^3 uniform
    box grid (16.8 M cells, 50.1 M faces) from AcousticAMRGrid.build_uniform_box
    with an 8-cell PML-typed shell under the anechoic border condition, one
    pressure pulse; the order-2 AMR oracle (C, OpenMP) is the CPU baseline."""
    from afdtd.amr import AcousticAMR, AcousticAMRGrid
    n, dx, n_pml = 256, 0.01, 8
    grid = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=n * dx, y_min=0, y_max=n * dx,
                                             z_min=0, z_max=n * dx, dx=dx).with_pml_shell(n_pml)
    order = a.amr_order
    args = (grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction, grid.cell_types,
            grid.face_cell_neg, grid.face_cell_pos, grid.face_area, grid.face_open_fraction,
            grid.face_distance, 343.0, 1.21, dx, order)
    g = AcousticAMR(*args)
    g.record_layout = a.amr_layout
    g.set_border_condition(0, 3.0, 0.0, 0.0, n_pml, grid.bounds_min, grid.bounds_max)
    layout = g.record_layout
    g.inject_pressure_nearest([1.0, 1.2, 1.4], 1.0)
    steps = a.steps if a.steps != 4000 else 1000
    warm = min(a.warmup, 20)
    if warm:
        g.step(warm)
    chunk = 250
    t0 = time.perf_counter()
    ms_f = ms_c = ms_t = 0.0
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        rc, t, f, c = g.step_timed(k)
        if rc != 0:
            raise SystemExit(f"amr_step_timed failed rc={rc}")
        ms_t, ms_f, ms_c = ms_t + t, ms_f + f, ms_c + c
        done += k
    wall = time.perf_counter() - t0
    if not np.all(np.isfinite(g.get_pressure())):
        raise SystemExit("non-finite AMR pressure")
    C_, F, E = g.n_cells, g.n_faces, g.n_csr_entries
    # algorithmic HBM bytes per step (each array touched once):
    #   faces kernel: v read 4 + v write 4 per face, p read 4 per cell, plus the
    #                 face record: direct 16 B (order 8: + 64 B of stencil),
    #                 compact 8 B {neg, class} (class tables stay in cache)
    #   cells kernel: row start 4 + p read/write 8 per cell, CSR index 4 per
    #                 entry, v and face weight 4 + 4 per face, plus the cell
    #                 record: direct 16 B, compact a 4 B class id
    rec_f = 8 if layout == "compact" else (16 + (64 if order == 8 else 0))
    rec_c = 4 if layout == "compact" else 16
    b_faces = (8 + rec_f) * F + 4 * C_
    b_cells = (12 + rec_c) * C_ + 4 * E + 8 * F
    fa = b_faces / (ms_f / steps * 1e-3) / 1e9
    ca = b_cells / (ms_c / steps * 1e-3) / 1e9
    dom = "cells" if ms_c >= ms_f else "faces"
    # HBM bytes per k_amr_cells launch counted with rocprofv3 on this workload
    amr_traffic = None
    prof = os.path.join(ROOT, "profiles",
                        f"amr_uniform_box_256cube_order{order}_anechoic8_{layout}.txt")
    if os.path.exists(prof):
        for ln in open(prof):
            if ln.startswith("traffic_bytes_per_launch:"):
                amr_traffic = int(ln.split(":")[1])
    ach = ca if dom == "cells" else fa
    line = {"metric": f"cell-updates/s (AMR order-{order} CSR stepper, 1 GPU)",
            "value": C_ * steps / wall, "unit": "cell-updates/s", "n_gpus": 1, "steps": steps,
            "warmup": warm, "ms_per_step": wall * 1e3 / steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (uniform box grid, one pulse)",
            "config": {"workload": f"amr-uniform-box-{n}x{n}x{n}-order{order}-anechoic{n_pml}",
                       "cells": C_, "faces": F, "csr_entries": E, "record_layout": layout},
            "kernel_ms_per_step": {"faces": ms_f / steps, "cells": ms_c / steps},
            "event_ms_total": ms_t,
            "roofline": {"bound": "hbm", "kernel": f"k_amr_{dom}", "achieved": ach,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                         "traffic": amr_traffic if dom == "cells" else None,
                         "bytes_per_launch": b_cells if dom == "cells" else b_faces,
                         "bytes_per_cell_update": (b_faces + b_cells) / C_,
                         "faces_GBps": fa, "cells_GBps": ca,
                         "step_GBps": (b_faces + b_cells) / ((ms_f + ms_c) / steps * 1e-3) / 1e9}}
    del g
    if not a.no_cpu:
        from oracle.amr_oracle import OracleAMR
        o = OracleAMR(*args)
        o.set_border_condition(0, 3.0, 0.0, 0.0, n_pml, grid.bounds_min, grid.bounds_max)
        o.inject_pressure_nearest([1.0, 1.2, 1.4], 1.0)
        o.step(1)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 12.0:
            o.step(1)
            k += 1
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = {"value": C_ * k / dt, "unit": "cell-updates/s", "cores": cores,
                                "kind": "port",
                                "sample": f"{k} steps of the same grid ({dt:.1f} s, float32 AMR "
                                          f"oracle, OpenMP over faces and cells)"}
    print(json.dumps(line), flush=True)
    return 0


def bench_amr_audio(a):
    """The acoustic half of the co-evolver's per-sample loop on the AMR
    stepper: 96^3 box grid (884,736 cells) with an 8-cell anechoic shell, 4
    mics, `substeps` steps per audio sample (the co-evolver's sub-stepping at
    48 kHz for dx = 5 mm), each sample's polar-mixed mic output recorded on
    the device (amr_record_mic_samples); the oracle running the same loop
    (step, sample_mics, mix) is the CPU baseline."""
    from afdtd.amr import AcousticAMR, AcousticAMRGrid
    n, dx, n_pml = a.amr_n or 96, 0.005, 8
    grid = AcousticAMRGrid.build_uniform_box(x_min=0, x_max=n * dx, y_min=0, y_max=n * dx,
                                             z_min=0, z_max=n * dx, dx=dx).with_pml_shell(n_pml)
    plate_kw = None
    if a.amr_plate:
        # a plate layer across the box's mid height inside the absorbing shell,
        # driven through three bridge nodes between blocks
        act = np.zeros((n, n), np.uint8)
        act[n_pml + 2:n - n_pml - 2, n_pml + 2:n - n_pml - 2] = 1
        grid, plate_kw = grid.with_plate_layer(n, n, n, n // 2, act)
    args = (grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction, grid.cell_types,
            grid.face_cell_neg, grid.face_cell_pos, grid.face_area, grid.face_open_fraction,
            grid.face_distance, 343.0, 1.21, dx, 2)
    sr = 48000.0
    box = n * dx
    g = AcousticAMR(*args)
    substeps = int(np.ceil((1.0 / sr) / g.dt))
    mics = (np.array([[0.11, 0.23, 0.31], [0.35, 0.2, 0.12], [0.24, 0.24, 0.4],
                      [0.3, 0.37, 0.25]]) * (box / 0.48)).astype(np.float32)
    axis = np.tile(np.float32([0, 0, 1]), (4, 1))
    pa, pb, gain = np.float32([1, 0.5, 0, 1]), np.float32([0, 0.5, 1, 0]), np.ones(4, np.float32)
    rho_c = np.float32(np.float32(1.21) * np.float32(343.0))

    def setup(m):
        m.set_border_condition(0, 3.0, 0.0, 0.0, n_pml, grid.bounds_min, grid.bounds_max)
        m.inject_pressure_nearest(np.array([0.2, 0.25, 0.22]) * (box / 0.48), 1.0)
        m.setup_mic_samplers(mics, axis, True)
        if plate_kw is not None:
            m.setup_plate(**plate_kw)
            fl = plate_kw["plate_active_flat_idx"]
            m.set_bridge_plate_sources(fl[[fl.size // 3, fl.size // 2, 2 * fl.size // 3]],
                                       [0.5, 0.3, 0.2])

    drive = np.float32([1e-3, -2e-3, 1.5e-3])

    setup(g)
    # blocks of `block` samples, as a co-evolver would hand them over; the
    # first (warm-up) block captures the launch sequence, the timed blocks
    # replay it
    block = 64
    samples = (a.steps if a.steps != 4000 else 2048) // block * block
    g.record_mic_samples(block, substeps, pa, pb, axis, gain, rho_c)
    t0 = time.perf_counter()
    for _ in range(samples // block):
        if plate_kw is not None:
            g.inject_bridge_drive(drive, 1.0)
        out, rc, done = g.record_mic_samples(block, substeps, pa, pb, axis, gain, rho_c)
        if rc != 0 or not np.all(np.isfinite(out)):
            raise SystemExit(f"amr_record_mic_samples failed rc={rc}")
    wall = time.perf_counter() - t0
    line = {"metric": "audio samples/s (AMR order-2 stepper, sub-stepped, mic polar mix, 1 GPU)",
            "value": samples / wall, "unit": "samples/s", "n_gpus": 1, "steps": samples,
            "warmup": block, "ms_per_step": wall * 1e3 / samples,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (uniform box grid, one pulse)",
            "config": {"workload": f"amr-audio-box-{n}x{n}x{n}-dx5mm-48kHz"
                                   + ("-plate" if plate_kw is not None else ""),
                       "cells": g.n_cells, "faces": g.n_faces, "substeps": substeps, "mics": 4,
                       "plate_nodes": 0 if plate_kw is None else
                       int(plate_kw["plate_active_flat_idx"].size),
                       "samples_per_block": block,
                       "record_layout": g.record_layout},
            "realtime_factor_48k": samples / wall / sr,
            "cell_updates_per_s": samples * substeps * g.n_cells / wall}
    del g
    if not a.no_cpu:
        from oracle.amr_oracle import OracleAMR
        o = OracleAMR(*args)
        setup(o)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < 10.0:
            if plate_kw is not None and k % block == 0:
                o.inject_bridge_drive(drive, 1.0)
            o.step(substeps)
            sp, vx, vy, vz = o.sample_mics()
            _ = (sp * pa + (pb * rho_c) * (vx * axis[:, 0] + vy * axis[:, 1] + vz * axis[:, 2])) * gain
            k += 1
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = {"value": k / dt, "unit": "samples/s", "cores": cores,
                                "kind": "port",
                                "sample": f"{k} samples of the same loop ({dt:.1f} s, float32 AMR "
                                          f"oracle, OpenMP over faces and cells)"}
    print(json.dumps(line), flush=True)
    return 0


def bench_body(a):
    """A voxelised two-bout body at the bridge factory's spacing (dx = 4 mm,
    the spacing at which the plate's explicit update with the reference's
    default plate constants is stable: its stiffness-proportional damping
    term bK/dt * D dt^2/(m dx^4) grows as dx shrinks), scaled 4x so that the
    grid is HBM-sized (256-vertex outline 1.44 m x 2 m, rib height 0.4 m,
    24 cells of padding, 10-cell PML): the soundhole cut, cut-cell face
    fractions, an active plate and three Gaussian bridge sources
    (FDTDCoEvolver.from_outline), i.e. the geometry the bridge feeds the
    solver, which runs the two-step sweep with cell geometry (--kernel fused:
    the single-step sweep with cell geometry; --kernel generic: the
    two-kernel path).  Timed: the steps after one bridge drive
    (the stencil), and the bridge's per-sample loop at 44.1 kHz (drive,
    steps_per_sample steps).  Algorithmic bytes per cell-update: the four
    fields read and written (32 B) and one geometry code byte (walls, closed,
    fractional and plate faces) = 33 B; the fractions of cut faces and the
    plate (O(surface)) are left out.  The reference's own layout would read
    the cell type and three fraction arrays, 45 B."""
    from afdtd.bridge import FDTDCoEvolver
    from afdtd.geometry import two_bout_outline
    scale, dx, pad = 4.0, 0.004, 24
    poly = (two_bout_outline(256) * scale).astype(np.float32)
    body_h = 0.1 * scale
    bridges = [(-0.08, -0.52), (0.0, -0.52), (0.08, -0.52)]
    hole = (0.0, 0.24, 0.18)
    ev = FDTDCoEvolver.from_outline(poly, body_h, bridges, soundhole=hole, dx=dx,
                                    pad_cells=pad, n_pml=10, kernel=a.kernel,
                                    tile_y=a.tile_y, seg_x=a.seg_x)
    info, f = ev.info, ev.fdtd
    cells = info["Nx"] * info["Ny"] * info["Nz"]
    steps = a.steps if a.steps != 4000 else 400
    nb = len(ev.bridge_cells)
    f.inject_bridge_drive(np.full(nb, 1e-3, np.float32), 1.0)
    f.step(a.warmup)
    t0 = time.perf_counter()
    ms_events, ms_stencil = f.step_timed(steps)
    wall = time.perf_counter() - t0
    if not np.all(np.isfinite(f.get_pressure_field())):
        raise SystemExit("non-finite pressure")
    bpc = 33
    achieved = cells * bpc / (ms_stencil * 1e-3) / 1e9
    traffic = None
    pair = f.steps_per_sweep == 2
    if f.kernel_family == "fused" and a.tile_y in (0, 8) and a.seg_x == 0:
        prof = os.path.join(ROOT, "profiles",
                            ("fused4_geo" if pair else "fused_geo") +
                            "_body_x4_395x549x148_order4_pml10.txt")
        if os.path.exists(prof):
            for ln in open(prof):
                if ln.startswith("traffic_bytes_per_launch:"):
                    traffic = int(ln.split(":")[1])
    # the bridge's per-sample loop: one drive + steps_per_sample steps per sample
    ns = 64
    drives = (1e-4 * np.random.default_rng(3).standard_normal((ns, nb))).astype(np.float32)
    t0 = time.perf_counter()
    ev.drive_bridge_block(drives)
    wall_s = time.perf_counter() - t0
    line = {"metric": "cell-updates/s (3D FDTD, voxelised body, 1 GPU)",
            "value": cells * steps / wall, "unit": "cell-updates/s", "n_gpus": 1,
            "steps": steps, "warmup": a.warmup, "ms_per_step": wall * 1e3 / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (two-bout outline, soundhole, bridge drive)",
            "config": {"workload": f"body-two-bout-x4-{info['Nx']}x{info['Ny']}x{info['Nz']}-dx4mm",
                       "cells": cells, "kernel": f.kernel_family + ("-two-step" if pair else ""),
                       "bridge_cells": nb, "steps_per_sample": ev.steps_per_sample},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "bytes_per_cell_update": bpc,
                         "cell_updates_per_launch": cells * f.steps_per_sweep,
                         "stencil_ms_per_step": ms_stencil, "event_ms_total": ms_events,
                         "traffic_GBps": (traffic / (f.steps_per_sweep * ms_stencil * 1e-3) / 1e9)
                         if traffic else None},
            "bridge_loop": {"samples": ns, "samples_per_s": ns / wall_s,
                            "realtime_factor_44k1": ns / wall_s / 44100.0}}
    del ev, f
    if not a.no_cpu:
        from afdtd.geometry import apply_soundhole_cutout
        from oracle import geometry as og
        from oracle.oracle import OracleFDTD
        ct, pa, g = og.voxelise(poly, body_h, dx, pad)
        apply_soundhole_cutout(ct, pa, g, hole)
        o = OracleFDTD(g["Nx"], g["Ny"], g["Nz"], dx, c=343.0, rho_air=1.21,
                       cell_type=ct.ravel(), plate_iz=g["plate_iz"], plate_active=pa.ravel(),
                       n_pml=10, precision="f32")
        o.set_face_fractions(*og.face_fractions(ct, pa, poly, g))
        o.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 15.0:
            o.step(2)
            n += 2
        el = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        line["cpu_baseline"] = {"value": cells * n / el, "unit": "cell-updates/s",
                                "cores": threads, "kind": "port",
                                "sample": f"{n} steps of the same body grid ({el:.1f} s, "
                                          f"float32 oracle, OpenMP over x slabs)"}
    print(json.dumps(line), flush=True)
    return 0


def coevolver_workload(dx=0.004, pad=10, n_pml=10, n_strings=6, n_segs=240):
    """The instrument the reference's scene factory builds
    (acoustic_fdtd_bridge.py:1105-1302) on a uniform body: a two-bout
    outline (rib 0.1 m) voxelised at dx with pad cells and a PML, three
    bridge saddles, six strings of n_segs segments with the shared neck
    mode, one bridge single-coil, a cardioid 75 mm above the top plus a ring
    of eight (44.1 kHz, modal stride 16)."""
    from afdtd.geometry import two_bout_outline, voxelise_guitar_body
    from afdtd.instrument import body_dict, guitar_defs
    poly = two_bout_outline(96)
    body_h = 0.1
    bridges = [(-0.02, -0.13), (0.0, -0.13), (0.02, -0.13)]
    ct, pa, info = voxelise_guitar_body(poly, body_h, dx=dx, pad_cells=pad, n_pml=n_pml)
    S, P, M = guitar_defs(body_h, bridges, n_strings=n_strings, n_segs=n_segs)
    B = body_dict(ct, pa, info, bridges, body_h, n_pml=n_pml)
    return S, P, M, B


def bench_coevolver(a):
    """The co-evolver's per-sample loop on the device (coevolver_step): per
    audio sample, K body sub-steps, each one string / bridge / neck kernel
    and the body FDTD step (single-step sweep with cell geometry), then the
    pickups and mics.  Metric: audio samples per second (1 GPU); the body
    sweep's roofline is reported at 33 B per cell-update (four fields read
    and written, one geometry code byte).  CPU baseline: the co-evolver
    oracle (float32, OpenMP body FDTD) on the same instrument."""
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, B = coevolver_workload()
    co = AcousticCoEvolver(S, P, M, B, 44100.0, modal_stride=16)
    cells = B["Nx"] * B["Ny"] * B["Nz"]
    K = co.substeps
    samples = a.steps if a.steps != 4000 else 2048
    for si in range(len(S)):
        co.schedule_pluck(8 * si, si, 0.2 + 0.1 * si, 0.002, 0)
    co.step(max(a.warmup, 1))
    t0 = time.perf_counter()
    ms_dev = co.step_timed(samples)
    wall = time.perf_counter() - t0
    mic = co.get_mic_output(0, min(samples, 4096))
    if not np.all(np.isfinite(mic)) or not np.any(mic):
        raise SystemExit("co-evolver produced no finite mic output")
    # the body sweep alone, on the co-evolver's own FDTD handle
    from afdtd import lib
    import ctypes as C
    h = co.fdtd_handle()
    mt, msw = C.c_float(), C.c_float()
    nst = 200
    rc = lib().fdtd_step_timed(h, nst, C.byref(mt), C.byref(msw))
    if rc:
        raise SystemExit(f"fdtd_step_timed: {rc}")
    bpc = 33
    achieved = cells * bpc / (msw.value * 1e-3) / 1e9
    line = {"metric": "audio samples/s (instrument co-evolver, uniform body, 1 GPU)",
            "value": samples / wall, "unit": "samples/s", "n_gpus": 1, "steps": samples,
            "warmup": a.warmup, "ms_per_step": wall * 1e3 / samples, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (two-bout body, six plucked strings)",
            "config": {"workload": f"coevolver-6str-240seg-body-{B['Nx']}x{B['Ny']}x{B['Nz']}-dx4mm",
                       "cells": cells, "substeps_per_sample": K, "strings": len(S),
                       "string_nodes": sum(np.asarray(d["path_xyz"]).size // 3 for d in S),
                       "pickups": len(P), "mics": len(M), "sample_rate": 44100},
            "realtime_factor_44k1": samples / wall / 44100.0,
            "device_ms_per_sample": ms_dev / samples,
            "device_us_per_substep": ms_dev * 1e3 / (samples * K),
            "roofline": {"bound": "hbm", "kernel": "body sweep (fused, cell geometry)",
                         "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                         "bytes_per_cell_update": bpc, "cell_updates_per_launch": cells,
                         "stencil_ms_per_step": msw.value,
                         "step_share_of_substep": msw.value / (ms_dev / (samples * K))}}
    del co
    if not a.no_cpu:
        from oracle import coevo
        o = coevo.OracleCoEvolver(S, P, M, B, 44100.0, modal_stride=16)
        for si in range(len(S)):
            o.schedule_pluck(8 * si, si, 0.2 + 0.1 * si, 0.002, 0)
        o.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 15.0 and n < samples:
            o.step(8)
            n += 8
        el = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        line["cpu_baseline"] = {"value": n / el, "unit": "samples/s", "cores": threads,
                                "kind": "port",
                                "sample": f"{n} samples of the same instrument ({el:.1f} s, "
                                          f"float32 co-evolver oracle, OpenMP body FDTD)"}
    print(json.dumps(line), flush=True)
    return 0


def coevolver_amr_workload(nx=96, ny=128, nz=56, kp=20, dx=0.004, n_pml=8):
    """The same instrument on an AMR body (coevolver_create_amr): a uniform
    AMR box of nx x ny x nz cells at dx (688 K cells, about the uniform
    body's 718 K) with an 8-cell anechoic shell, the plate in layer kp
    (nodes 5 dx apart), bridge and neck plate patches; six 240-segment
    strings, one bridge single-coil, a cardioid plus a ring of eight."""
    from afdtd.instrument import amr_box_desc, guitar_defs
    org = (-0.192, -0.300, 0.0)
    body_h = org[2] + kp * dx
    bridges = [(-0.02, -0.13), (0.0, -0.13), (0.02, -0.13)]
    D = amr_box_desc(nx, ny, nz, kp, dx, org, bridges, neck_positions=[(0.0, 0.2)],
                     n_pml=n_pml, border_mode=0)
    S, P, M = guitar_defs(body_h, bridges)
    return S, P, M, D


def bench_coevolver_amr(a):
    """The co-evolver's per-sample loop on an AMR body (coevolver_step on a
    coevolver_create_amr handle): per audio sample, K sub-steps of one
    string / bridge / neck kernel and the AMR step (faces, plate faces,
    cells, plate), then pickups and the mics' stencil sums.  Metric: audio
    samples per second (1 GPU).  Roofline: the AMR step's dominant kernel on
    a plain AMR handle of the same grid, with the AMR stepper's byte
    accounting (order 2, the handle's record layout).  CPU baseline: the co-evolver
    oracle on the AMR oracle, same instrument."""
    from afdtd.coevolver import AcousticCoEvolver
    S, P, M, D = coevolver_amr_workload()
    co = AcousticCoEvolver.from_amr_desc(D, S, P, M, 44100.0, modal_stride=16)
    K = co.substeps
    samples = a.steps if a.steps != 4000 else 2048
    for si in range(len(S)):
        co.schedule_pluck(8 * si, si, 0.2 + 0.1 * si, 0.002, 0)
    co.step(max(a.warmup, 1))
    t0 = time.perf_counter()
    ms_dev = co.step_timed(samples)
    wall = time.perf_counter() - t0
    mic = co.get_mic_output(0, min(samples, 4096))
    if not np.all(np.isfinite(mic)) or not np.any(mic):
        raise SystemExit("AMR co-evolver produced no finite mic output")
    del co
    # the AMR step alone on a plain handle of the same grid
    from afdtd.amr import AcousticAMR
    g = AcousticAMR(D["cell_centers"], D["cell_volumes"], D["open_volume_frac"],
                    D["cell_types"], D["face_cell_neg"], D["face_cell_pos"], D["face_area"],
                    D["face_open_frac"], D["face_distance"], D["c"], D["rho_air"], D["min_dx"],
                    D["gradient_order"])
    g.set_border_condition(0, 3.0, 0.0, 0.0, D["n_pml"], D["bounds_min"], D["bounds_max"])
    layout = g.record_layout                 # the layout the co-evolver's body picked too
    nst = 200
    rc, ms_t, ms_f, ms_c = g.step_timed(nst)
    if rc:
        raise SystemExit(f"amr_step_timed failed rc={rc}")
    C_, F, E = g.n_cells, D["n_faces"], g.n_csr_entries
    rec_f, rec_c = (8, 4) if layout == "compact" else (16, 16)
    b_faces = (8 + rec_f) * F + 4 * C_
    b_cells = (12 + rec_c) * C_ + 4 * E + 8 * F
    dom = "cells" if ms_c >= ms_f else "faces"
    ach = (b_cells / (ms_c / nst * 1e-3) if dom == "cells" else
           b_faces / (ms_f / nst * 1e-3)) / 1e9
    del g
    line = {"metric": "audio samples/s (instrument co-evolver, AMR body, 1 GPU)",
            "value": samples / wall, "unit": "samples/s", "n_gpus": 1, "steps": samples,
            "warmup": a.warmup, "ms_per_step": wall * 1e3 / samples, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
            "data": "synthetic (AMR box body with plate, six plucked strings)",
            "config": {"workload": f"coevolver-amr-6str-240seg-box-{D['plate_Nx']}x"
                                   f"{D['plate_Ny']}x{D['n_cells'] // (D['plate_Nx'] * D['plate_Ny'])}-dx4mm",
                       "cells": D["n_cells"], "faces": D["n_faces"], "substeps_per_sample": K,
                       "strings": len(S),
                       "string_nodes": sum(np.asarray(d["path_xyz"]).size // 3 for d in S),
                       "pickups": len(P), "mics": len(M), "sample_rate": 44100},
            "realtime_factor_44k1": samples / wall / 44100.0,
            "device_ms_per_sample": ms_dev / samples,
            "device_us_per_substep": ms_dev * 1e3 / (samples * K),
            "roofline": {"bound": "hbm", "kernel": f"k_amr_{dom} (plain handle, same grid, "
                                                   f"{layout} records)",
                         "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS, "traffic": None,
                         "bytes_per_launch": b_cells if dom == "cells" else b_faces,
                         "amr_step_ms": (ms_f + ms_c) / nst,
                         "step_share_of_substep": ((ms_f + ms_c) / nst)
                                                  / (ms_dev / (samples * K))}}
    if not a.no_cpu:
        from oracle import coevo
        o = coevo.OracleCoEvolver.from_amr(D, S, P, M, 44100.0, modal_stride=16)
        for si in range(len(S)):
            o.schedule_pluck(8 * si, si, 0.2 + 0.1 * si, 0.002, 0)
        o.step(1)
        t0 = time.perf_counter()
        n = 0
        while time.perf_counter() - t0 < 15.0 and n < samples:
            o.step(8)
            n += 8
        el = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        line["cpu_baseline"] = {"value": n / el, "unit": "samples/s", "cores": threads,
                                "kind": "port",
                                "sample": f"{n} samples of the same instrument ({el:.1f} s, "
                                          f"float32 co-evolver oracle on the AMR oracle, OpenMP)"}
    print(json.dumps(line), flush=True)
    return 0


def _timed_room(w, storage="f32", kernel="auto", steps=40, warmup=4):
    """Steps of workload w on one handle: (cell-updates/s wall clock, average
    stencil ms per launch, kernel family)."""
    s = make_solver(w, storage, kernel)
    total = warmup + steps
    s.set_point_sources(w.src, w.signals(total))
    s.set_receivers(w.rec)
    s.record_receivers(total)
    s.step(warmup)
    t0 = time.perf_counter()
    _, ms_stencil = s.step_timed(steps)
    wall = time.perf_counter() - t0
    fam = s.kernel_family
    if not np.all(np.isfinite(s.get_receiver_traces())):
        raise RuntimeError("non-finite receiver traces")
    del s
    per_launch = 2 if fam == "fused2" else 1
    return w.cells * steps / wall, ms_stencil * per_launch, fam


def measure_forms(steps=40):
    """The three forms beside the plain drop-in and the reference's own
    scheme, each at a reduced step count in the same run as the headline
    line: batched rooms against per-room launches, 16-bit storage against
    FP32 on the precision room, the adjoint's reverse rate, the order-4
    PML room, and the voxelised body on the two-step and the single-step
    sweeps with geometry.  A form that fails reports its error instead of a
    number."""
    from afdtd import RoomBatch, AcousticFDTD, workloads as W
    forms = {}

    def guard(name, fn):
        try:
            t0 = time.perf_counter()
            forms[name] = fn()
            forms[name]["bench_seconds"] = round(time.perf_counter() - t0, 2)
        except Exception as e:                       # reported, not fatal
            forms[name] = {"error": f"{type(e).__name__}: {e}"}
        print(f"[bench] form {name}: {forms[name]}", file=sys.stderr, flush=True)

    def reference_scheme():
        w = W.config2_order4_pml(steps=steps + 4)
        v, ms, fam = _timed_room(w, steps=steps)
        achieved = 2 * w.cells * 32 / (ms * 1e-3) / 1e9 if fam == "fused2" else \
            w.cells * 32 / (ms * 1e-3) / 1e9
        return {"workload": w.name, "value": v, "unit": "cell-updates/s", "kernel": fam,
                "steps": steps, "roofline_frac": achieved / HBM_PEAK_GBS}

    def batched():
        c = W.config3(steps=steps)
        dims = c["dims"]
        cells = int(np.prod(dims, axis=1).sum())
        b = RoomBatch(dims, 0.01, c["courant"], c["refl"], c=343.0, rho_air=1.2, order=2)
        tuned = b.autotune(trial_steps=6)
        b.set_sources(c["src"], c["signals"])
        b.set_receivers(c["rec"])
        b.record(steps + 4)
        b.step(4)
        t0 = time.perf_counter()
        b.step(steps, graph=True)
        wall_b = time.perf_counter() - t0
        del b
        t_rooms = 0.0
        for r, (Nx, Ny, Nz) in enumerate(dims):
            Nx, Ny, Nz = int(Nx), int(Ny), int(Nz)
            h = AcousticFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.2,
                             cell_type=np.zeros(Nx * Ny * Nz, np.uint8), plate_iz=Nz // 2,
                             plate_active=np.zeros(Nx * Ny, np.uint8), n_pml=0, order=2,
                             courant=float(c["courant"][r]), boundary="impedance",
                             refl=c["refl"][r])
            h.set_point_sources([c["src"][r]], c["signals"][r:r + 1])
            h.set_receivers(c["rec"][r])
            h.step(4)
            t0 = time.perf_counter()
            h.step(steps)
            t_rooms += time.perf_counter() - t0
            del h
        return {"workload": "config3-256-rooms-48to160-order2", "rooms": int(len(dims)),
                "cells": cells, "steps": steps, "value": cells * steps / wall_b,
                "unit": "cell-updates/s", "per_room_value": cells * steps / t_rooms,
                "speedup_vs_per_room": t_rooms / wall_b, "autotuned": tuned}

    def precision():
        w = W.config4(steps=steps + 4)
        out = {"workload": w.name, "steps": steps, "unit": "cell-updates/s"}
        for st, kern in (("f32", "auto"), ("f32", "fused"), ("bf16", "auto"), ("bf16", "fused2"),
                         ("f16", "auto"), ("f16", "fused2"), ("bfp16", "auto"), ("f16s", "auto")):
            v, _, fam = _timed_room(w, storage=st, kernel=kern, steps=steps)
            out[f"{st}_{fam}"] = v
        base = out["f32_fused2"]          # the FP32 path a handle gets by default
        out["bf16_vs_f32"] = max(out["bf16_fused"], out["bf16_fused2"]) / base
        out["bfp16_vs_f32"] = out["bfp16_fused"] / base
        out["f16s_vs_f32"] = out["f16s_fused"] / base
        return out

    def adjoint():
        import torch
        from afdtd.adjoint import AdjointSolver, simulate_traces
        c = W.config5(steps=10 * steps)
        shape, n = c["shape"], c["steps"]
        N = int(np.prod(shape))
        sol = AdjointSolver(*shape, c["dtdx"], dtype="float32", checkpoint_every=100)
        sol.set_sources(c["src"], c["signals"])
        sol.set_receivers(c["rec"])
        dev = lambda x: torch.tensor(np.ascontiguousarray(x).ravel(), dtype=torch.float32,
                                     device="cuda")
        with torch.no_grad():
            target = sol.forward(dev(c["m"] + c["pert"][0]), dev(c["l"] + c["pert"][1]),
                                 dev(c["eta"]), n).clone()
        mg, lg, eg = (dev(x).requires_grad_() for x in (c["m"], c["l"], c["eta"]))
        ((simulate_traces(sol, mg, lg, eg, n) - target) ** 2).sum().backward()
        mg.grad = lg.grad = eg.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = ((simulate_traces(sol, mg, lg, eg, n) - target) ** 2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return {"workload": "config5-128cube-ckpt100", "steps": n, "cells": N,
                "value": N * n / (t2 - t1), "unit": "reverse cell-updates/s",
                "forward_value": N * n / (t1 - t0), "roofline_frac":
                N * n / (t2 - t1) * 104 / 1e9 / HBM_PEAK_GBS}

    def voxelised_body():
        from afdtd.bridge import FDTDCoEvolver
        from afdtd.geometry import two_bout_outline
        poly = (two_bout_outline(256) * 4.0).astype(np.float32)
        out = {"workload": "body-two-bout-x4-395x549x148-dx4mm", "steps": steps,
               "unit": "cell-updates/s"}
        for kern in ("auto", "fused"):
            ev = FDTDCoEvolver.from_outline(poly, 0.4, [(-0.08, -0.52), (0.0, -0.52), (0.08, -0.52)],
                                            soundhole=(0.0, 0.24, 0.18), dx=0.004, pad_cells=24,
                                            n_pml=10, kernel=kern)
            f = ev.fdtd
            cells = ev.info["Nx"] * ev.info["Ny"] * ev.info["Nz"]
            f.inject_bridge_drive(np.full(len(ev.bridge_cells), 1e-3, np.float32), 1.0)
            f.step(4)
            t0 = time.perf_counter()
            _, ms = f.step_timed(steps)
            wall = time.perf_counter() - t0
            name = "two_step" if f.steps_per_sweep == 2 else "single_step"
            out[name] = cells * steps / wall
            out[name + "_roofline_frac"] = cells * 33 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS
            del ev, f
        out["two_step_vs_single_step"] = out["two_step"] / out["single_step"]
        return out

    guard("reference_scheme_order4", reference_scheme)
    guard("batched", batched)
    guard("precision", precision)
    guard("adjoint", adjoint)
    guard("voxelised_body", voxelised_body)
    return forms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="room")
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--kernel", default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-y", type=int, default=0)
    ap.add_argument("--seg-x", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--amr-order", type=int, default=2, choices=(2, 8),
                    help="amr: gradient order of the AMR stepper")
    ap.add_argument("--amr-n", type=int, default=0,
                    help="amr-audio: cells per box edge (default 96)")
    ap.add_argument("--amr-plate", action="store_true",
                    help="amr-audio: a plate layer across the box, driven between blocks")
    ap.add_argument("--amr-layout", default="auto", choices=("auto", "direct", "compact"),
                    help="amr: record layout of the step kernels")
    ap.add_argument("--no-fd", action="store_true",
                    help="adjoint: float32 timing only (no float64 pass, no finite differences)")
    ap.add_argument("--no-forms", action="store_true",
                    help="room: skip the forms measured beside the headline line")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one MI355X: the multi-GPU (slab-sharded) path is a "
              "design only in this build; run with --gpus 1", file=sys.stderr)
        return 2
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.config == "batch":
        if a.steps == 4000:
            a.steps = 1000
        return bench_batch(a)
    if a.config == "adjoint":
        if a.steps == 4000:
            a.steps = 2000
        return bench_adjoint(a)
    if a.config == "geometry":
        return bench_geometry(a)
    if a.config == "body":
        return bench_body(a)
    if a.config == "coevolver":
        return bench_coevolver(a)
    if a.config == "coevolver-amr":
        return bench_coevolver_amr(a)
    if a.config == "amr":
        return bench_amr(a)
    if a.config == "amr-audio":
        return bench_amr_audio(a)
    w = make_workload(a.config, a.warmup + a.steps)
    s = make_solver(w, a.storage, a.kernel, a.tile_y, a.seg_x)
    total = a.warmup + a.steps
    s.set_point_sources(w.src, w.signals(total))
    s.set_receivers(w.rec)
    s.record_receivers(total)
    if a.warmup:
        s.step(a.warmup)
    t0 = time.perf_counter()
    ms_events, ms_stencil = s.step_timed(a.steps)
    wall = time.perf_counter() - t0
    traces = s.get_receiver_traces()
    if not np.all(np.isfinite(traces)):
        raise SystemExit("non-finite receiver traces")
    value = w.cells * a.steps / wall
    bpc = w.bytes_per_cell_update_for(a.storage)
    # Algorithmic bytes: B per cell-update (every field read and written once
    # per step) times the cell-updates one launch performs.  The two-step
    # sweep ("fused2") performs two steps per launch but moves each field
    # through HBM only once, so its minimum HBM bytes per launch are half its
    # algorithmic bytes; both are reported.  ms_stencil is the average
    # stencil time per step (HIP events on the handle's stream).
    steps_per_launch = 2 if s.kernel_family == "fused2" else 1
    ms_launch = ms_stencil * steps_per_launch
    alg_bytes_launch = steps_per_launch * w.cells * bpc
    achieved = alg_bytes_launch / (ms_launch * 1e-3) / 1e9
    once_bytes_launch = w.cells * bpc
    once_rate = once_bytes_launch / (ms_launch * 1e-3) / 1e9
    traffic = measured_traffic(w.name, s.kernel_family, a.storage)
    line = {
        "metric": "cell-updates/s (3D FDTD, 1 GPU)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "fp32", "bf16": "bf16-storage/fp32", "f16": "fp16-storage/fp32",
                  "bfp16": "int16-block-float-storage/fp32",
                  "f16s": "fp16-row-scaled-storage/fp32"}[a.storage],
        "data": "synthetic (seeded media, sources and receivers)",
        "config": {"workload": w.name, "cells": w.cells, "stencil_order": w.order,
                   "courant": w.courant, "boundary": w.boundary, "heterogeneous": w.hetero,
                   "sources": len(w.src), "receivers": int(len(w.rec)),
                   "kernel": s.kernel_family, "storage": a.storage,
                   "tile_y": a.tile_y or (24 if (s.kernel_family == "fused2" and w.order == 4)
                                          else 16),
                   "seg_x": a.seg_x or (64 if s.kernel_family == "fused2" else 32),
                   "steps_per_launch": steps_per_launch},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "bytes_per_cell_update": bpc,
                     "cell_updates_per_launch": steps_per_launch * w.cells,
                     "bytes_per_launch": alg_bytes_launch,
                     "stencil_ms_per_launch": ms_launch, "event_ms_total": ms_events,
                     # each field read and written once per launch: the least
                     # HBM traffic a launch can have, and its rate
                     "once_bytes_per_launch": once_bytes_launch,
                     "once_GBps": once_rate,
                     "once_frac": once_rate / HBM_PEAK_GBS,
                     # measured HBM bytes (counters) over the launch time
                     "traffic_GBps": (traffic / (ms_launch * 1e-3) / 1e9) if traffic else None},
    }
    del s
    # rate of a plain streaming kernel with one step's read/write mix: above 1
    # the path beats any sweep that passes over HBM once per step
    ceiling = stream_ceiling(w, a.storage)
    line["roofline"]["stream_ceiling_GBps"] = ceiling
    line["roofline"]["frac_of_stream_ceiling"] = achieved / ceiling
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w)
    if a.config == "room" and not a.no_forms:
        line["forms"] = measure_forms()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
