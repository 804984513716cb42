"""Benchmark: cell-updates/s and HBM-roofline fraction of the FDTD step on one
MI355X, with the CPU oracle timed beside it on the host cores.

A "step" is one full leapfrog step (velocity + pressure half steps, sources,
receiver gather) over the whole grid.  Default workload: the single large
room, 512x384x256 cells, order-2 stencil at Courant 0.55, heterogeneous
density and bulk modulus, locally reacting walls with per-side reflection
coefficients, 4 sources and 64 receivers (synthetic, seeded).

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
    key = {"config2-room-512x384x256-hetero-order2": "fused_room_512x384x256_hetero_order2",
           "room-512x384x256-order4-pml10": "fused_room_512x384x256_order4_pml10"}.get(workload)
    if key is None or kernel != "fused" or storage != "f32":
        return None
    for path in glob.glob(os.path.join(ROOT, "profiles", key + ".txt")):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    return None


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
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return 0
    if a.gpus != 1:
        print("bench.py measures one MI355X: the multi-GPU (slab-sharded) path is a "
              "design only in this build; run with --gpus 1", file=sys.stderr)
        return 2
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
    bpc = w.bytes_per_cell_update if a.storage == "f32" else w.bytes_per_cell_update // 2
    achieved = w.cells * bpc / (ms_stencil * 1e-3) / 1e9
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
        "dtype": {"f32": "fp32", "bf16": "bf16-storage/fp32", "f16": "fp16-storage/fp32"}[a.storage],
        "data": "synthetic (seeded media, sources and receivers)",
        "config": {"workload": w.name, "cells": w.cells, "stencil_order": w.order,
                   "courant": w.courant, "boundary": w.boundary, "heterogeneous": w.hetero,
                   "sources": len(w.src), "receivers": int(len(w.rec)),
                   "kernel": s.kernel_family, "storage": a.storage,
                   "tile_y": a.tile_y or 16, "seg_x": a.seg_x or 32},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(w.name, s.kernel_family, a.storage),
                     "bytes_per_cell_update": bpc,
                     "stencil_ms_per_launch": ms_stencil, "event_ms_total": ms_events},
    }
    del s
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
