"""Writes config1_golden.npz: receiver traces of the 64^3 workload after
2000 steps, made by the float32 oracle, in the configuration as written
(order 2, Courant 0.5, impedance walls R = 0.9) and in the reference's own
scheme (order 4, reference dt, rigid walls), with the SHA-256 of each final
pressure field.

The reference ships no golden vectors for this path and cannot be built
here, so this is a regression fixture: the oracle it comes from is pinned
by the closed-form checks in tests/test_oracle_pins.py, and the fixture
freezes the oracle's and the HIP path's output for the full 2000 steps.

    python tests/golden/make_config1_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "spectral-analyzer_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

OUT = os.path.join(HERE, "config1_golden.npz")


def run_case(make_solver, w, steps):
    """Sources, receivers and `steps` steps of workload w on a solver made by
    make_solver(**kwargs); returns (traces float32, sha256 of final P)."""
    from afdtd.signals import differentiated_gaussian
    N = w.Nx
    s = make_solver(N)
    sig = differentiated_gaussian(steps, 8.0, 48.0, 1.0)[None, :]
    s.set_point_sources(w.src, sig)
    s.set_receivers(w.rec)
    s.record_receivers(steps)
    assert s.step(steps) == 0
    tr = np.asarray(s.get_receiver_traces(), np.float32)
    p = np.ascontiguousarray(np.asarray(s.get_pressure_field(), np.float32).ravel())
    return tr, hashlib.sha256(p.tobytes()).hexdigest()


def oracle_solver(w):
    from oracle import oracle as orc
    orc.build()

    def make(N):
        kw = dict(n_pml=0, order=w.order, courant=w.courant or 0.0,
                  boundary=1 if w.boundary == "impedance" else 0, refl=w.refl)
        return orc.OracleFDTD(N, N, N, w.dx, c=343.0, rho_air=1.2,
                              cell_type=np.zeros(N ** 3, np.uint8), plate_iz=N // 2,
                              plate_active=np.zeros(N * N, np.uint8), precision="f32", **kw)
    return make


def main():
    from afdtd import workloads as W
    out = {}
    for tag, w in (("as_written", W.config1()), ("parity", W.config1(parity=True))):
        tr, sha = run_case(oracle_solver(w), w, 2000)
        out[f"traces_{tag}"] = tr
        out[f"p_sha256_{tag}"] = np.array(sha)
        print(tag, tr.shape, float(np.abs(tr).max()), sha)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
