"""The HIP path reproduces the committed 64^3 regression fixture bit for
bit: receiver traces of 2000 steps and the SHA-256 of the final pressure
field, with the default kernel choice (the two-step sweep, for the order-2
configuration as written and for the reference's order-4 scheme) and with
the single-step sweep forced."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


@pytest.mark.parametrize("tag,kernel,family", [("as_written", "auto", "fused2"),
                                               ("parity", "auto", "fused2"),
                                               ("as_written", "fused", "fused"),
                                               ("parity", "fused", "fused")])
def test_hip_reproduces_config1_golden(tag, kernel, family):
    import make_config1_golden as mk
    from afdtd import AcousticFDTD, workloads as W
    gold = np.load(os.path.join(HERE, "golden", "config1_golden.npz"))
    w = W.config1(parity=(tag == "parity"))
    print(f"golden config1 {tag}", flush=True)
    made = []

    def make(N):
        s = AcousticFDTD(N, N, N, w.dx, c=343.0, rho_air=1.2, cell_type=np.zeros(N ** 3, np.uint8),
                         plate_iz=N // 2, plate_active=np.zeros(N * N, np.uint8), n_pml=0,
                         order=w.order, courant=w.courant, boundary=w.boundary, refl=w.refl,
                         kernel=kernel)
        made.append(s)
        return s
    tr, sha = mk.run_case(make, w, 2000)
    assert made[0].kernel_family == family
    assert np.array_equal(tr.view(np.uint32), gold[f"traces_{tag}"].view(np.uint32))
    assert sha == str(gold[f"p_sha256_{tag}"])
