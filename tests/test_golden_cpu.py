"""The float32 oracle reproduces the committed 64^3 regression fixture
(tests/golden/config1_golden.npz, written by make_config1_golden.py): receiver
traces bit for bit and the final pressure field by SHA-256, in both the
configuration as written and the reference's own scheme."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))


@pytest.mark.parametrize("tag", ["as_written", "parity"])
def test_oracle_reproduces_config1_golden(tag):
    import make_config1_golden as mk
    from afdtd import workloads as W
    gold = np.load(os.path.join(HERE, "golden", "config1_golden.npz"))
    w = W.config1(parity=(tag == "parity"))
    tr, sha = mk.run_case(mk.oracle_solver(w), w, 2000)
    assert np.array_equal(tr.view(np.uint32), gold[f"traces_{tag}"].view(np.uint32))
    assert sha == str(gold[f"p_sha256_{tag}"])
