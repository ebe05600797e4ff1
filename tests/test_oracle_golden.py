"""The oracle reproduces the committed golden fixtures (tests/golden/golden_parity.npz).

The fixtures were written by tests/golden/make_golden.py; this pins the oracle against
accidental change: any edit to its formulas that moves lnL, the response or the waveform at
the 1e-12 level fails here before it can move the GPU parity tests' reference.
"""

import json
import os

import numpy as np
import pytest

from tests.golden import make_golden as mg

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_parity.npz"))


@pytest.mark.parametrize("name", list(mg.CASES))
def test_oracle_reproduces_golden(name):
    spec = json.loads(str(FIX[f"{name}/spec"]))
    assert spec == json.loads(json.dumps(mg.case_spec(name)))
    r = mg.compute(spec)
    np.testing.assert_allclose(r["lnl"], FIX[f"{name}/lnl"], rtol=1e-12)
    np.testing.assert_array_equal(r["idx"], FIX[f"{name}/idx"])
    for k in ("resp", "wf"):
        a, b = r[k], FIX[f"{name}/{k}"]
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b)), k
