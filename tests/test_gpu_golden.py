"""The HIP library against the compiled reference's own outputs
(tests/golden/*.npz, written by oracle/_ref through
tests/golden/make_golden.py), FP64 path at 1e-10: step sizes, loop state,
conserved fields and the distribution functions."""
import glob
import os

import numpy as np
import pytest

from hugks import capi
from metrics import field_errors
from test_oracle_golden import case_of, load_fixture

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))
TOL = 1e-10


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_gpu_matches_reference_outputs(have_gpu, path):
    fx, dim, params, gpu = load_fixture(path)
    case = case_of(fx, dim, params, gpu, capi.FP64)
    steps = int(fx["steps"])
    print(f"fixture {os.path.basename(path)} steps={steps}", flush=True)
    dts = []
    with capi.Solver(case) as s:
        for _ in range(steps):
            dts.append(s.step().dt)
        w, _ = s.macro()
        hf, bf = s.get_f()
        t, it = s.time()
    assert it == int(fx["iter"])
    np.testing.assert_allclose(dts, fx["dts"], rtol=1e-12, atol=0)
    assert abs(t - float(fx["simtime"])) <= 1e-12 * abs(float(fx["simtime"]))
    err = field_errors(w, fx["w"], thermal=True)
    fe = max(np.max(np.abs(hf - fx["h"])) / np.max(np.abs(fx["h"])),
             np.max(np.abs(bf - fx["b"])) / np.max(np.abs(fx["b"])))
    print(os.path.basename(path), "W err", err, "f err", fe, flush=True)
    assert np.all(err <= TOL), err
    assert fe <= TOL
