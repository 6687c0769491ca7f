"""Newton-Cotes velocity grids (uniform nodes with the closed 4-panel
Newton-Cotes weights: 2d/src/velocity.f90:29-61, 1d/src/velocity.f90:34-63)
against the compiled reference's own outputs (tests/golden/*nc*.npz,
written by oracle/_ref through tests/golden/make_golden.py):

- 2D force-driven channel, nv = 21 (node u = 0 on the grid), 20 steps;
- 2D Cartesian Rayleigh-Taylor, nv = 17, 12 steps;
- 1D Sod tube under gravity on the reference's 1D default grid
  (Newton-Cotes, nv = 100), 50 steps.

FP64 with every 2D kernel schedule: conserved fields and distributions
within 1e-10, the same iteration index and simulated time.  Mixed paths
(FP32 f; Hermite contractions on FP32 tiles, and as bf16x3 split
products in 2D): conserved fields within 1e-5.
"""
import glob
import os

import numpy as np
import pytest

from hugks import capi
from metrics import field_errors
from test_oracle_golden import case_of, load_fixture

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# (the finite-difference fixtures on Newton-Cotes grids are test_gpu_finite_difference.py's)
NC = sorted(p for p in glob.glob(os.path.join(GOLDEN, "*_nc*.npz")) if "_fd_" not in p)
NC2D = [p for p in NC if os.path.basename(p).startswith("p2d")]
SCHED = [(f, u) for f in (capi.FLUX_FUSED, capi.FLUX_SPLIT, capi.FLUX_STENCIL, capi.FLUX_CELLSUM)
         for u in (capi.UPDATE_LDS, capi.UPDATE_MFMA, capi.UPDATE_MFMA_NODE)] + [
    (capi.FLUX_CELLSUM, capi.UPDATE_MFMA_FACE), (capi.FLUX_SPLIT, capi.UPDATE_MFMA_FACE)]


def _ids(paths):
    return [os.path.basename(p)[:-4] for p in paths]


def test_fixtures_present():
    assert len(NC) == 3 and len(NC2D) == 2, NC


def _run(case, steps, sched=None):
    with capi.Solver(case) as s:
        if sched is not None:
            s.set_schedule(*sched)
        info = s.advance(steps)
        w = s.macro(with_q=False)
        hf, bf = s.get_f()
    return info, w, hf, bf


@pytest.mark.parametrize("sched", SCHED, ids=[f"flux{f}-update{u}" for f, u in SCHED])
@pytest.mark.parametrize("path", NC2D, ids=_ids(NC2D))
def test_newton_cotes_2d_every_schedule_fp64(have_gpu, path, sched):
    fx, dim, params, gpu = load_fixture(path)
    case = case_of(fx, dim, params, gpu, capi.FP64)
    steps = int(fx["steps"])
    print(f"{os.path.basename(path)} nv={case.desc.nv} schedule={sched} steps={steps}",
          flush=True)
    info, w, hf, bf = _run(case, steps, sched)
    assert info.iter == int(fx["iter"])
    assert abs(info.simtime - float(fx["simtime"])) <= 1e-12 * float(fx["simtime"])
    err = field_errors(w, fx["w"], thermal=True)
    fe = max(np.max(np.abs(hf - fx["h"])) / np.max(np.abs(fx["h"])),
             np.max(np.abs(bf - fx["b"])) / np.max(np.abs(fx["b"])))
    print(f"  W err {err} f err {fe:.3e}", flush=True)
    assert np.all(err <= 1e-10) and fe <= 1e-10, (err, fe)


@pytest.mark.parametrize("path", NC, ids=_ids(NC))
def test_newton_cotes_mixed(have_gpu, path):
    fx, dim, params, gpu = load_fixture(path)
    precs = [capi.MIXED_FP32] + ([capi.MIXED_BF16X3] if dim == 2 else [])
    for prec in precs:
        case = case_of(fx, dim, params, gpu, prec)
        steps = int(fx["steps"])
        print(f"{os.path.basename(path)} prec={prec} steps={steps}", flush=True)
        info, w, _, _ = _run(case, steps)
        assert info.iter == int(fx["iter"])
        assert abs(info.simtime - float(fx["simtime"])) <= 1e-5 * float(fx["simtime"])
        err = field_errors(w, fx["w"], thermal=True)
        print(f"  W err {err}", flush=True)
        assert np.all(err <= 1e-5), err
