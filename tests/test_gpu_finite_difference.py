"""The finite-difference force (force_type 'finite_difference':
cal_fsource1, 2d/src/gravity.f90:108-187, and cal_fv1,
1d/src/gravity.f90:127-154) against the compiled reference's own outputs
(tests/golden/*_fd_*.npz, written by oracle/_ref through
tests/golden/make_golden.py):

- 2D force-driven channel on a Newton-Cotes grid (nv = 21), 20 steps;
- 2D Cartesian Rayleigh-Taylor on a Gauss-Hermite grid (nv = 12, the
  reference's uniform difference step on non-uniform nodes), 12 steps;
- 1D Sod tube under gravity, Newton-Cotes nv = 100, 50 steps.

The reference's differences, reproduced as written: the u-axis interior
difference of b takes three of its terms from h, and the last v node's
one-sided difference has the forward-difference sign (the oracle with
either changed fails these fixtures).

FP64 with each 2D flux schedule (the update is the LDS kernel, the only
one with this force): conserved fields and distributions within 1e-10,
the same iteration index and simulated time.  Mixed FP32: conserved
fields within 1e-5.  The MFMA updates and the bf16x3 mode, which have no
difference operator, are refused with HUGKS_E_UNSUPPORTED.
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
FD = sorted(glob.glob(os.path.join(GOLDEN, "*_fd_*.npz")))
FD2D = [p for p in FD if os.path.basename(p).startswith("p2d")]
FLUX = {"fused": capi.FLUX_FUSED, "split": capi.FLUX_SPLIT, "stencil": capi.FLUX_STENCIL,
        "cellsum": capi.FLUX_CELLSUM}


def _ids(paths):
    return [os.path.basename(p)[:-4] for p in paths]


def test_fixtures_present():
    assert len(FD) == 3 and len(FD2D) == 2, FD


def _check(fx, info, w, tol, hf=None, bf=None):
    assert info.iter == int(fx["iter"])
    ttol = 1e-12 if tol < 1e-9 else tol
    assert abs(info.simtime - float(fx["simtime"])) <= ttol * float(fx["simtime"])
    err = field_errors(w, fx["w"], thermal=True)
    fe = 0.0
    if hf is not None:
        fe = max(np.max(np.abs(hf - fx["h"])) / np.max(np.abs(fx["h"])),
                 np.max(np.abs(bf - fx["b"])) / np.max(np.abs(fx["b"])))
    print(f"  W err {err} f err {fe:.3e}", flush=True)
    assert np.all(err <= tol) and fe <= tol, (err, fe)


@pytest.mark.parametrize("flux", sorted(FLUX))
@pytest.mark.parametrize("path", FD2D, ids=_ids(FD2D))
def test_fd_2d_fp64_every_flux(have_gpu, path, flux):
    fx, dim, params, gpu = load_fixture(path)
    case = case_of(fx, dim, params, gpu, capi.FP64)
    steps = int(fx["steps"])
    print(f"{os.path.basename(path)} flux={flux} steps={steps}", flush=True)
    with capi.Solver(case) as s:
        assert s.schedule()[1] == capi.UPDATE_LDS
        s.set_schedule(FLUX[flux], capi.UPDATE_LDS)
        info = s.advance(steps)
        w = s.macro(with_q=False)
        hf, bf = s.get_f()
    _check(fx, info, w, 1e-10, hf, bf)


@pytest.mark.parametrize("path", FD, ids=_ids(FD))
def test_fd_fp64_and_mixed(have_gpu, path):
    fx, dim, params, gpu = load_fixture(path)
    steps = int(fx["steps"])
    for prec, tol in ((capi.FP64, 1e-10), (capi.MIXED_FP32, 1e-5)):
        case = case_of(fx, dim, params, gpu, prec)
        print(f"{os.path.basename(path)} prec={prec} steps={steps}", flush=True)
        with capi.Solver(case) as s:
            info = s.advance(steps)
            w = s.macro(with_q=False)
            hf, bf = s.get_f() if prec == capi.FP64 else (None, None)
        _check(fx, info, w, tol, hf, bf)


def test_fd_refusals(have_gpu):
    fx, dim, params, gpu = load_fixture(FD2D[0])
    case = case_of(fx, dim, params, gpu, capi.FP64)
    with capi.Solver(case) as s:
        for uk in (capi.UPDATE_MFMA, capi.UPDATE_MFMA_NODE):
            with pytest.raises(capi.HugksError) as e:
                s.set_schedule(capi.FLUX_SPLIT, uk)
            assert e.value.code == capi.E_UNSUPPORTED
    b3 = case_of(fx, dim, params, gpu, capi.MIXED_BF16X3)
    with pytest.raises(capi.HugksError) as e:
        capi.Solver(b3)
    assert e.value.code == capi.E_UNSUPPORTED
