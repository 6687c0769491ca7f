"""The named 2D configurations at full size against the compiled reference
solver's own outputs (tests/golden/large/*.npz, written by oracle/_ref
through tests/golden/make_golden.py; conserved variables, step sizes and
loop state only):

- Poiseuille 4 x 64, GH24, Nm = 8, Kn = 0.1, phi = 0.05: all 2000 steps;
- Cartesian Rayleigh-Taylor 120 x 120, GH28, Nm = 10, Kn = 1e-3: 12 steps;
- a 512 x 16-cell band of the 512 x 512 Rayleigh-Taylor case (the same
  cell size, GH40, Nm = 12): 4 steps.

FP64 path: every conserved field within 1e-10 of the reference, the same
iteration index and simulated time.  Mixed paths (FP32 f and flux, FP64
moments; Hermite contractions on FP32 tiles or as bf16x3 split products):
within 1e-5 on the same measure; their per-field errors in rho, U, V, T
are printed.  Momentum is measured against the
largest momentum of the developed channel flow, and against the thermal
momentum sqrt(max rho * max rho E) in the Rayleigh-Taylor cases, which
are a few steps from rest (velocities of 1e-4).  The reference
being matched: 2d/src/solver.f90:13-395, 2d/src/flux.f90:28-689,
2d/src/hermite_symmetric2d.f90:23-104.
"""
import glob
import os

import numpy as np
import pytest

from hugks import capi
from metrics import field_errors, prim_errors
from test_oracle_golden import case_of, load_fixture

pytestmark = pytest.mark.gpu

LARGE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large")
FIXTURES = sorted(glob.glob(os.path.join(LARGE, "*.npz")))
TOL = {capi.FP64: 1e-10, capi.MIXED_FP32: 1e-5, capi.MIXED_BF16X3: 1e-5}


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED_FP32, capi.MIXED_BF16X3],
                         ids=["fp64", "mixed", "bf16x3"])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_full_size_against_reference(have_gpu, path, precision):
    fx, dim, params, gpu = load_fixture(path)
    case = case_of(fx, dim, params, gpu, precision)
    steps = int(fx["steps"])
    d = case.desc
    print(f"{os.path.basename(path)} {d.nx}x{d.ny} nv={d.nv} Nm={d.nm} prec={precision} "
          f"steps={steps}", flush=True)
    with capi.Solver(case) as s:
        info = s.advance(steps)
        w = s.macro(with_q=False)
    assert info.steps == steps
    assert info.iter == int(fx["iter"])
    # dt follows W: exact in FP64, within the field tolerance when mixed
    ttol = 1e-12 if precision == capi.FP64 else TOL[precision]
    assert abs(info.simtime - float(fx["simtime"])) <= ttol * float(fx["simtime"])
    assert abs(info.dt - float(fx["dts"][-1])) <= ttol * float(fx["dts"][-1])
    wr = fx["w"]
    err = field_errors(w, wr, thermal=int(params["case_id"]) == 6)
    pe = prim_errors(w, wr, d.gamma)
    print(f"  W err {err}  (rho, U, V, T) err {pe}", flush=True)
    assert np.all(err <= TOL[precision]), err
