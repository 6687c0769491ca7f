"""The batched sweep workload: 64 seeded force-driven channel cases of
uneven size (hugks.cases.sweep_c4: Kn log-uniform on [0.01, 1], phi on
[0.01, 0.2], ny on [32, 128], nv in {16, 20, 24, 28, 32}, Nm = 8, half of
them Fourier-type channels with a warmer upper wall), advanced together by
one graph replay per step with the grouped launches.

- Every case of the batch equals, bitwise, the same case stepped on its
  own with the same kernel schedule (mixed and FP64 batches).
- The smallest case of each velocity grid agrees with the FP64 oracle:
  FP64 within 1e-10, mixed within 1e-5.
- After 2000 steps (developed channel flows) the mixed batch stays within
  1e-5 of the FP64 batch in every case in rho and T, in the velocities
  against the thermal speed and in the heat flux against its thermal
  scale.  Against each case's own largest flow speed the velocity error
  stays under 1e-4: the slowest channels (Mach 0.01) carry an absolute
  velocity error of about 1e-6 of the thermal speed, which the FP32
  distribution function itself sets (the FP64 oracle with h and b
  rounded to FP32 after every step shows the same error).
The reference runs each case as its own process (2d/src/main.f90:13-137).
"""
import numpy as np
import pytest

from hugks import capi, cases
import oracle_py
from metrics import field_errors, prim_errors, prim_errors_thermal, q_error, q_error_thermal

pytestmark = pytest.mark.gpu

STEPS = 10


def sweep(prec):
    return [cases.build_2d(p, prec) for p in cases.sweep_c4()]


@pytest.mark.parametrize("prec", [capi.MIXED_FP32, capi.FP64], ids=["mixed", "fp64"])
def test_sweep_batch_equals_cases_alone_and_oracle(have_gpu, prec):
    cs = sweep(prec)
    print(f"c4 sweep: {len(cs)} cases, {sum(c.ncell * c.nn for c in cs)} points, prec={prec}",
          flush=True)
    with capi.Batch(cs, autotune=True) as b:
        assert b.grouped_cases == len(cs)
        infos = b.advance(STEPS)
        assert all(i.steps == STEPS for i in infos)
        wb = [s.macro(with_q=False) for s in b.solvers]
        scheds = [s.schedule() for s in b.solvers]
        waves = [s.update_waves() for s in b.solvers]
    print(f"  tuned waves per cell: {sorted(set(waves))}", flush=True)
    for k, (c, sc, nw) in enumerate(zip(cs, scheds, waves)):
        with capi.Solver(c) as s:
            s.set_schedule(*sc)
            s.set_update_waves(nw)
            s.advance(STEPS)
            w = s.macro(with_q=False)
        assert np.array_equal(w, wb[k]), k
    # the smallest case of each velocity grid against the oracle
    tol = 1e-10 if prec == capi.FP64 else 1e-5
    picks = {}
    for k, c in enumerate(cs):
        nv = c.desc.nv
        if nv not in picks or c.ncell < cs[picks[nv]].ncell:
            picks[nv] = k
    for nv, k in sorted(picks.items()):
        o = oracle_py.Oracle(cs[k].with_precision(capi.FP64))
        o.step(STEPS)
        err = field_errors(wb[k], o.get()[0], thermal=True)
        print(f"  case {k} nv={nv} ny={cs[k].desc.ny} vs oracle {err}", flush=True)
        assert np.all(err <= tol), (k, err)


def test_sweep_mixed_against_fp64_developed(have_gpu):
    steps = 2000
    res = {}
    for prec in (capi.MIXED_FP32, capi.FP64):
        cs = sweep(prec)
        print(f"c4 sweep {steps} steps prec={prec}", flush=True)
        with capi.Batch(cs) as b:
            w0 = [s.macro(with_q=False) for s in b.solvers]
            infos = b.advance(steps)
            assert all(i.steps == steps for i in infos)
            res[prec] = [s.macro() for s in b.solvers], w0, cs
    (m, w0m, cs), (f, _, _) = res[capi.MIXED_FP32], res[capi.FP64]
    worst_flow = np.zeros(4)
    worst_th = np.zeros(4)
    worst_q = worst_qth = drift = 0.0
    within = 0
    for k in range(len(cs)):
        (wm, qm), (wf, qf) = m[k], f[k]
        g = cs[k].desc.gamma
        pe = prim_errors(wm, wf, g)
        pt = prim_errors_thermal(wm, wf, g)
        worst_flow = np.maximum(worst_flow, pe)
        worst_th = np.maximum(worst_th, pt)
        worst_q = max(worst_q, q_error(qm, qf))
        worst_qth = max(worst_qth, q_error_thermal(qm, qf, wf, g))
        within += int(np.all(pe <= 1e-5))
        # relative mass drift from t = 0 of the mixed path (periodic in x,
        # walls of zero net mass flux)
        m0 = w0m[k][:, 0].sum()
        drift = max(drift, abs(wm[:, 0].sum() - m0) / m0)
    print(f"  worst over 64 cases: (rho, U, V, T) flow scale {worst_flow}, thermal scale "
          f"{worst_th}; q {worst_q:.3e} (own scale), {worst_qth:.3e} (thermal); "
          f"{within} of 64 cases within 1e-5 on the flow scale; mass drift {drift:.3e}",
          flush=True)
    assert np.all(worst_th <= 1e-5) and worst_qth <= 1e-5
    assert np.all(worst_flow <= 1e-4)
    assert drift <= 1e-5
