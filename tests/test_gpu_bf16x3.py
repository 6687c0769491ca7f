"""The bf16x3 split-product Hermite contraction (HUGKS_MIXED_BF16X3): FP32
f and flux, FP64 moments and W, and the Hermite projection and force
contractions of the MFMA node update as bf16 MFMA products of operands
split exactly into three bf16 pieces, accumulated in FP32
(2d/src/hermite_symmetric2d.f90:69-98 is the contraction replaced).

- The headline velocity grid (40 x 40 nodes, Nm = 12), Rayleigh-Taylor and
  channel, with every flux schedule: after 6 steps every conserved field
  within 1e-5 of the FP64 oracle.
- One step from the same state: the distributions differ from those of
  the FP32-tile contraction by less than 1e-6 of their largest value.  The
  two paths share every other kernel, so this bounds the contraction's own
  error at the FP32 level: a contraction in plain bf16 (8-bit products)
  differs by about dt |F| 2^-8, 2e-5 of f in these cases.
- Batched bf16x3 cases (which keep their own launches) equal the same
  cases stepped alone, bitwise, next to grouped FP32 cases.
- The mode runs only the MFMA node update: other update kernels and
  Hermite orders above 14 are refused with HUGKS_E_UNSUPPORTED.
The full-size configurations against the compiled reference, in this mode
too, are in test_gpu_reference_fullsize.py.
"""
import numpy as np
import pytest

from hugks import capi, cases, namelist
from metrics import field_errors
from test_gpu_headline_grid import CASES, FLUX, STEPS, oracle_after

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flux", sorted(FLUX))
def test_bf16x3_headline_grid(have_gpu, name, flux):
    wr, _, _, (t, it, _) = oracle_after(name, STEPS)
    case = CASES[name](capi.MIXED_BF16X3)
    print(f"bf16x3 {name} nv=40 Nm=12 flux={flux} steps={STEPS}", flush=True)
    with capi.Solver(case) as s:
        s.set_schedule(FLUX[flux], capi.UPDATE_MFMA_NODE)
        info = s.advance(STEPS)
        w = s.macro(with_q=False)
    assert info.steps == STEPS and info.iter == it
    assert abs(info.simtime - t) <= TOL * t
    err = field_errors(w, wr)
    print(f"  W err {err}", flush=True)
    assert np.all(err <= TOL), err


@pytest.mark.parametrize("name", sorted(CASES))
def test_bf16x3_contraction_at_fp32_level(have_gpu, name):
    out = {}
    for prec in (capi.MIXED_FP32, capi.MIXED_BF16X3):
        case = CASES[name](prec)
        print(f"bf16x3 vs FP32 tiles {name} prec={prec}", flush=True)
        with capi.Solver(case) as s:
            s.set_schedule(capi.FLUX_SPLIT, capi.UPDATE_MFMA_NODE)
            s.advance(1)
            out[prec] = s.get_f()
    (h32, b32), (h3, b3) = out[capi.MIXED_FP32], out[capi.MIXED_BF16X3]
    dh = np.max(np.abs(h3 - h32)) / np.max(np.abs(h32))
    db = np.max(np.abs(b3 - b32)) / np.max(np.abs(b32))
    print(f"  max |f_bf16x3 - f_fp32| / max |f|: h {dh:.3e}, b {db:.3e}", flush=True)
    assert dh <= 1e-6 and db <= 1e-6, (dh, db)


def test_bf16x3_batch_equals_cases_alone(have_gpu):
    ps = cases.sweep_c4(seed=3, n=6)
    precs = [capi.MIXED_BF16X3, capi.MIXED_FP32] * 3
    cs = [cases.build_2d(p, pr) for p, pr in zip(ps, precs)]
    print(f"bf16x3 batch: {len(cs)} cases, nv {[c.desc.nv for c in cs]}", flush=True)
    with capi.Batch(cs) as b:
        assert b.grouped_cases == 3      # the FP32 ones
        infos = b.advance(5)
        assert all(i.steps == 5 for i in infos)
        wb = [s.macro(with_q=False) for s in b.solvers]
        scheds = [s.schedule() for s in b.solvers]
        waves = [s.update_waves() for s in b.solvers]
    for k, (c, sc, nw) in enumerate(zip(cs, scheds, waves)):
        with capi.Solver(c) as s:
            s.set_schedule(*sc)
            s.set_update_waves(nw)
            s.advance(5)
            w = s.macro(with_q=False)
        assert np.array_equal(w, wb[k]), k


def test_bf16x3_refusals(have_gpu):
    case = CASES["channel"](capi.MIXED_BF16X3)
    with capi.Solver(case) as s:
        assert s.schedule()[1] == capi.UPDATE_MFMA_NODE
        for uk in (capi.UPDATE_MFMA, capi.UPDATE_LDS):
            with pytest.raises(capi.HugksError) as e:
                s.set_schedule(capi.FLUX_SPLIT, uk)
            assert e.value.code == capi.E_UNSUPPORTED
            assert "BF16X3" in str(e.value)
    p = dict(nx=4, ny=8, nv=16, Nm=15)
    big = cases.build_2d({**namelist.DEFAULTS_2D, **p}, capi.MIXED_BF16X3)
    with pytest.raises(capi.HugksError) as e:
        capi.Solver(big)
    assert e.value.code == capi.E_UNSUPPORTED
    assert "BF16X3" in str(e.value)
