"""hugks_get_macro_noalloc (the output read-back that allocates no device
memory, staging q in the library's static device array) against
hugks_get_macro on the same handle: bitwise equal W and q, on two calls in
a row on cases of decreasing size (the second call finds the first one's
values in the shared array and must overwrite every one it reads back),
in both storage types and in 1D; sizes above the static staging are
refused with HUGKS_E_UNSUPPORTED while W alone can still be read."""
import numpy as np
import pytest

from hugks import capi, cases
from test_gpu_parity import poiseuille, rt
from test_gpu_parity_1d import from_file

pytestmark = pytest.mark.gpu


def read_both(s):
    w_a, q_a = s.macro()
    w_b, q_b = s.macro_noalloc()
    return w_a, q_a, w_b, q_b


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED_FP32])
def test_noalloc_matches_get_macro_on_shrinking_cases(have_gpu, prec):
    big = cases.build_2d(rt(nx=24, ny=32, nv=12), prec, rt_variant="cartesian", seed=4)
    small = cases.build_2d(poiseuille(nx=4, ny=6, nv=12), prec)
    for case, steps in ((big, 3), (small, 2)):
        d = case.desc
        print(f"noalloc {d.nx}x{d.ny} nv={d.nv} prec={prec}", flush=True)
        with capi.Solver(case) as s:
            s.advance(steps)
            w_a, q_a, w_b, q_b = read_both(s)
            assert np.array_equal(w_a, w_b)
            assert np.array_equal(q_a, q_b)
            assert np.all(np.isfinite(q_b)) and np.max(np.abs(q_b)) > 0
            # W alone, q alone
            w_c = s.macro_noalloc(with_q=False)
            assert np.array_equal(w_a, w_c)


def test_noalloc_1d(have_gpu):
    case = cases.build_1d(from_file("p1d_sod_40.namelist"), capi.FP64)
    print("noalloc 1D sod40", flush=True)
    with capi.Solver(case) as s:
        s.advance(5)
        w_a, q_a, w_b, q_b = read_both(s)
    assert np.array_equal(w_a, w_b) and np.array_equal(q_a, q_b)


def test_noalloc_refuses_sizes_above_its_staging(have_gpu):
    # 1030 x 520 cells: 2 * 535600 heat-flux values > 2^20
    p = dict(poiseuille(nx=1030, ny=520, nv=4, nm=2), xlength=0.1 * 1030 / 4)
    case = cases.build_2d(p, capi.MIXED_FP32)
    assert 2 * case.ncell > 1 << 20
    print(f"noalloc refusal {case.desc.nx}x{case.desc.ny}", flush=True)
    with capi.Solver(case) as s:
        with pytest.raises(capi.HugksError) as ei:
            s.macro_noalloc()
        assert ei.value.code == capi.E_UNSUPPORTED
        assert "exceed" in str(ei.value)
        w = s.macro_noalloc(with_q=False)
        assert np.array_equal(w, s.macro(with_q=False))
