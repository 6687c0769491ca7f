"""Device-detected failures reach the caller as status codes, as the
reference's error stops do (2d/src/hermite_symmetric2d.f90:39-40 on
T <= 0; 1d/src/flux.f90:338-341 on a zero wall coefficient).

The conditions are set up from the host, so every kernel runs to its
normal end on finite or non-finite numbers and only the device flag
word tells: a state with negative internal energy in one cell (set
through hugks_set_f) gives HUGKS_E_NUMERIC; weights that are zero on
every node leaving a diffuse wall give HUGKS_E_WALL.  A batch reports the
index of the case that failed, and its other cases keep going.
"""
import numpy as np
import pytest

from hugks import capi, cases
from test_gpu_parity import poiseuille, rt
from test_gpu_parity_1d import from_file

pytestmark = pytest.mark.gpu


def negative_energy_state(case, cell=3):
    """The initial f of the case and a W whose cell `cell` has a total
    energy below its kinetic energy (T < 0)."""
    with capi.Solver(case) as s:
        w = s.macro(with_q=False).copy()
        hf, bf = s.get_f()
    w[cell, -1] = -abs(w[cell, -1])
    return hf, bf, w


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED_FP32])
@pytest.mark.parametrize("name", ["poiseuille", "rt"])
def test_negative_temperature_is_numeric_error(have_gpu, name, prec):
    p = poiseuille(nx=4, ny=8, nv=12) if name == "poiseuille" else rt(nx=6, ny=8, nv=12)
    case = cases.build_2d(p, prec, rt_variant="cartesian", seed=2)
    print(f"E_NUMERIC {name} prec={prec}", flush=True)
    hf, bf, w = negative_energy_state(case)
    with capi.Solver(case) as s:
        s.set_f(hf, bf, w)
        with pytest.raises(capi.HugksError) as ei:
            s.advance(3)
        assert ei.value.code == capi.E_NUMERIC
        assert "T <= 0" in str(ei.value)
        # stopped at the failing step; later calls run no step
        t0, it0 = s.time()
        with pytest.raises(capi.HugksError):
            s.advance(2)
        assert s.time()[1] == it0


def test_negative_temperature_1d(have_gpu):
    case = cases.build_1d(from_file("p1d_sod_40.namelist"), capi.FP64)
    print("E_NUMERIC 1D sod40", flush=True)
    hf, bf, w = negative_energy_state(case, cell=20)
    with capi.Solver(case) as s:
        s.set_f(hf, bf, w)
        with pytest.raises(capi.HugksError) as ei:
            s.advance(2)
        assert ei.value.code == capi.E_NUMERIC


def zero_outgoing_weights_2d(case):
    """Weights zero on every node with v > 0: the lower diffuse wall of the
    channel has no node to re-emit into the gas."""
    nv = case.desc.nv
    w = case.arrays["weight"].reshape(nv, nv)     # [jv, iu]
    w[case.arrays["v"] > 0, :] = 0.0
    return case


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED_FP32])
def test_zero_wall_coefficient_2d(have_gpu, prec):
    case = zero_outgoing_weights_2d(cases.build_2d(poiseuille(nx=4, ny=8, nv=12), prec))
    print(f"E_WALL 2D prec={prec}", flush=True)
    with capi.Solver(case) as s:
        with pytest.raises(capi.HugksError) as ei:
            s.step()
        assert ei.value.code == capi.E_WALL
        assert "wall" in str(ei.value)


def test_zero_wall_coefficient_1d(have_gpu):
    p = dict(from_file("ref1d_fourier_smoke.namelist"))
    case = cases.build_1d(p, capi.FP64)
    case.arrays["weight"][case.arrays["u"] > 0] = 0.0
    print("E_WALL 1D fourier", flush=True)
    with capi.Solver(case) as s:
        with pytest.raises(capi.HugksError) as ei:
            s.step()
        assert ei.value.code == capi.E_WALL


@pytest.mark.parametrize("grouped", [True, False])
def test_batch_names_failing_case(have_gpu, monkeypatch, grouped):
    """One bad case among good ones, grouped launches or one case per lane:
    hugks_batch_advance returns the failure of that case and names it; the
    other cases advance normally."""
    if not grouped:
        monkeypatch.setenv("HUGKS_BATCH_GROUPED", "0")
    good = [cases.build_2d(poiseuille(nx=4, ny=8 + 2 * k, nv=12), capi.MIXED_FP32)
            for k in range(3)]
    bad = zero_outgoing_weights_2d(cases.build_2d(poiseuille(nx=4, ny=9, nv=12),
                                                  capi.MIXED_FP32))
    cs = good[:2] + [bad] + good[2:]
    print(f"batch with failing case 2, grouped={grouped}", flush=True)
    with capi.Batch(cs) as b:
        assert (b.grouped_cases == len(cs)) == grouped
        infos = (capi.StepInfo * len(cs))()
        rc = b.lib.hugks_batch_advance(b.b, 4, infos)
        msg = b.lib.hugks_batch_last_error(b.b).decode()
        assert rc == capi.E_WALL, (rc, msg)
        assert msg.startswith("case 2:"), msg
        assert infos[2].device_flags & 2
        # the failing case stopped at its first step, as the reference stops
        assert infos[2].stop_reason == capi.STOP_ERROR and infos[2].steps == 1
        for k in (0, 1, 3):
            assert infos[k].device_flags == 0 and infos[k].steps == 4
            assert np.all(np.isfinite(b.solvers[k].macro(with_q=False)))
