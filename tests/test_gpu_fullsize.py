"""The named 2D workloads at their full sizes.

- 2D Poiseuille, 4 x 64 cells, GH24, Nm=8, Kn=0.1, phi=0.05: FP64 against
  the oracle over 300 steps; mixed against FP64 over the whole 2000 steps.
- Cartesian Rayleigh-Taylor, 120 x 120 cells, GH28, Nm=10, Kn=1e-3: FP64
  and mixed against the oracle over 2 steps.
- Cartesian Rayleigh-Taylor, 512 x 512 cells, GH40, Nm=12 (4.2e8
  phase-space points; too large for the oracle): properties that do not
  depend on size - mass conserved to round-off between specular walls,
  bitwise-repeatable results, mixed within 1e-5 of FP64.
"""
import numpy as np
import pytest

from hugks import capi, cases
import oracle_py
from metrics import field_errors

pytestmark = pytest.mark.gpu


def poiseuille_4x64():
    return dict(nx=4, ny=64, xlength=0.1, ylength=1.0, nv=24, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8, cfl=0.8,
                maxtime=200.0, case_id=5, knudsen=0.1, phi=0.05, TwL=1.0, TwR=1.0)


def rt_cartesian(n, nv, nm):
    return dict(nx=n, ny=n, xlength=0.5, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=nm, cfl=0.8,
                maxtime=200.0, case_id=6, knudsen=1e-3, phi=1.0, TwL=1.0, TwR=1.0)


def gpu_run(case, steps, sched=None):
    d = case.desc
    print(f"run {d.nx}x{d.ny} nv={d.nv} nm={d.nm} prec={d.precision} steps={steps} "
          f"sched={sched}", flush=True)
    with capi.Solver(case) as s:
        if sched:
            s.set_schedule(*sched)
        w0 = s.macro(with_q=False)
        info = s.advance(steps)
        w = s.macro(with_q=False)
    return info, w0, w


def test_poiseuille_4x64_fp64_against_oracle(have_gpu):
    case = cases.build_2d(poiseuille_4x64(), capi.FP64)
    info, _, w = gpu_run(case, 300)
    o = oracle_py.Oracle(case)
    o.step(300)
    wr = o.get()[0]
    err = field_errors(w, wr)
    print("poiseuille 300 steps W err", err, flush=True)
    assert np.all(err <= 1e-10), err


def test_poiseuille_4x64_mixed_against_fp64_2000_steps(have_gpu):
    p = poiseuille_4x64()
    _, _, w64 = gpu_run(cases.build_2d(p, capi.FP64), 2000)
    info, w0, w32 = gpu_run(cases.build_2d(p, capi.MIXED_FP32), 2000)
    err = field_errors(w32, w64)
    mass = abs(w32[:, 0].sum() - w0[:, 0].sum()) / w0[:, 0].sum()
    print("poiseuille 2000 steps mixed vs fp64", err, "mass drift", mass, flush=True)
    assert info.steps == 2000
    assert np.all(err <= 1e-5), err


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-10), (capi.MIXED_FP32, 1e-5)])
def test_rt_120_against_oracle(have_gpu, precision, tol):
    case = cases.build_2d(rt_cartesian(120, 28, 10), precision, rt_variant="cartesian", seed=1)
    _, _, w = gpu_run(case, 2)
    o = oracle_py.Oracle(case.with_precision(capi.FP64))
    o.step(2)
    err = field_errors(w, o.get()[0])
    print("rt120 prec", precision, "W err", err, flush=True)
    assert np.all(err <= tol), err


def test_rt_512_properties(have_gpu):
    p = rt_cartesian(512, 40, 12)
    steps = 3
    c64 = cases.build_2d(p, capi.FP64, rt_variant="cartesian", seed=1)
    _, w0, w64 = gpu_run(c64, steps)
    m0 = w0[:, 0].sum()
    assert abs(w64[:, 0].sum() - m0) / m0 <= 1e-13
    c32 = cases.build_2d(p, capi.MIXED_FP32, rt_variant="cartesian", seed=1)
    _, _, w32a = gpu_run(c32, steps)
    _, _, w32b = gpu_run(c32, steps)
    assert np.array_equal(w32a, w32b)
    assert abs(w32a[:, 0].sum() - m0) / m0 <= 1e-13
    assert np.all(np.isfinite(w32a))
    err = field_errors(w32a, w64)
    print("rt512 mixed vs fp64", err, flush=True)
    assert np.all(err <= 1e-5), err


@pytest.mark.parametrize("prec", [capi.MIXED_FP32, capi.FP64])
def test_arrays_over_4gib_match_narrow_channel(have_gpu, prec):
    """Distribution arrays past 4 GiB take the 64-bit addressing of the face
    kernels (32-bit buffer offsets cover smaller ones).  A channel flow that
    is uniform along x, 11000 x 64 cells at GH40 (4.5 GB per array in FP32
    storage, 9 GB in FP64), must give in every column exactly what the same
    channel 4 cells wide gives (x-face fluxes cancel exactly between equal
    cells).  Both run the schedule the wide channel takes by default (the
    face-forming update needs arrays under 4 GiB)."""
    p = dict(nx=4, ny=64, xlength=0.1, ylength=1.0, nv=40, vmax=5.0, Tref=1.0,
             quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8, cfl=0.8,
             maxtime=200.0, case_id=5, knudsen=0.1, phi=0.05, TwL=1.0, TwR=1.2)
    steps = 3
    wide_p = dict(p, nx=11000, xlength=0.1 * 11000 / 4)
    wide = cases.build_2d(wide_p, prec)
    assert wide.ncell * wide.desc.nv ** 2 * 4 > 2 ** 32
    sched = (capi.FLUX_CELLSUM, capi.UPDATE_MFMA_NODE)
    narrow = cases.build_2d(p, prec)
    _, _, wn = gpu_run(narrow, steps, sched)
    print(f"run 11000x64 nv=40 prec={prec} steps={steps} sched={sched}", flush=True)
    with capi.Solver(wide) as s:
        # the default where the face-forming update cannot run, which refuses
        assert s.schedule() == sched
        with pytest.raises(capi.HugksError) as ei:
            s.set_schedule(capi.FLUX_CELLSUM, capi.UPDATE_MFMA_FACE)
        assert ei.value.code == capi.E_UNSUPPORTED
        assert s.schedule() == sched
        s.advance(steps)
        ww = s.macro(with_q=False)
    wn = wn.reshape(64, 4, -1)
    ww = ww.reshape(64, 11000, -1)
    for i in (0, 1, 5000, 10999):
        assert np.array_equal(ww[:, i, :], wn[:, 0, :]), i
    assert np.array_equal(ww, np.repeat(ww[:, :1, :], 11000, axis=1))


@pytest.mark.parametrize("prec", [capi.MIXED_FP32, capi.FP64])
def test_face_update_arrays_near_4gib_match_narrow_channel(have_gpu, prec):
    """The face-forming update and its per-cell divergence address whole
    distribution arrays with 32-bit byte offsets, which the library allows up
    to 4 GiB - 16 bytes.  A channel uniform along x with each array just
    under that (10485 x 64 cells at GH40 in FP32 storage, 5242 x 64 in FP64)
    runs that schedule and gives in every column exactly what the channel 4
    cells wide gives."""
    p = dict(nx=4, ny=64, xlength=0.1, ylength=1.0, nv=40, vmax=5.0, Tref=1.0,
             quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8, cfl=0.8,
             maxtime=200.0, case_id=5, knudsen=0.1, phi=0.05, TwL=1.0, TwR=1.2)
    steps = 3
    nx = 10485 if prec == capi.MIXED_FP32 else 5242
    es = 4 if prec == capi.MIXED_FP32 else 8
    sched = (capi.FLUX_CELLSUM, capi.UPDATE_MFMA_FACE)
    wide = cases.build_2d(dict(p, nx=nx, xlength=0.1 * nx / 4), prec)
    nbytes = wide.ncell * 1600 * es
    assert 2 ** 32 - 2 ** 26 < nbytes < 2 ** 32 - 16
    narrow = cases.build_2d(p, prec)
    _, _, wn = gpu_run(narrow, steps, sched)
    print(f"run {nx}x64 nv=40 prec={prec} steps={steps} sched={sched} ({nbytes} B per array)",
          flush=True)
    with capi.Solver(wide) as s:
        assert s.schedule() == sched
        s.advance(steps)
        ww = s.macro(with_q=False)
    wn = wn.reshape(64, 4, -1)
    ww = ww.reshape(64, nx, -1)
    for i in (0, 1, nx // 2, nx - 1):
        assert np.array_equal(ww[:, i, :], wn[:, 0, :]), i
    assert np.array_equal(ww, np.repeat(ww[:, :1, :], nx, axis=1))
