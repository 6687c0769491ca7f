"""Parity of the HIP library (through the C-ABI) with the FP64 oracle.

FP64 path: after N steps every conserved field agrees with the oracle to
1e-10 relative to that field's largest magnitude.  Mixed path (FP32 f,
FP64 moments): 1e-5 on the same measure.
"""
import numpy as np
import pytest

from hugks import capi, cases
import oracle_py

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5


def poiseuille(nx=4, ny=16, nv=12, nm=6, kn=0.1, phi=0.05, force="hermite_symmetric", tw=(1.0, 1.0)):
    return dict(nx=nx, ny=ny, xlength=0.1, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type=force, Nm=nm, cfl=0.8, maxtime=200.0,
                case_id=5, knudsen=kn, phi=phi, TwL=tw[0], TwR=tw[1])


def rt(nx=12, ny=16, nv=12, nm=6, kn=1e-3, phi=1.0, xlength=0.5):
    return dict(nx=nx, ny=ny, xlength=xlength, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=nm, cfl=0.5,
                maxtime=0.14, case_id=6, knudsen=kn, phi=phi, TwL=1.0, TwR=1.0)


from metrics import field_errors as field_err, q_error


def run_pair(case, steps):
    print(f"case nx={case.desc.nx} ny={case.desc.ny} nv={case.desc.nv} nm={case.desc.nm} "
          f"prec={case.desc.precision} steps={steps}", flush=True)
    with capi.Solver(case) as s:
        info = s.advance(steps)
        w, q = s.macro()
        hf, bf = s.get_f()
    o = oracle_py.Oracle(case.with_precision(capi.FP64))
    o.step(steps)
    wr, hr, br = o.get()
    qr = o.heat_flux()
    t, it, dt = o.time()
    return dict(info=info, w=w, q=q, h=hf, b=bf, wr=wr, qr=qr, hr=hr, br=br, t=t, it=it, dt=dt)


@pytest.mark.parametrize("name,params,variant,steps", [
    ("poiseuille", poiseuille(), "reference", 20),
    ("poiseuille_legacy_force", poiseuille(force="hermite", tw=(1.0, 1.3)), "reference", 10),
    ("rt_reference", rt(xlength=1.0, phi=-1.5, kn=0.01), "reference", 10),
    ("rt_cartesian", rt(), "cartesian", 10),
])
def test_fp64_parity(have_gpu, name, params, variant, steps):
    case = cases.build_2d(params, capi.FP64, rt_variant=variant, seed=3)
    r = run_pair(case, steps)
    assert r["info"].steps == steps
    assert r["info"].iter == r["it"]
    assert abs(r["info"].simtime - r["t"]) <= 1e-12 * r["t"]
    err = field_err(r["w"], r["wr"])
    fe = np.max(np.abs(r["h"] - r["hr"])) / np.max(np.abs(r["hr"]))
    print(name, "W err", err, "f err", fe, flush=True)
    assert np.all(err <= TOL_FP64), err
    assert fe <= TOL_FP64, fe
    qe = q_error(r["q"], r["qr"])
    print(name, "q err", qe, "f err", fe)
    assert qe <= 1e-8, qe


@pytest.mark.parametrize("name,params,variant,steps", [
    ("poiseuille", poiseuille(), "reference", 20),
    ("rt_cartesian", rt(), "cartesian", 10),
])
def test_mixed_parity(have_gpu, name, params, variant, steps):
    case = cases.build_2d(params, capi.MIXED_FP32, rt_variant=variant, seed=3)
    r = run_pair(case, steps)
    err = field_err(r["w"], r["wr"])
    print(name, "mixed W err", err)
    assert np.all(err <= TOL_MIXED), err


def test_advance_until_output_time(have_gpu):
    """advance_until stops after the first step ending at or past t_stop,
    the reference main loop's write_macro points (2d/src/main.f90:88-99);
    the state there equals the oracle's after the same number of steps."""
    case = cases.build_2d(rt(), capi.FP64, rt_variant="cartesian", seed=3)
    o = oracle_py.Oracle(case)
    ts = []
    for _ in range(12):
        ts.append(o.step().simtime)
    t_stop = 0.5 * (ts[6] + ts[7])      # reached by step 8
    with capi.Solver(case) as s:
        info = s.advance_until(0, t_stop)
        assert info.stop_reason == capi.STOP_TIME and info.steps == 8
        w = s.macro(with_q=False)
        info2 = s.advance_until(0, ts[11])
        assert info2.stop_reason == capi.STOP_TIME and info2.steps == 4
    o2 = oracle_py.Oracle(case)
    o2.step(8)
    assert np.all(field_err(w, o2.get()[0]) <= TOL_FP64)


@pytest.mark.parametrize("mode", ["fused", "split", "split-nodes", "stencil", "cellsum"])
def test_flux_kernels_agree(have_gpu, monkeypatch, mode):
    """The interior-face flux implementations (one fused kernel; the
    five-kernel pipeline reading the slope arrays, with the Maxwellian sums
    from 1D moments of the tensor-product quadrature or node by node; the
    pipeline forming the slopes from the cell stencils; the pipeline with
    the face sums added from per-cell half-space sums) against the oracle,
    FP64 and mixed."""
    monkeypatch.setenv("HUGKS_FLUX_KERNEL", mode)
    for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
        case = cases.build_2d(rt(nx=10, ny=12, nv=16, nm=8), prec, rt_variant="cartesian", seed=5)
        r = run_pair(case, 6)
        err = field_err(r["w"], r["wr"])
        print(mode, prec, "W err", err, flush=True)
        assert np.all(err <= tol), err


@pytest.mark.parametrize("mode", ["lds", "mfma", "mfma-node", "mfma-face"])
def test_update_kernels_agree(have_gpu, monkeypatch, mode):
    """The cell-update implementations (LDS dot products; MFMA Hermite
    tiles; MFMA node update; the node update forming the interior faces'
    equilibrium flux) against the oracle, FP64 and mixed, at a velocity
    grid that is not a multiple of 16 and Nm = 10."""
    monkeypatch.setenv("HUGKS_UPDATE_KERNEL", mode)
    for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
        case = cases.build_2d(rt(nx=8, ny=10, nv=20, nm=10), prec, rt_variant="cartesian", seed=7)
        r = run_pair(case, 6)
        err = field_err(r["w"], r["wr"])
        print(mode, prec, "W err", err, flush=True)
        assert np.all(err <= tol), err


def test_flux_weights_not_a_tensor_product(have_gpu):
    """Weights perturbed by 1e-3 so that they no longer factor as
    wx[i] wy[j]: the pipeline must take the node-by-node Maxwellian sums
    (the 1D-moment form would be off by about 1e-4)."""
    case = cases.build_2d(rt(nx=8, ny=10, nv=14, nm=8), capi.FP64, rt_variant="cartesian", seed=6)
    w = case.arrays["weight"]
    w *= 1.0 + 1e-3 * np.random.default_rng(3).uniform(-1.0, 1.0, w.shape)
    r = run_pair(case, 6)
    err = field_err(r["w"], r["wr"])
    print("perturbed weights W err", err, flush=True)
    assert np.all(err <= TOL_FP64), err


@pytest.mark.parametrize("mode", ["split", "stencil", "fused", "cellsum"])
def test_odd_velocity_grids(have_gpu, monkeypatch, mode):
    """nv^2 odd (the face kernels read node pairs, the last pair ending on
    the first pad node) and nv = 14 (a pair of nodes on both sides of the
    upwind switch), FP64 and mixed, against the oracle."""
    monkeypatch.setenv("HUGKS_FLUX_KERNEL", mode)
    for nv in (13, 14, 15):
        for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
            case = cases.build_2d(rt(nx=6, ny=8, nv=nv, nm=8), prec, rt_variant="cartesian", seed=nv)
            r = run_pair(case, 4)
            err = field_err(r["w"], r["wr"])
            print(mode, "nv", nv, "prec", prec, "W err", err, flush=True)
            assert np.all(err <= tol), err


@pytest.mark.parametrize("nv", [24, 28, 32])
def test_compiled_in_grid_sizes(have_gpu, nv):
    """The update kernel instances compiled for common Gauss-Hermite grids
    (nv = 24, 28, 32 here; 16 and 20 in test_flux_kernels_agree and
    test_update_kernels_agree, 40 in test_gpu_headline_grid.py), FP64 and
    mixed, against the oracle."""
    for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
        case = cases.build_2d(poiseuille(nx=4, ny=6, nv=nv, nm=8), prec)
        r = run_pair(case, 3)
        err = field_err(r["w"], r["wr"])
        print("nv", nv, "prec", prec, "W err", err, flush=True)
        assert np.all(err <= tol), err


def test_cellsum_refuses_unsplit_nodes(have_gpu):
    """The cell-sum schedule needs the v nodes negative before
    non-negative (its rows are taken in those two passes): a case whose v
    list is reversed is refused with HUGKS_E_UNSUPPORTED, and the other
    schedules still run it to the oracle."""
    case = cases.build_2d(rt(nx=6, ny=8, nv=12, nm=6), capi.FP64, rt_variant="cartesian", seed=4)
    nv = case.desc.nv
    case.arrays["v"][:] = case.arrays["v"][::-1].copy()
    w = case.arrays["weight"].reshape(nv, nv)
    case.arrays["weight"][:] = w[::-1, :].reshape(-1).copy()
    with capi.Solver(case) as s:
        with pytest.raises(capi.HugksError) as ei:
            s.set_schedule(capi.FLUX_CELLSUM, capi.UPDATE_MFMA_NODE)
        assert ei.value.code == capi.E_UNSUPPORTED
    r = run_pair(case, 3)
    err = field_err(r["w"], r["wr"])
    assert np.all(err <= TOL_FP64), err
