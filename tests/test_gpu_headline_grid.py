"""The headline velocity grid (40 x 40 Gauss-Hermite nodes) and Hermite
order (Nm = 12) against the FP64 oracle, on small grids.

The large benchmark case runs exactly these kernel instances: the update
kernels compiled for nv = 40 (FP32 storage: 7 nodes per thread, 256
threads; FP64: 5 nodes per thread, 320 threads), the MFMA Hermite tiles
with r + s up to 12 in the 16-row tiles, and the interior-face pipeline at
1600 nodes per face.  Each flux schedule (split, stencil, fused, cell-sum)
is run with each update kernel (MFMA Hermite tiles, MFMA node update, LDS),
and the cell-sum and split fluxes also with the node update that forms the
equilibrium part of the interior faces' node fluxes itself (the rest as a
per-cell divergence in place of the pipeline's per-node kernel):

- after 6 steps every conserved field agrees with the oracle to 1e-10
  (FP64) and 1e-5 (mixed) of the field's scale, the distributions to
  1e-10 (FP64);
- after one step the intermediate arrays of the step (van Leer slopes,
  per-node face fluxes and face conserved fluxes of every interior face)
  agree with the oracle's per-face routines (2d/src/solver.f90:64-159,
  2d/src/flux.f90:28-472) evaluated on the GPU's own step-start state.

Two cases: the Cartesian Rayleigh-Taylor set-up (specular walls, uniform
gravity, Shakhov with Pr = 2/3) and the force-driven channel (periodic in
x, diffuse walls of unequal temperature).
"""
import numpy as np
import pytest

from hugks import capi, cases
import kernels_ref
import oracle_py
from metrics import field_errors

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-10
TOL_MIXED = 1e-5
STEPS = 6
FLUX = {"split": capi.FLUX_SPLIT, "stencil": capi.FLUX_STENCIL, "fused": capi.FLUX_FUSED,
        "cellsum": capi.FLUX_CELLSUM}
UPDATE = {"mfma": capi.UPDATE_MFMA, "lds": capi.UPDATE_LDS, "mfma-node": capi.UPDATE_MFMA_NODE,
          "mfma-face": capi.UPDATE_MFMA_FACE}
# every flux schedule with every update; the face-forming update with the
# schedules that store the slope arrays and the per-face coefficients
SCHED = [(f, u) for f in sorted(FLUX) for u in sorted(UPDATE)
         if u != "mfma-face" or f in ("cellsum", "split")]


def rt_params():
    return dict(nx=8, ny=10, xlength=0.5, ylength=1.0, nv=40, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=12, cfl=0.5,
                maxtime=0.14, case_id=6, knudsen=1e-3, phi=1.0, TwL=1.0, TwR=1.0)


def channel_params():
    return dict(nx=4, ny=8, xlength=0.1, ylength=1.0, nv=40, vmax=5.0, Tref=1.0,
                quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=12, cfl=0.8,
                maxtime=200.0, case_id=5, knudsen=0.1, phi=0.05, TwL=1.0, TwR=1.3)


CASES = {
    "rt": lambda prec: cases.build_2d(rt_params(), prec, rt_variant="cartesian", seed=11),
    "channel": lambda prec: cases.build_2d(channel_params(), prec),
}
_oracle_cache = {}


def oracle_after(name, steps):
    key = (name, steps)
    if key not in _oracle_cache:
        o = oracle_py.Oracle(CASES[name](capi.FP64))
        o.step(steps)
        w, h, b = o.get()
        _oracle_cache[key] = (w, h, b, o.time())
    return _oracle_cache[key]


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flux,update", SCHED)
def test_headline_grid_steps(have_gpu, name, flux, update):
    wr, hr, br, (t, it, _) = oracle_after(name, STEPS)
    for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
        case = CASES[name](prec)
        print(f"{name} nv=40 Nm=12 flux={flux} update={update} prec={prec} steps={STEPS}",
              flush=True)
        with capi.Solver(case) as s:
            s.set_schedule(FLUX[flux], UPDATE[update])
            assert s.schedule() == (FLUX[flux], UPDATE[update])
            info = s.advance(STEPS)
            w = s.macro(with_q=False)
            hf, bf = s.get_f()
        assert info.steps == STEPS and info.iter == it
        # dt follows W: exact in FP64, within the field error when mixed
        assert abs(info.simtime - t) <= (1e-12 if prec == capi.FP64 else tol) * t
        err = field_errors(w, wr)
        fe = max(np.max(np.abs(hf - hr)) / np.max(np.abs(hr)),
                 np.max(np.abs(bf - br)) / np.max(np.abs(br)))
        print(f"  W err {err} f err {fe:.3e}", flush=True)
        assert np.all(err <= tol), err
        if prec == capi.FP64:
            assert fe <= TOL_FP64, fe


def _interior_faces(case):
    """(dir, face index, left cell, right cell) of every interior face;
    dir 1 = vertical (x-normal), 2 = horizontal."""
    d = case.desc
    nx, ny = d.nx, d.ny
    xper, yper = d.bc[0] == capi.BC_PERIODIC, d.bc[2] == capi.BC_PERIODIC
    out = []
    for j in range(ny):
        for i in (range(nx + 1) if xper else range(1, nx)):
            out.append((1, i + (nx + 1) * j, (i - 1) % nx + nx * j, i % nx + nx * j))
    for j in (range(ny + 1) if yper else range(1, ny)):
        for i in range(nx):
            out.append((2, i + nx * j, i + nx * ((j - 1) % ny), i + nx * (j % ny)))
    return out


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("flux", sorted(FLUX))
def test_headline_grid_intermediates(have_gpu, name, flux):
    for prec, tol in ((capi.FP64, TOL_FP64), (capi.MIXED_FP32, TOL_MIXED)):
        case = CASES[name](prec)
        print(f"{name} intermediates nv=40 Nm=12 flux={flux} prec={prec}", flush=True)
        with capi.Solver(case) as s:
            s.set_schedule(FLUX[flux], capi.UPDATE_MFMA)
            s.advance(2)          # a state that is no longer the initial Maxwellian
            w0 = s.macro(with_q=False)
            h0, b0 = s.get_f()
            info = s.step()
            dt = info.dt
            got = {a: s.debug_fetch(a) for a in (capi.ARR_SHX, capi.ARR_SHY, capi.ARR_SBX,
                                                 capi.ARR_SBY, capi.ARR_VFH, capi.ARR_VFB,
                                                 capi.ARR_HFH, capi.ARR_HFB, capi.ARR_VFLUX,
                                                 capi.ARR_HFLUX)}
        shx, shy = kernels_ref.slopes_2d(case, h0)
        sbx, sby = kernels_ref.slopes_2d(case, b0)
        rel = lambda a, r: np.max(np.abs(a - r)) / max(np.max(np.abs(r)), 1e-300)
        if flux != "stencil":   # the stencil schedule forms slopes inside the face kernels
            for a, ref in ((capi.ARR_SHX, shx), (capi.ARR_SHY, shy), (capi.ARR_SBX, sbx),
                           (capi.ARR_SBY, sby)):
                e = rel(got[a], ref)
                assert e <= tol, (a, e)
        acc_v = case.arrays["accel_vface"].reshape(-1, 2)
        acc_h = case.arrays["accel_hface"].reshape(-1, 2)
        # the slopes the oracle's face routine takes are the exact ones
        # (FP64) of the GPU's step-start f, in both precisions
        sh, sb = (shx, shy), (sbx, sby)
        ef = ec = 0.0
        sc_f = sc_c = 0.0
        for dir_, face, cl, cr in _interior_faces(case):
            acc = acc_v[face] if dir_ == 1 else acc_h[face]
            flux_r, fh_r, fb_r = kernels_ref.face_flux_interior(case, dir_, dt, cl, cr, acc, h0,
                                                                b0, sh, sb, w0)
            gh = got[capi.ARR_VFH if dir_ == 1 else capi.ARR_HFH][face]
            gb = got[capi.ARR_VFB if dir_ == 1 else capi.ARR_HFB][face]
            gc = got[capi.ARR_VFLUX if dir_ == 1 else capi.ARR_HFLUX][face]
            ef = max(ef, np.max(np.abs(gh - fh_r)), np.max(np.abs(gb - fb_r)))
            sc_f = max(sc_f, np.max(np.abs(fh_r)), np.max(np.abs(fb_r)))
            ec = max(ec, np.max(np.abs(gc - flux_r) / np.maximum(np.abs(flux_r), 1e-300 +
                                                                  np.max(np.abs(flux_r)))))
        print(f"  face flux err {ef / sc_f:.3e}, conserved face flux err {ec:.3e}", flush=True)
        assert ef / sc_f <= tol
        assert ec <= tol
