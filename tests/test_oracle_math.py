"""The oracle's function-level pieces against what mathematics fixes.

Restates the reference's own unit tests (2d/tests/*.f90, 1d/tests/*.f90)
on the C oracle, plus closed forms that the reference's tests leave open:
the Gauss-Hermite rule against numpy's, and the interface flux of a
uniform equilibrium state against the Euler flux of that state.
"""
import ctypes as C
import math

import numpy as np
import pytest

from hugks import capi, cases, namelist
from hugks.quadrature import PI_REF, gauss_hermite_1d, gauss_hermite_2d
import oracle_py
from oracle_py import ptr

RC = 0.5  # gas constant of the reference's units: lambda = 1/(2 Rc T)


def desc_2d(nv, nm=6, force_type="hermite_symmetric"):
    """A 2D descriptor on an nv x nv Gauss-Hermite grid (Tref = 1)."""
    p = dict(namelist.DEFAULTS_2D)
    p.update(nx=2, ny=2, nv=nv, Nm=nm, force_type=force_type, case_id=6, phi=-0.1)
    case = cases.build_2d(p)
    return case, case.bind()


def grid_2d(case):
    nv = case.desc.nv
    u = np.tile(case.arrays["u"], nv)          # u fastest
    v = np.repeat(case.arrays["v"], nv)
    return u, v, case.arrays["weight"]


def maxwell_2d(u, v, prim, innerk=1.0):
    """reduced_maxwell of 2d/src/gastheory2d.f90:6-13 (input generator)."""
    rho, U, V, lam = prim
    h = rho * (lam / PI_REF) * np.exp(-lam * ((u - U) ** 2 + (v - V) ** 2))
    return h, h * innerk / (2.0 * lam)


def force_2d(case, d, variant, order, h, b, prim, gx, gy):
    lib = oracle_py.load()
    nn = case.nn
    fh, fb = np.empty(nn), np.empty(nn)
    eh, eb = C.c_double(), C.c_double()
    prim = np.asarray(prim, dtype=np.float64)
    lib.oracle2d_hermite_force(C.addressof(d), variant, order, ptr(h), ptr(b), ptr(prim), gx, gy,
                               ptr(fh), ptr(fb), C.cast(C.byref(eh), C.POINTER(C.c_double)),
                               C.cast(C.byref(eb), C.POINTER(C.c_double)))
    return fh, fb, eh.value, eb.value


# --------------------------------------------------------------------------
# velocity quadrature
# --------------------------------------------------------------------------

@pytest.mark.parametrize("nv", [22, 34, 36])
def test_gauss_hermite_1d_properties(nv):
    """1d/tests/test_gauss_hermite.f90: ordered, symmetric, positive, and
    integrates the unit Maxwellian's mass and second moment."""
    u, w, _, _ = gauss_hermite_1d(nv, 1.0)
    assert np.all(np.diff(u) > 0)
    assert np.max(np.abs(u + u[::-1])) < 1e-13
    assert np.all(w > 0)
    g = np.exp(-u ** 2 / 2.0) / math.sqrt(2.0 * PI_REF)
    assert abs(np.sum(w * g) - 1.0) < 5e-10
    assert abs(np.sum(w * u ** 2 * g) - 1.0) < 2e-9


@pytest.mark.parametrize("nv", [8, 16, 20, 28, 40])
def test_gauss_hermite_matches_numpy(nv):
    """The Newton-iterated rule equals numpy's eigenvalue-based hermgauss."""
    x, wx = np.polynomial.hermite.hermgauss(nv)
    u, w, roots, wgh = gauss_hermite_1d(nv, 1.0)
    assert np.max(np.abs(roots - x)) < 1e-12 * max(1.0, np.max(np.abs(x)))
    # the rule's weight normalisation carries the reference's PI: pi**-0.5
    # enters squared through the recurrence's start value
    ratio = math.sqrt(PI_REF / math.pi)
    np.testing.assert_allclose(wgh, wx * ratio, rtol=1e-10)


@pytest.mark.parametrize("nv", [16, 18, 20])
def test_gauss_hermite_2d_moments(nv):
    """2d/tests/test_velocity_force.f90:20-44."""
    u1, v1, w = gauss_hermite_2d(nv, 1.0)
    assert np.all(np.diff(u1) > 0)
    assert np.max(np.abs(u1 + u1[::-1])) < 2e-13
    assert np.max(np.abs(v1 + v1[::-1])) < 2e-13
    assert np.all(w > 0)
    u, v = np.tile(u1, nv), np.repeat(v1, nv)
    g = np.exp(-(u ** 2 + v ** 2) / 2.0) / (2.0 * PI_REF)
    tol = 1e-6
    assert abs(np.sum(w * g) - 1.0) < tol
    assert abs(np.sum(w * u * g)) < tol and abs(np.sum(w * v * g)) < tol
    assert abs(np.sum(w * u * u * g) - 1.0) < tol and abs(np.sum(w * v * v * g) - 1.0) < tol


# --------------------------------------------------------------------------
# Hermite force
# --------------------------------------------------------------------------

def manufactured(case):
    """Rank-0..3 Hermite expansion of 2d/tests/test_symmetric_hermite.f90:41-57."""
    u, v, w = grid_2d(case)
    rt = RC / 1.0
    ux, uy = u / math.sqrt(rt), v / math.sqrt(rt)
    om = np.exp(-0.5 * (ux ** 2 + uy ** 2)) / (2.0 * PI_REF)
    He = {
        "x2": ux ** 2 - 1, "x3": ux ** 3 - 3 * ux, "x4": ux ** 4 - 6 * ux ** 2 + 3,
        "y2": uy ** 2 - 1, "y3": uy ** 3 - 3 * uy, "y4": uy ** 4 - 6 * uy ** 2 + 3,
    }
    h = om * (1.1 + 0.14 * He["x2"] / 2 + 0.07 * ux * uy + 0.03 * He["y3"] / 6)
    b = om * (0.8 - 0.05 * ux + 0.09 * He["y2"] / 2 + 0.04 * He["x3"] / 6)
    return h, b, ux, uy, om, He, rt


@pytest.mark.parametrize("variant", [capi.FORCE_HERMITE_SYMMETRIC, capi.FORCE_HERMITE_SYMMETRIC_B])
def test_hermite_force_manufactured_rank3(variant):
    """2d/tests/test_symmetric_hermite.f90:59-91: the force a.grad_v f of a
    rank-3 expansion is, term by term, the next Hermite polynomial up
    (d/dx [om He_n] = -om He_{n+1} / sqrt(RT)); the closed form below is
    that identity applied to each term, not the operator's algorithm."""
    case, d = desc_2d(20, nm=3)
    h, b, ux, uy, om, He, rt = manufactured(case)
    gx, gy = 0.13, -0.09
    fh, fb, eh, eb = force_2d(case, d, variant, 3, h, b, [1.0, 0.0, 0.0, 1.0], gx, gy)
    s = -om / math.sqrt(rt)
    exp_h = s * (1.1 * (gx * ux + gy * uy)
                 + 0.14 / 2 * (gx * He["x3"] + gy * He["x2"] * uy)
                 + 0.07 * (gx * He["x2"] * uy + gy * ux * He["y2"])
                 + 0.03 / 6 * (gx * ux * He["y3"] + gy * He["y4"]))
    exp_b = s * (0.8 * (gx * ux + gy * uy)
                 - 0.05 * (gx * He["x2"] + gy * ux * uy)
                 + 0.09 / 2 * (gx * ux * He["y2"] + gy * He["y3"])
                 + 0.04 / 6 * (gx * He["x4"] + gy * He["x3"] * uy))
    assert np.max(np.abs(fh - exp_h)) < 1e-9
    assert np.max(np.abs(fb - exp_b)) < 1e-9
    assert eh < 1e-9 and eb < 1e-9


def test_hermite_force_closed_form_separates_mistakes():
    """The closed form is sharp enough to catch swapped acceleration
    components: they miss it by far more than the 1e-9 bar."""
    case, d = desc_2d(20, nm=3)
    h, b, *_ = manufactured(case)
    prim = [1.0, 0, 0, 1.0]
    f = force_2d(case, d, capi.FORCE_HERMITE_SYMMETRIC, 3, h, b, prim, 0.13, -0.09)[0]
    swapped = force_2d(case, d, capi.FORCE_HERMITE_SYMMETRIC, 3, h, b, prim, -0.09, 0.13)[0]
    assert np.max(np.abs(swapped - f)) > 1e-3


@pytest.mark.parametrize("a", [(0.1, 0.0), (-0.08, 0.06)])
def test_hermite_force_variants_agree_and_conserve(a):
    """2d/tests/test_symmetric_hermite.f90:93-135 and
    test_velocity_force.f90:66-103: symmetric, symmetric-b and legacy
    operators agree on a shifted Maxwellian, and the force moments are
    (0, -rho a, -rho U.a)."""
    ax, ay = a
    case, d = desc_2d(20, nm=6)
    u, v, w = grid_2d(case)
    prim = [1.2, 0.23, -0.17, 1.0]
    h, b = maxwell_2d(u, v, prim)
    res = {var: force_2d(case, d, var, 6, h, b, prim, ax, ay)
           for var in (capi.FORCE_HERMITE_SYMMETRIC, capi.FORCE_HERMITE_SYMMETRIC_B,
                       capi.FORCE_HERMITE_LEGACY)}
    sh, sb = res[capi.FORCE_HERMITE_SYMMETRIC][:2]
    bh, bb = res[capi.FORCE_HERMITE_SYMMETRIC_B][:2]
    lh, lb = res[capi.FORCE_HERMITE_LEGACY][:2]
    assert np.max(np.abs(bh - sh)) < 1e-12 and np.max(np.abs(bb - sb)) < 1e-12
    assert np.linalg.norm(sh - lh) / np.linalg.norm(lh) < 1e-12
    assert np.linalg.norm(sb - lb) / np.linalg.norm(lb) < 1e-12
    for fh, fb in ((sh, sb), (lh, lb)):
        s0 = np.sum(w * fh)
        sx = np.sum(w * u * fh) + prim[0] * ax
        sy = np.sum(w * v * fh) + prim[0] * ay
        se = 0.5 * np.sum(w * ((u ** 2 + v ** 2) * fh + fb)) + prim[0] * (prim[1] * ax + prim[2] * ay)
        assert abs(s0) < 2e-9
        assert abs(sx) < 2e-8 and abs(sy) < 2e-8
        assert abs(se) < 5e-8


def test_hermite_force_1d_moments():
    """1d/tests/test_force_moments.f90: nv = 34 GH, Nm = 6, prim = (1, 0.3, 0.5)."""
    p = dict(namelist.DEFAULTS_1D)
    p.update(nx=4, nv=34, Nm=6, quad_type="gauss_hermite", case_id=3)
    case = cases.build_1d(p)
    d = case.bind()
    u, w = case.arrays["u"], case.arrays["weight"]
    prim = np.array([1.0, 0.3, 0.5])
    h = prim[0] * math.sqrt(prim[2] / PI_REF) * np.exp(-prim[2] * (u - prim[1]) ** 2)
    b = h * d.innerk / (2.0 * prim[2])
    dvh, dvb = np.empty_like(u), np.empty_like(u)
    oracle_py.load().oracle1d_hermite_force(C.addressof(d), 6, ptr(h), ptr(b), ptr(prim),
                                            ptr(dvh), ptr(dvb))
    assert abs(np.sum(w * dvh)) < 1e-10
    assert abs(np.sum(w * u * dvh) + prim[0]) < 1e-9
    assert abs(0.5 * np.sum(w * (u ** 2 * dvh + dvb)) + prim[0] * prim[1]) < 1e-9


# --------------------------------------------------------------------------
# wall fluxes
# --------------------------------------------------------------------------

RN, RY = 1, -1  # 2d/src/data2d.f90:15-16


def wall_flux(case, d, kind, direction, rot, tw, h, b):
    nn = case.nn
    sh = np.zeros(2 * nn)
    sb = np.zeros(2 * nn)
    flux, fh, fb = np.empty(4), np.empty(nn), np.empty(nn)
    oracle_py.load().oracle2d_flux_wall(C.addressof(d), kind, direction, rot, tw, 1.0, 1.0,
                                        ptr(h), ptr(b), ptr(sh), ptr(sb), 1.0,
                                        ptr(flux), ptr(fh), ptr(fb))
    return flux, fh, fb


def wall_cell(case):
    u, v, w = grid_2d(case)
    h, b = maxwell_2d(u, v, [1.0, 0.17, -0.11, 1.0])
    return h * (1 + 0.04 * u + 0.03 * v), b * (1 - 0.02 * u + 0.01 * v), u, v


def test_specular_walls_half_range():
    """2d/tests/test_boundaries.f90:32-76: zero mass flux, the outgoing half
    range kept, the incoming one mirrored."""
    case, d = desc_2d(16)
    nv = d.nv
    h, b, u, v = wall_cell(case)
    H = h.reshape(nv, nv)                                   # [jv, iu]
    for direction, vel in ((1, u), (2, v)):
        flux, fh, _ = wall_flux(case, d, capi.BC_SPECULAR, direction, RN, 1.0, h, b)
        assert abs(flux[0]) < 2e-11
        rec = (fh / vel).reshape(nv, nv)
        mirror = H[:, ::-1] if direction == 1 else H[::-1, :]
        out = (vel < 0).reshape(nv, nv)
        assert np.max(np.abs(rec[out] - H[out])) < 2e-12
        assert np.max(np.abs(rec[~out] - mirror[~out])) < 2e-12
        flux, _, _ = wall_flux(case, d, capi.BC_SPECULAR, direction, RY, 1.0, h, b)
        assert abs(flux[0]) < 2e-11


@pytest.mark.parametrize("rot", [RN, RY])
def test_diffuse_wall_zero_mass_flux(rot):
    """2d/tests/test_boundaries.f90:78-84 (and 1d/tests/test_diffuse_wall.f90)."""
    case, d = desc_2d(16)
    h, b, u, v = wall_cell(case)
    for direction in (1, 2):
        flux, _, _ = wall_flux(case, d, capi.BC_DIFFUSE, direction, rot, 1.0, h, b)
        assert abs(flux[0]) < 2e-11


def test_diffuse_wall_reemits_wall_maxwellian():
    """The incoming half range at a diffuse wall is a Maxwellian at the
    wall's temperature with zero mean velocity: fh/vn is proportional to
    exp(-|c|^2 / Tw) there, whatever the gas state."""
    case, d = desc_2d(16)
    h, b, u, v = wall_cell(case)
    tw = 1.3
    _, fh, _ = wall_flux(case, d, capi.BC_DIFFUSE, 2, RN, tw, h, b)
    inc = v > 0
    ratio = fh[inc] / v[inc] / np.exp(-(u[inc] ** 2 + v[inc] ** 2) / tw)
    assert np.max(np.abs(ratio / ratio[0] - 1.0)) < 1e-12


# --------------------------------------------------------------------------
# interior interface flux
# --------------------------------------------------------------------------

@pytest.mark.parametrize("direction", [1, 2])
@pytest.mark.parametrize("prim", [(1.0, 0.0, 0.0, 1.0), (1.2, 0.23, -0.17, 0.8)])
def test_interior_flux_of_uniform_equilibrium(direction, prim):
    """With the same Maxwellian on both sides and zero slopes the UGKS flux
    over dt is dt times the Euler flux of that state, for any tau: mass
    rho Un, momentum rho Un U + p n, energy Un (rho E + p), with
    p = rho / (2 lambda) and rho E = rho |U|^2 / 2 + rho (K + 2) / (4 lambda).
    Per node the flux is dt vn M."""
    case, d = desc_2d(28)
    u, v, w = grid_2d(case)
    h, b = maxwell_2d(u, v, prim, d.innerk)
    nn = case.nn
    z = np.zeros(2 * nn)
    wc = np.array([np.sum(w * h), np.sum(w * u * h), np.sum(w * v * h),
                   0.5 * np.sum(w * ((u ** 2 + v ** 2) * h + b))])
    lens = np.array([0.05, 0.05])
    dt = 1e-3
    flux, fh, fb = np.empty(4), np.empty(nn), np.empty(nn)
    oracle_py.load().oracle2d_flux_interior(
        C.addressof(d), direction, dt, 0.0, 0.0, 1.0,
        ptr(h), ptr(b), ptr(z), ptr(z), ptr(wc), ptr(lens),
        ptr(h), ptr(b), ptr(z), ptr(z), ptr(wc), ptr(lens),
        ptr(flux), ptr(fh), ptr(fb))
    rho, U, V, lam = prim
    p = rho / (2.0 * lam)
    rhoE = 0.5 * rho * (U * U + V * V) + rho * (d.innerk + 2.0) / (4.0 * lam)
    Un = U if direction == 1 else V
    nx, ny = (1.0, 0.0) if direction == 1 else (0.0, 1.0)
    euler = dt * np.array([rho * Un, rho * Un * U + p * nx, rho * Un * V + p * ny,
                           Un * (rhoE + p)])
    np.testing.assert_allclose(flux, euler, rtol=0, atol=1e-12 * dt * max(1.0, rhoE))
    vn = u if direction == 1 else v
    scale = np.max(np.abs(vn * h))
    assert np.max(np.abs(fh - dt * vn * h)) <= 1e-12 * dt * scale
    assert np.max(np.abs(fb - dt * vn * b)) <= 1e-12 * dt * scale
