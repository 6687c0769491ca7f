"""Case builders: turn the reference's ``&params`` values into a library
descriptor (host-side initialisation that the reference does in
initialize(), 2d/src/initialization.f90:11-237 and
1d/src/initialization.f90:10-194).

Cases:
  2D case_id 5  force-driven Poiseuille flow, x-periodic, diffuse walls
  2D case_id 6  Rayleigh-Taylor; ``rt_variant='reference'`` is the
                reference's quarter-domain polar set-up,
                ``rt_variant='cartesian'`` the Cartesian heavy-over-light
                set-up with uniform gravity (0, -|phi|)
  1D case_id 2  hydrostatic equilibrium, 3 Sod shock tube, 4 Fourier flow
"""
import math

import numpy as np

from . import capi
from .quadrature import PI_REF, gauss_hermite_1d, gauss_hermite_2d, newton_cotes_1d, newton_cotes_2d

SHAKHOV_COEF = float(np.float32(0.8))   # default-REAL literal 0.8 (gastheory2d.f90:22)
OMEGA_VHS = 0.72                        # 2d/src/initialization.f90:130
EPS_2D = 1e-7                           # 2d/src/data2d.f90:11
SMV_2D = 1e-8                           # 2d/src/data2d.f90:10
SMV_1D = np.finfo(np.float64).tiny      # 1d/src/data1d.f90:10
PR_1D = 0.666666666667                  # 1d/src/data1d.f90:63


def get_mu(kn, alpha=1.0, omega=0.5, pi=PI_REF):
    """Hard-sphere reference viscosity, 2d/src/gastheory2d.f90:101-111."""
    return 5.0 * (alpha + 1.0) * (alpha + 2.0) * math.sqrt(pi) / (
        4 * alpha * (5 - 2 * omega) * (7 - 2 * omega)) * kn


def minstd(x):
    return (48271 * x) % 2147483647


def cell_uniform(seed, i, j):
    """Uniform (0,1) value of 1-based cell (i, j): Park-Miller MINSTD,
    three steps from a hashed start (same recipe as the oracle harness)."""
    x = (seed * 1000003 + j * 7919 + i * 104729 + 12345) % 2147483646 + 1
    for _ in range(3):
        x = minstd(x)
    return x / 2147483647.0


def _velocity_2d(p):
    nv, tref = int(p["nv"]), float(p["Tref"])
    if str(p["quad_type"]).strip() == "gauss_hermite":
        u, v, w = gauss_hermite_2d(nv, tref)
        umax, vmax = u[-1], v[-1]
        span = u[-1] - u[0]                      # lengthU, 2d/src/velocity.f90:131-134
    else:
        u, v, w = newton_cotes_2d(nv, float(p["vmax"]), tref)
        umax = vmax = float(p["vmax"]) * math.sqrt(tref)
        span = umax - (-umax)                    # lengthU, 2d/src/config.f90:73-80
    gap = min(np.min(np.diff(u)), np.min(np.diff(v)))   # 2d/src/solver.f90:53-54
    return u, v, w, umax, vmax, gap, span


def _force_variant(force_type):
    ft = str(force_type).strip()
    if ft == "hermite_symmetric":
        return capi.FORCE_HERMITE_SYMMETRIC
    if ft == "hermite_symmetric_b":
        return capi.FORCE_HERMITE_SYMMETRIC_B
    if ft == "hermite":
        return capi.FORCE_HERMITE_LEGACY
    if ft == "finite_difference":
        return capi.FORCE_FINITE_DIFFERENCE
    raise ValueError(f"unknown force_type '{ft}'")


def build_2d(p, precision=capi.FP64, rt_variant="reference", seed=0):
    p = dict(p)
    nx, ny, nv = int(p["nx"]), int(p["ny"]), int(p["nv"])
    case_id = int(p["case_id"])
    phi = float(p["phi"])
    xlength, ylength = float(p["xlength"]), float(p["ylength"])
    dx, dy = xlength / nx, ylength / ny          # 2d/src/initialization.f90:69-70
    u, v, w, umax, vmax, gap, span = _velocity_2d(p)
    kn = float(p["knudsen"])
    innerk = 1.0
    gamma = (innerk + 4.0) / (innerk + 2.0)      # 2d/src/gastheory2d.f90:70-76
    i1 = np.arange(1, nx + 1)
    j1 = np.arange(1, ny + 1)
    xc = (i1 - 0.5) * dx
    yc = (j1 - 0.5) * dy
    X, Y = np.meshgrid(xc, yc)                    # [ny, nx], x fastest when raveled
    prim = np.zeros((ny, nx, 4))
    d = capi.Desc()
    d.abi_version = capi.ABI_VERSION
    d.dim = 2
    d.nx, d.ny, d.nv, d.nm = nx, ny, nv, int(p["Nm"])
    d.precision = precision
    d.force_variant = _force_variant(p["force_type"])
    d.fd_span = span
    d.clamp_dt = 1
    d.dx, d.dy = dx, dy
    d.pi, d.smv, d.innerk, d.gamma = PI_REF, SMV_2D, innerk, gamma
    d.shakhov_coef, d.omega = SHAKHOV_COEF, OMEGA_VHS
    d.cfl, d.maxtime, d.eps_res = float(p["cfl"]), float(p["maxtime"]), EPS_2D
    d.umax_dt, d.vmax_dt = umax, vmax
    d.force_rate = abs(phi) / gap
    d.tref = float(p["Tref"])
    meta = {"case_id": case_id, "rt_variant": rt_variant, "seed": seed}
    if case_id == 5:
        # Poiseuille: x-periodic, diffuse isothermal walls in y
        # (2d/src/solver.f90:204-236, 2d/src/initialization.f90:172-186)
        d.bc[:] = [capi.BC_PERIODIC, capi.BC_PERIODIC, capi.BC_DIFFUSE, capi.BC_DIFFUSE]
        d.tw[:] = [1.0, 1.0, float(p["TwL"]), float(p["TwR"])]
        d.tau_model = capi.TAU_MU_OVER_RHO
        d.dt_rule = capi.DT_2D_WALL_NORMAL
        d.stop_on_residual = 1
        d.prandtl = 1.0
        d.muref = 0.5 * math.sqrt(PI_REF) * kn
        prim[..., 0] = 1.0
        prim[..., 3] = 1.0 / float(p["Tref"])
        ax = np.full((ny, nx), phi)
        ay = np.zeros((ny, nx))
        vtheta = htheta = None
        ax_v, ay_v = np.full((ny, nx + 1), phi), np.zeros((ny, nx + 1))
        ax_h, ay_h = np.full((ny + 1, nx), phi), np.zeros((ny + 1, nx))
    elif case_id == 6:
        # specular walls on all sides (2d/src/solver.f90:171-201)
        d.bc[:] = [capi.BC_SPECULAR] * 4
        d.tw[:] = [1.0] * 4
        d.tau_model = capi.TAU_VHS
        d.dt_rule = capi.DT_2D_GENERAL
        d.stop_on_residual = 0
        d.prandtl = 2.0 / 3.0
        d.muref = get_mu(kn)
        if rt_variant == "reference":
            # polar initial condition, 2d/src/initialization.f90:146-170
            r = np.sqrt(X ** 2 + Y ** 2)
            th = np.arctan2(Y, X)
            radius = 0.6 * (1.0 + 0.02 * np.cos(20 * th))
            inner = r <= radius
            prim[..., 0] = np.where(inner, np.exp(-2.68 * (r + 0.258)), np.exp(-5.53 * (r - 0.308)))
            pin = r <= 0.62324965
            prim[..., 3] = np.where(
                pin, 0.5 * prim[..., 0] / 1.5 / np.exp(-2.68 * (r + 0.258)) * 2.68,
                0.5 * prim[..., 0] / 1.5 / np.exp(-5.53 * (r - 0.308)) * 5.53)
            ctheta = th
            # face angles, 2d/src/initialization.f90:86-114
            vtheta = np.empty((ny, nx + 1))
            vtheta[:, :nx] = np.arctan2(Y, X - 0.5 * dx)
            vtheta[:, nx] = np.arctan2(Y[:, -1], X[:, -1] + 0.5 * dx)
            htheta = np.empty((ny + 1, nx))
            htheta[:ny, :] = np.arctan2(Y - 0.5 * dy, X)
            htheta[ny, :] = np.arctan2(Y[-1, :] + 0.5 * dy, X[-1, :])
        elif rt_variant == "cartesian":
            # heavy isothermal layer (T=1) over light isothermal layer
            # (T=2), density ratio 2 at y_i(x) = ylength/2 + 0.01 cos(2 pi x /
            # xlength), gravity (0, -|phi|), y-velocity noise +-1e-4 per cell
            g = abs(phi)
            phi = g
            th_, tl_ = 1.0, 2.0
            yi = 0.5 * ylength + 0.01 * np.cos(2.0 * math.pi * X / xlength)
            heavy = Y >= yi
            prim[..., 0] = np.where(heavy, np.exp(-2.0 * g * (Y - yi) / th_),
                                    0.5 * np.exp(2.0 * g * (yi - Y) / tl_))
            prim[..., 3] = np.where(heavy, 1.0 / th_, 1.0 / tl_)
            noise = np.array([[cell_uniform(seed, i, j) for i in range(1, nx + 1)]
                              for j in range(1, ny + 1)])
            prim[..., 2] = 1e-4 * (2.0 * noise - 1.0)
            down = math.atan2(-1.0, 0.0)
            ctheta = np.full((ny, nx), down)
            vtheta = np.full((ny, nx + 1), down)
            htheta = np.full((ny + 1, nx), down)
        else:
            raise ValueError(f"unknown rt_variant {rt_variant}")
        # cell acceleration (2d/src/solver.f90:291-297), face acceleration
        # phi*(cos, sin) of the face angle (2d/src/flux.f90:21-22)
        ax, ay = phi * np.cos(ctheta), phi * np.sin(ctheta)
        ax_v, ay_v = phi * np.cos(vtheta), phi * np.sin(vtheta)
        ax_h, ay_h = phi * np.cos(htheta), phi * np.sin(htheta)
    else:
        raise ValueError(f"2D case_id {case_id} is not one of the reference's cases (5, 6)")
    # projection onto the face normal/tangent, 2d/src/flux.f90:24-25:
    # vertical faces (cosa, sina) = (1, 0), horizontal (0, 1)
    fn_v = ax_v * 1.0 + ay_v * 0.0
    ft_v = ay_v * 1.0 - ax_v * 0.0
    fn_h = ax_h * 0.0 + ay_h * 1.0
    ft_h = ay_h * 0.0 - ax_h * 1.0
    arrays = {
        "u": np.ascontiguousarray(u), "v": np.ascontiguousarray(v),
        "weight": np.ascontiguousarray(w),
        "prim0": np.ascontiguousarray(prim.reshape(-1)),
        "accel_cell": np.ascontiguousarray(np.stack([ax, ay], -1).reshape(-1)),
        "accel_vface": np.ascontiguousarray(np.stack([fn_v, ft_v], -1).reshape(-1)),
        "accel_hface": np.ascontiguousarray(np.stack([fn_h, ft_h], -1).reshape(-1)),
    }
    meta.update(xc=xc, yc=yc, params=p)
    return capi.Case(d, arrays, meta)


def build_1d(p, precision=capi.FP64):
    p = dict(p)
    nx, nv = int(p["nx"]), int(p["nv"])
    case_id = int(p["case_id"])
    phi = float(p["phi"])
    tref = float(p["Tref"])
    dx = float(p["xlength"]) / nx                  # 1d/src/initialization.f90:68
    if str(p["quad_type"]).strip() == "gauss_hermite":
        u, w, _, _ = gauss_hermite_1d(nv, tref)
        umin, umax = u[0], u[-1]
        du = (umax - umin) / (nv - 1)              # 1d/src/velocity.f90:146
        span = umax - umin                         # lengthU, 1d/src/velocity.f90:147
    else:
        u, w = newton_cotes_1d(nv, float(p["vmax"]), tref)
        umax = float(p["vmax"]) * math.sqrt(tref)
        du = (2 * umax) / (nv - 1)
        span = umax - (-umax)                      # lengthU, 1d/src/config.f90:75-78
    innerk = 2.0
    gamma = (innerk + 3.0) / (innerk + 1.0)        # 1d/src/gastheory.f90:60-66
    x = (np.arange(1, nx + 1) - 0.5) * dx
    prim = np.zeros((nx, 3))
    d = capi.Desc()
    d.abi_version = capi.ABI_VERSION
    d.dim = 1
    d.nx, d.ny, d.nv, d.nm = nx, 1, nv, int(p["Nm"])
    d.precision = precision
    ft = str(p["force_type"]).strip()
    if ft not in ("hermite", "finite_difference"):         # 1d/src/solver.f90:233-242
        raise ValueError(f"unknown 1D force_type '{ft}'")
    d.force_variant = (capi.FORCE_HERMITE_LEGACY if ft == "hermite"
                       else capi.FORCE_FINITE_DIFFERENCE)
    d.fd_span = span
    d.clamp_dt = 0
    d.dx, d.dy = dx, 1.0
    d.pi, d.smv, d.innerk, d.gamma = PI_REF, SMV_1D, innerk, gamma
    d.prandtl, d.shakhov_coef, d.omega = PR_1D, SHAKHOV_COEF, OMEGA_VHS
    d.muref = get_mu(float(p["knudsen"]))
    d.tau_model = capi.TAU_VHS
    d.dt_rule = capi.DT_1D
    d.cfl, d.maxtime, d.eps_res = float(p["cfl"]), float(p["maxtime"]), 1e-6
    d.umax_dt, d.vmax_dt = umax, umax
    d.force_rate = abs(phi) / du
    d.tref = tref
    d.phi = phi
    if case_id == 2:
        prim[:, 0] = np.exp(-x)
        prim[:, 2] = 0.5 * prim[:, 0] / (np.exp(-x) + 0.01 * np.exp(-100.0 * (x - 0.5) ** 2))
        d.bc[:] = [capi.BC_SPECULAR, capi.BC_SPECULAR, 0, 0]
    elif case_id == 3:
        half = nx // 2
        prim[:half] = [1.0, 0.0, 0.5]
        prim[half:] = [0.125, 0.0, 0.625]
        d.bc[:] = [capi.BC_SPECULAR, capi.BC_SPECULAR, 0, 0]
    elif case_id == 4:
        prim[:, 0] = 1.0
        prim[:, 2] = 1.0 / (float(p["TwL"]) + x * (float(p["TwR"]) - float(p["TwL"])))
        d.bc[:] = [capi.BC_DIFFUSE, capi.BC_DIFFUSE, 0, 0]
        # steady-state exit of the Fourier case, 1d/src/main.f90:108-126
        d.steady_steps = int(p["steady_steps"])
        d.steady_tol, d.mass_tol = float(p["steady_tol"]), float(p["mass_tol"])
    else:
        raise ValueError(f"1D case_id {case_id} is not one of the reference's cases (2, 3, 4)")
    d.tw[:] = [float(p["TwL"]), float(p["TwR"]), 1.0, 1.0]
    arrays = {"u": np.ascontiguousarray(u), "weight": np.ascontiguousarray(w),
              "prim0": np.ascontiguousarray(prim.reshape(-1))}
    return capi.Case(d, arrays, {"case_id": case_id, "xc": x, "params": p})


PRECISIONS = {"fp64": capi.FP64, "mixed": capi.MIXED_FP32, "mixed_fp32": capi.MIXED_FP32,
              "mixed_bf16x3": capi.MIXED_BF16X3, "bf16x3": capi.MIXED_BF16X3}


def from_namelist(path, dim, precision=None):
    from . import namelist
    params, gpu = namelist.read(path, dim)
    prec = PRECISIONS[str(precision or gpu["precision"]).lower()]
    if dim == 2:
        return build_2d(params, prec, str(gpu["rt_variant"]), int(gpu["seed"])), gpu
    return build_1d(params, prec), gpu


def sweep_c4(seed=7, n=64):
    """The batched Kn / force sweep: n force-driven channel cases (case 5),
    nx = 4, Nm = 8, Gauss-Hermite grids; per case Kn log-uniform on
    [0.01, 1], phi uniform on [0.01, 0.2], ny uniform integer on [32, 128],
    nv from {16, 20, 24, 28, 32}; every second case is a Fourier-type
    channel with TwR uniform on [1.1, 1.5] (TwL = 1).  Drawn with numpy's
    PCG64 from a fixed seed.  Returns the &params dicts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(n):
        kn = float(10 ** rng.uniform(-2.0, 0.0))
        phi = float(rng.uniform(0.01, 0.2))
        ny = int(rng.integers(32, 129))
        nv = int(rng.choice([16, 20, 24, 28, 32]))
        twr = float(rng.uniform(1.1, 1.5)) if k % 2 else 1.0
        out.append(dict(nx=4, ny=ny, xlength=0.1, ylength=1.0, nv=nv, vmax=5.0, Tref=1.0,
                        quad_type="gauss_hermite", force_type="hermite_symmetric", Nm=8,
                        cfl=0.8, maxtime=200.0, case_id=5, knudsen=kn, phi=phi, TwL=1.0,
                        TwR=twr))
    return out
