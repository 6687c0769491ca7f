"""Error measures used by the parity tests, smoke and bench.

field_errors(w, wr): per conserved component, max over cells of |w - wr|
divided by the field's scale: max |rho| for density, max |E| for energy
and, for both momentum components, the largest momentum magnitude
|(rho U, rho V)| (a component that is zero up to round-off, such as the
wall-normal momentum of Poiseuille flow, is measured against the momentum
the flow actually carries).  With thermal=True momentum is measured
against the thermal momentum sqrt(max rho * max rho E) instead: for a
state at hydrostatic rest rho U is the small difference of pressure and
force terms, and a one-ulp perturbation of the initial f moves it by
1e-10 of its own size within tens of steps.

prim_errors(w, wr): per primitive field (rho, U, V, T) on the same idea:
velocities against the largest speed, T against max T.

prim_errors_thermal(w, wr): as prim_errors with velocities against the
thermal speed sqrt(max T) (T = 1/lambda, so sqrt(T) is the most probable
molecular speed sqrt(2RT)); q_error_thermal(q, qr, wr): the heat flux
against max rho * (max T)^(3/2).  For low-speed flows (Mach 0.01) these
are the scales on which an FP32 distribution function resolves the
macroscopic fields.
"""
import numpy as np


def field_errors(w, wr, thermal=False):
    w = np.asarray(w)
    wr = np.asarray(wr)
    d = w.shape[1] - 2
    err = np.empty(w.shape[1])
    err[0] = np.max(np.abs(w[:, 0] - wr[:, 0])) / np.max(np.abs(wr[:, 0]))
    if thermal:
        ms = np.sqrt(np.max(np.abs(wr[:, 0])) * np.max(np.abs(wr[:, -1])))
    else:
        ms = max(np.max(np.sqrt(np.sum(wr[:, 1:1 + d] ** 2, axis=1))), 1e-300)
    for q in range(1, 1 + d):
        err[q] = np.max(np.abs(w[:, q] - wr[:, q])) / ms
    err[-1] = np.max(np.abs(w[:, -1] - wr[:, -1])) / np.max(np.abs(wr[:, -1]))
    return err


def primitives(w, gamma):
    w = np.asarray(w)
    d = w.shape[1] - 2
    rho = w[:, 0]
    vel = w[:, 1:1 + d] / rho[:, None]
    ke = 0.5 * np.sum(w[:, 1:1 + d] ** 2, axis=1) / rho
    lam = 0.5 * rho / (gamma - 1.0) / (w[:, -1] - ke)
    return rho, vel, 1.0 / lam


def prim_errors(w, wr, gamma):
    """Relative errors of rho, each velocity component and T."""
    r, v, t = primitives(w, gamma)
    rr, vr, tr = primitives(wr, gamma)
    out = [np.max(np.abs(r - rr)) / np.max(np.abs(rr))]
    vs = max(np.max(np.sqrt(np.sum(vr ** 2, axis=1))), 1e-300)
    for q in range(v.shape[1]):
        out.append(np.max(np.abs(v[:, q] - vr[:, q])) / vs)
    out.append(np.max(np.abs(t - tr)) / np.max(np.abs(tr)))
    return np.array(out)


def q_error(q, qr):
    qs = max(np.max(np.sqrt(np.sum(np.asarray(qr) ** 2, axis=1))), 1e-300)
    return np.max(np.abs(np.asarray(q) - np.asarray(qr))) / qs


def prim_errors_thermal(w, wr, gamma):
    """rho and T as prim_errors; velocities against sqrt(max T)."""
    r, v, t = primitives(w, gamma)
    rr, vr, tr = primitives(wr, gamma)
    cth = np.sqrt(np.max(np.abs(tr)))
    out = [np.max(np.abs(r - rr)) / np.max(np.abs(rr))]
    for q in range(v.shape[1]):
        out.append(np.max(np.abs(v[:, q] - vr[:, q])) / cth)
    out.append(np.max(np.abs(t - tr)) / np.max(np.abs(tr)))
    return np.array(out)


def q_error_thermal(q, qr, wr, gamma):
    rr, _, tr = primitives(wr, gamma)
    scale = np.max(np.abs(rr)) * np.max(np.abs(tr)) ** 1.5
    return np.max(np.abs(np.asarray(q) - np.asarray(qr))) / scale
