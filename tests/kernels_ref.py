"""Function-level references for the intermediate arrays of one 2D step,
built from the oracle's per-face routines (test infrastructure)."""
import ctypes as C

import numpy as np

import oracle_py


def slopes_2d(case, h):
    """van Leer / one-sided slopes of one cell array h [ncell, nn] in x and y
    (2d/src/solver.f90:125-159), with periodic wrap or wall one-sided."""
    d = case.desc
    nx, ny = d.nx, d.ny
    f = h.reshape(ny, nx, -1)
    out = []
    for ax, n, ln, per in ((1, nx, d.dx, d.bc[0] == 0), (0, ny, d.dy, d.bc[2] == 0)):
        den = 0.5 * ln + 0.5 * ln
        L = np.roll(f, 1, axis=ax)
        R = np.roll(f, -1, axis=ax)
        sL = (f - L) / den
        sR = (R - f) / den
        s = (np.copysign(1.0, sR) + np.copysign(1.0, sL)) * np.abs(sR) * np.abs(sL) / (
            np.abs(sR) + np.abs(sL) + d.smv)
        if not per:
            idx = [slice(None)] * 3
            idx[ax] = 0
            i1 = list(idx); i1[ax] = 1
            s[tuple(idx)] = (f[tuple(i1)] - f[tuple(idx)]) / den
            idx[ax] = n - 1
            i2 = list(idx); i2[ax] = n - 2
            s[tuple(idx)] = (f[tuple(idx)] - f[tuple(i2)]) / den
        out.append(s.reshape(nx * ny, -1))
    return out


def face_flux_interior(case, dir_, dt, L, R, acc, h, b, sh, sb, w):
    """Oracle calc_flux_mtforcev/h for the face between cells L and R."""
    lib = oracle_py.load()
    d = case.bind()
    nn = case.nn
    ptr = oracle_py.ptr
    cell = lambda c: (np.ascontiguousarray(h[c]), np.ascontiguousarray(b[c]),
                      np.ascontiguousarray(np.concatenate([sh[0][c], sh[1][c]])),
                      np.ascontiguousarray(np.concatenate([sb[0][c], sb[1][c]])),
                      np.ascontiguousarray(w[c]))
    hl, bl, shl, sbl, wl = cell(L)
    hr, br, shr, sbr, wr = cell(R)
    ln = np.array([d.dx, d.dy])
    flux = np.empty(4)
    fh = np.empty(nn)
    fb = np.empty(nn)
    flen = d.dy if dir_ == 1 else d.dx
    lib.oracle2d_flux_interior(C.addressof(d), dir_, dt, acc[0], acc[1], flen, ptr(hl), ptr(bl),
                               ptr(shl), ptr(sbl), ptr(wl), ptr(ln), ptr(hr), ptr(br), ptr(shr),
                               ptr(sbr), ptr(wr), ptr(ln), ptr(flux), ptr(fh), ptr(fb))
    return flux, fh, fb
