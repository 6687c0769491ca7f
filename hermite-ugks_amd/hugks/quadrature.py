"""Velocity quadratures the reference builds at start-up (init-time, host
side): Gauss-Hermite and closed Newton-Cotes.

Gauss-Hermite nodes are the roots of the physicists' H_n, found by Newton
iteration on the orthonormal three-term recurrence (which stays finite for
large n), starting from the classical asymptotic guesses for the largest
roots and extrapolating inward; the weight of a root is 2/(p'_n)^2 in the
orthonormal normalisation.  This is the published method the reference
uses (2d/src/velocity.f90:63-137, 1d/src/velocity.f90:65-151); the result
is converted, as there, to plain integration weights w*exp(x^2)*sqrt(Tref)
and nodes scaled by sqrt(Tref).
"""
import math

import numpy as np

PI_REF = 3.1415926535  # the reference's truncated PI (2d/src/data2d.f90:9)


def gauss_hermite_1d(n, tref=1.0, pi=PI_REF, tol=1e-14):
    """Returns (nodes, weights) on the physical axis, nodes increasing."""
    if n < 2:
        raise ValueError("gauss_hermite: nv must be >= 2")
    m = (n + 1) // 2
    roots = np.zeros(n)
    wgh = np.zeros(n)
    pos = []
    z = 0.0
    p_start = pi ** (-0.25)
    for i in range(m):
        if i == 0:
            z = math.sqrt(2.0 * n + 1.0) - 1.85575 * (2.0 * n + 1.0) ** (-1.0 / 6.0)
        elif i == 1:
            z = z - 1.14 * float(n) ** 0.426 / z
        elif i == 2:
            z = 1.86 * z - 0.86 * pos[0]
        elif i == 3:
            z = 1.91 * z - 0.91 * pos[1]
        else:
            z = 2.0 * z - pos[i - 2]
        for _ in range(100):
            p1, p2 = p_start, 0.0
            for j in range(1, n + 1):
                p0, p2 = p2, p1
                p1 = z * math.sqrt(2.0 / j) * p2 - math.sqrt((j - 1.0) / j) * p0
            dp = math.sqrt(2.0 * n) * p2
            zold = z
            z = zold - p1 / dp
            if abs(z - zold) <= tol * (abs(z) + 1.0):
                break
        else:
            raise RuntimeError("gauss_hermite: root iteration did not converge")
        pos.append(z)
        roots[i] = -z
        roots[n - 1 - i] = z
        wgh[i] = wgh[n - 1 - i] = 2.0 / (dp * dp)
    scale = math.sqrt(tref)
    nodes = roots * scale
    weights = np.array([wgh[i] * math.exp(roots[i] ** 2) * scale for i in range(n)])
    return nodes, weights, roots, wgh


def gauss_hermite_2d(n, tref=1.0, pi=PI_REF):
    """Tensor-product grid: u, v nodes [n] and weight [n*n] (u fastest),
    evaluated in the reference's operation order."""
    nodes, _, roots, wgh = gauss_hermite_1d(n, tref, pi)
    scale = math.sqrt(tref)
    w = np.empty(n * n)
    for j in range(n):
        for i in range(n):
            w[i + n * j] = (wgh[i] * math.exp(roots[i] ** 2) * scale * wgh[j] *
                            math.exp(roots[j] ** 2) * scale)
    return nodes.copy(), nodes.copy(), w


def newton_cotes_coeff(idx, num):
    """Closed 4-panel Newton-Cotes coefficient of node idx (1-based),
    2d/src/velocity.f90:49-61."""
    if idx == 1 or idx == num:
        return 14.0 / 45.0
    if (idx - 5) % 4 == 0:
        return 28.0 / 45.0
    if (idx - 3) % 4 == 0:
        return 24.0 / 45.0
    return 64.0 / 45.0


def newton_cotes_2d(n, vmax, tref=1.0):
    umin, umax = -vmax * math.sqrt(tref), vmax * math.sqrt(tref)
    du = (umax - umin) / (n - 1)
    u = np.array([umin + i * du for i in range(n)])
    w = np.empty(n * n)
    for j in range(n):
        for i in range(n):
            w[i + n * j] = (newton_cotes_coeff(i + 1, n) * du) * (newton_cotes_coeff(j + 1, n) * du)
    return u, u.copy(), w


def newton_cotes_1d(n, vmax, tref=1.0):
    umin, umax = -vmax * math.sqrt(tref), vmax * math.sqrt(tref)
    du = (umax - umin) / (n - 1)
    u = np.array([umin + i * du for i in range(n)])
    w = np.array([newton_cotes_coeff(i + 1, n) * du for i in range(n)])
    return u, w
