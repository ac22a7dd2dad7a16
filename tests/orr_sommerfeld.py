"""Orr-Sommerfeld linear stability of the laminar profile U = 1.5 (1 - y^2), built
from a solver's own wall-normal operators: the known answer of Orszag (1971),
Re_c = 1e4, alpha = 1: least-stable c = 0.23752649 + 0.00373967 i in units of the
centreline velocity (here U_c = 1.5, so nu = U_c / Re_c = 1.5e-4).

v ~ e^{lambda t}: (DY2 - a^2) lambda v = A v with, on interior rows,
  A = -i a U (DY2 - a^2) + i a U'' + nu (DY4 - 2 a^2 DY2 + a^4)
and boundary rows v = 0, DY1 v = 0 at both walls (the rows the step's Av uses).
c = i lambda / a / U_c.
"""
import numpy as np
import scipy.linalg as sl

ORSZAG_C = 0.23752649 + 0.00373967j
NU, ALPHA, UC = 1.5e-4, 1.0, 1.5


def boundary_rows(A, B, DY1):
    N = A.shape[0]
    I = np.eye(N)
    for r, row in ((0, I[0]), (N - 1, I[N - 1]), (1, DY1[0]), (N - 2, DY1[N - 1])):
        A[r] = row
        B[r] = 0.0


def least_stable(A, B):
    lam, V = sl.eig(A, B)
    ok = np.isfinite(lam) & (np.abs(lam) < 1e3)
    lam, V = lam[ok], V[:, ok]
    i = np.argmax(lam.real)
    return lam[i], V[:, i]


def textbook_operator(y, DY1, DY2, DY4, nu=NU, a=ALPHA):
    N = len(y)
    I = np.eye(N)
    U = UC * (1 - y ** 2)
    A = -1j * a * np.diag(U) @ (DY2 - a * a * I) + 1j * a * (-2 * UC) * I + nu * (DY4 - 2 * a * a * DY2 + a ** 4 * I)
    B = DY2 - a * a * I
    boundary_rows(A, B, DY1)
    return A, B


def c_of(lam, a=ALPHA):
    return 1j * lam / a / UC


def eigenmode_fields(shape, y, DY1, phi, eps, a=ALPHA):
    """u, v, w ([NZ][NXH][NY]) of the laminar profile plus eps * phi as v at
    (kz = 0, kx = a), u from continuity (the step's horizontal recovery)."""
    u = np.zeros(shape, dtype=complex)
    v = np.zeros(shape, dtype=complex)
    w = np.zeros(shape, dtype=complex)
    u[0, 0] = UC * (1 - y ** 2)
    v[0, 1] = eps * phi
    u[0, 1] = (1j / a) * (DY1 @ v[0, 1])
    return u, v, w


def growth(step, field_v, phi, dt, n1, n2):
    """lambda measured between steps n1 and n2 from the projection on phi."""
    amps = {}
    for n in range(1, n2 + 1):
        step()
        if n in (n1, n2):
            amps[n] = np.vdot(phi, field_v()) / np.vdot(phi, phi)
    return np.log(amps[n2] / amps[n1]) / ((n2 - n1) * dt)
