"""Known answers of the time step that the mathematics fixes, shared by the
oracle's pins (test_oracle.py) and their GPU twins (test_gpu_pins.py). Each
helper drives a solver (the oracle, or the library through its C ABI) with
set_velocity / step / field and returns how far it is from the known answer.
None of them restates the step: each compares the solver with an identity, a
recursion on a single mode, or an eigenproblem built here from the solver's own
DY matrices.

  recovery_identities  horizontal recovery (src/solver.cpp:516-539): four exact
                       identities per pencil, and the mean pencil's vorticity
  squire_sequence      a kx = 0 omega_y mode without mean flow decays by the
                       exact BDF3 recursion at the Dirichlet eigenvalue of DY2
  mean_flow_recursion  the mean flow (src/solver.cpp:428-514, 306-327) follows
                       the implicit recursion with the friction and mean-Lamb
                       histories, level by level
  flow_rate            the friction force holds the bulk flow rate
  oblique_mode         an oblique (kx, kz != 0) eigenmode of the coupled
                       Orr-Sommerfeld / Squire operator grows, turns and keeps
                       its omega_y / v ratio as the eigenproblem says
"""
import numpy as np
import scipy.linalg as sl

G0 = 11.0 / 6.0
AL = (3.0, -1.5, 1.0 / 3.0)
BE = (3.0, -3.0, 1.0)


def grid_y(NY, alpha):
    s = 2.0 * np.arange(NY) / (NY - 1) - 1.0
    return np.tanh(alpha * s) / np.tanh(alpha)


def wavenumbers(NX, NZ, dkx, dkz):
    kx = np.arange(NX // 2) * dkx
    kz = np.array([k if k < NZ // 2 else k - NZ for k in range(NZ)], dtype=float) * dkz
    return kx, kz


def recovery_identities(s, kx, kz, DY1, DY2):
    """Residuals, relative to the size of each identity's terms, of
        i kx u + DY1 v + i kz w = 0            (divergence)
        omega_y = i kz u - i kx w                (u, w carry omega_y's terms)
        kx omega_x + kz omega_z = i DY1 omega_y  (omega_x, omega_z carry DY1 omega_y)
        kz omega_x - kx omega_z = i (DY2 - k^2) v
    on every k^2 > 0 pencil, and omega_x = DY1 W, omega_z = -DY1 U on the mean pencil."""
    u, v, w, ox, oy, oz = (s.field(n) for n in ("u", "v", "w", "omx", "omy", "omz"))
    res = np.zeros(5)
    NZ, NXH = u.shape[:2]
    for k in range(NZ):
        for i in range(NXH):
            a, b = kx[i], kz[k]
            k2 = a * a + b * b
            if k2 < 1e-8:
                r1 = np.abs(ox[k, i] - DY1 @ w[k, i]).max() / max(np.abs(ox[k, i]).max(), 1e-300)
                r2 = np.abs(oz[k, i] + DY1 @ u[k, i]).max() / max(np.abs(oz[k, i]).max(), 1e-300)
                res[4] = max(res[4], r1 if np.abs(ox[k, i]).max() > 0 else 0.0, r2)
                continue
            d1v = DY1 @ v[k, i]
            terms = [
                (1j * a * u[k, i] + d1v + 1j * b * w[k, i], np.abs(d1v).max()),
                (oy[k, i] - (1j * b * u[k, i] - 1j * a * w[k, i]), np.abs(oy[k, i]).max()),
                (a * ox[k, i] + b * oz[k, i] - 1j * (DY1 @ oy[k, i]), np.abs(DY1 @ oy[k, i]).max()),
                (b * ox[k, i] - a * oz[k, i] - 1j * (DY2 @ v[k, i] - k2 * v[k, i]),
                 np.abs(DY2 @ v[k, i] - k2 * v[k, i]).max()),
            ]
            for q, (r, scale) in enumerate(terms):
                if scale > 0:
                    res[q] = max(res[q], np.abs(r).max() / scale)
    return res


def squire_mode(DY2):
    """Least-damped real Dirichlet eigenpair (lambda, phi) of DY2's interior block."""
    N = DY2.shape[0]
    lam, V = np.linalg.eig(DY2[1:-1, 1:-1])
    real = np.abs(lam.imag) < 1e-9 * np.abs(lam)
    i = np.flatnonzero(real)[np.argmax(lam.real[real])]
    phi = np.zeros(N, dtype=complex)
    phi[1:-1] = V[:, i].real / np.abs(V[:, i].real).max()
    return lam[i].real, phi


def squire_sequence(s, shape, DY2, kz1, nu, dt, steps, eps=1e-8):
    """kx = 0, kz = kz1 (index 1), no mean flow: omega_y = eps phi, v = 0, so u =
    -(i kz / k^2) omega_y and Hg vanishes to O(eps^2). The amplitude must follow
        (G0 - z) a_{n+1} = AL0 a_n + AL1 a_{n-1} + AL2 a_{n-2},  z = dt nu (lambda - kz^2),
    from a_0 = a_{-1} = a_{-2} (the impulsive start). Returns the largest deviation
    relative to a_0."""
    lam, phi = squire_mode(DY2)
    NZ = shape[0]
    u = np.zeros(shape, dtype=complex)
    u[1, 0] = -(1j / kz1) * eps * phi
    u[NZ - 1, 0] = np.conj(u[1, 0])
    z0 = np.zeros(shape, dtype=complex)
    s.set_velocity(u, z0, z0)

    def amp():
        return np.vdot(phi, s.field("omy")[1, 0]) / np.vdot(phi, phi)
    a = [amp()]
    for _ in range(steps):
        s.step(1)
        a.append(amp())
    zz = dt * nu * (lam - kz1 * kz1)
    b = [a[0]] * 3
    for _ in range(steps):
        b.append((AL[0] * b[-1] + AL[1] * b[-2] + AL[2] * b[-3]) / (G0 - zz))
    return np.abs(np.array(a) - np.array(b[2:])).max() / abs(a[0])


def mean_flow_recursion(s, shape, y, DY1, DY2, kz1, nu, dt, steps):
    """Mean flow U(y) = 1.5 (1 - y^2) + 0.1 y (1 - y^2)^2 plus a streamwise vortex at
    (kx = 0, kz = kz1): v = 0.1 (1 - y^2)^2, w = i DY1 v / kz1. The vortex's mean Lamb
    vector drives W (Lz = -<v omega_x>), the friction forces fx, fz hold U and W, and
    nothing else acts on the mean pencil. Each step n -> n+1 of (U, W) must be
        (G0/dt - nu DY2) X_{n+1} = sum_j BE_j (L_{n-j} + f_{n-j}) + sum_j AL_j X_{n-j} / dt
    with Dirichlet rows, L the solver's own mean Lamb component of level n - j
    (field lmbx or lmbz, k = 0) and f = -nu / Ly ((DY1 X)_{N-1} - (DY1 X)_0) of that
    level (the impulsive start repeats level 0). Returns the largest one-step
    deviation relative to |X|, and the largest friction deviation of the solver's
    own friction series (diag) relative to fx."""
    N, NZ = len(y), shape[0]
    U0 = 1.5 * (1 - y ** 2) + 0.1 * y * (1 - y ** 2) ** 2
    u = np.zeros(shape, dtype=complex)
    v = np.zeros(shape, dtype=complex)
    w = np.zeros(shape, dtype=complex)
    u[0, 0] = U0
    v[1, 0] = 0.1 * (1 - y ** 2) ** 2
    w[1, 0] = 1j * (DY1 @ v[1, 0]) / kz1
    v[NZ - 1, 0] = np.conj(v[1, 0])
    w[NZ - 1, 0] = np.conj(w[1, 0])
    s.set_velocity(u, v, w)
    Ly = y[-1] - y[0]

    def fric(X):
        d = DY1 @ X
        return -nu / Ly * (d[N - 1].real - d[0].real)

    def level():
        U, W = s.field("u")[0, 0], s.field("w")[0, 0]
        return U, W, s.field("lmbx")[0, 0], s.field("lmbz")[0, 0]
    hist = [level()] * 3
    fxs = [fric(hist[0][0])]
    A = -nu * DY2 + (G0 / dt) * np.eye(N)
    A[0] = 0
    A[-1] = 0
    A[0, 0] = A[-1, -1] = 1.0
    dev = 0.0
    for _ in range(steps):
        s.step(1)
        cur = level()
        for c, lc in ((0, 2), (1, 3)):
            rhs = sum(BE[j] * (hist[j][lc] + fric(hist[j][c])) + AL[j] * hist[j][c] / dt for j in range(3))
            rhs[0] = rhs[-1] = 0.0
            pred = np.linalg.solve(A, rhs)
            dev = max(dev, np.abs(pred - cur[c]).max() / np.abs(cur[0]).max())
        hist = [cur] + hist[:2]
        fxs.append(fric(cur[0]))
    d = s.diag()[-(steps + 1):, 0]
    return dev, np.abs(d - np.array(fxs)).max() / np.abs(np.array(fxs)).max()


def flow_rate_drift(s, y, steps):
    """Relative change of the bulk flow rate sum_j q_j U_j (trapezoid weights)
    over `steps` steps from the field the solver holds."""
    q = np.empty_like(y)
    q[1:-1] = (y[2:] - y[:-2]) / 2
    q[0], q[-1] = (y[1] - y[0]) / 2, (y[-1] - y[-2]) / 2
    Q0 = (q * s.field("u")[0, 0].real).sum()
    s.step(steps)
    return abs((q * s.field("u")[0, 0].real).sum() - Q0) / abs(Q0)


def coupled_operator(y, DY1, DY2, DY4, a, b, nu, UC=1.5):
    """The linearized step about U = UC (1 - y^2) at (kx, kz) = (a, b), state (v, omega_y):
        (DY2 - k^2) dv/dt = [-i a U (DY2 - k^2) + i a U'' + nu (DY2 - k^2)^2] v   (rows of v,
                            with DY4 - 2 k^2 DY2 + k^4 for the fourth-order term)
        d omega_y/dt = -i b U' v - i a U omega_y + nu (DY2 - k^2) omega_y
    U' = DY1 U (as the step's Lamb vector carries it), U'' = -2 UC; boundary rows
    v = 0 and DY1 v = 0 at both walls, omega_y = 0 at both walls."""
    N = len(y)
    I = np.eye(N)
    k2 = a * a + b * b
    U = UC * (1 - y ** 2)
    Up = DY1 @ U
    L = DY2 - k2 * I
    Aos = -1j * a * np.diag(U) @ L + 1j * a * (-2 * UC) * I + nu * (DY4 - 2 * k2 * DY2 + k2 * k2 * I)
    Asq = -1j * a * np.diag(U) + nu * L
    A = np.zeros((2 * N, 2 * N), dtype=complex)
    B = np.zeros((2 * N, 2 * N), dtype=complex)
    A[:N, :N] = Aos
    B[:N, :N] = L
    A[N:, :N] = np.diag(-1j * b * Up)
    A[N:, N:] = Asq
    B[N:, N:] = I
    for r, row in ((0, I[0]), (N - 1, I[N - 1]), (1, DY1[0]), (N - 2, DY1[N - 1])):
        A[r] = 0
        A[r, :N] = row
        B[r] = 0
    for r in (N, 2 * N - 1):
        A[r] = 0
        A[r, r] = 1
        B[r] = 0
    return A, B


def oblique_mode(s, shape, y, DY1, DY2, DY4, a, b, nu, dt, n1, n2, eps=1e-6):
    """From the laminar profile plus eps times the least-stable Orr-Sommerfeld-type
    eigenmode (v part non-zero) of coupled_operator at pencil (kz index 1, kx index 1):
    returns (measured lambda from v between steps n1 and n2, the discrete eigenvalue,
    the ratio of the omega_y and v amplitudes at step n2 to the eigenvector's)."""
    N = len(y)
    A, B = coupled_operator(y, DY1, DY2, DY4, a, b, nu)
    lam, V = sl.eig(A, B)
    ok = np.isfinite(lam) & (np.abs(lam) < 1e3)
    vpart = np.linalg.norm(V[:N], axis=0) > 1e-3 * np.linalg.norm(V, axis=0)
    cand = np.flatnonzero(ok & vpart)
    i = cand[np.argmax(lam[cand].real)]
    phv, phg = V[:N, i], V[N:, i]
    sc = phv[np.argmax(np.abs(phv))]
    phv, phg = phv / sc, phg / sc
    k2 = a * a + b * b
    u = np.zeros(shape, dtype=complex)
    v = np.zeros(shape, dtype=complex)
    w = np.zeros(shape, dtype=complex)
    u[0, 0] = 1.5 * (1 - y ** 2)
    d1v = DY1 @ (eps * phv)
    v[1, 1] = eps * phv
    u[1, 1] = (1j * a / k2) * d1v - (1j * b / k2) * (eps * phg)
    w[1, 1] = (1j * a / k2) * (eps * phg) + (1j * b / k2) * d1v
    s.set_velocity(u, v, w)
    amps = {}
    for n in range(1, n2 + 1):
        s.step(1)
        if n in (n1, n2):
            av = np.vdot(phv, s.field("v")[1, 1]) / np.vdot(phv, phv)
            ag = np.vdot(phg, s.field("omy")[1, 1]) / np.vdot(phg, phg)
            amps[n] = (av, ag)
    lm = np.log(amps[n2][0] / amps[n1][0]) / ((n2 - n1) * dt)
    return lm, lam[i], amps[n2][1] / amps[n2][0]
