"""Pins of the CPU oracle (oracle/owt_oracle.cpp) against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures, and it cannot be built
in this image (its path needs FFTW, LAPACKE/CBLAS headers and parallel HDF5).
So each part of the oracle is checked here against an independent fact:

  transforms      brute-force DFT; dealiasing identity against a direct
                  spectral convolution with the reference's conventions
                  (z-Nyquist dropped, kx = 0 enters through Re, 3/2 padding)
  dense LU        LAPACK itself (numpy.linalg.solve / scipy lu_factor are
                  dgesv / dgetrf, the routines src/linalg.hpp calls)
  compact rows    exactness degree on local Taylor monomials (Lele's schemes),
                  the B3 top-row overwrite signature (src/compdiff.hpp:257-258)
  metric, DY_n    derivatives of tanh(a s)/tanh(a) and of smooth functions
  step            laminar Poiseuille balance fx_fric = 3 nu, Hermitian symmetry
                  of the kx = 0 plane, discrete divergence, seeded amplitude
  linear dynamics Orr-Sommerfeld at Re 1e4, alpha 1: the step's Lamb-vector path
                  linearized about the laminar profile equals the textbook operator
                  on the same DY matrices; with DY4 = DY2^2 that operator gives
                  Orszag's eigenvalue; the time-stepped eigenmode grows at the
                  discrete eigenvalue
  pressure        (no reference counterpart) laminar P = u^2/2 - mean; a gradient
                  Lamb vector L = grad phi gives P = phi; wall data nu v'' alone
                  gives A cosh(ky) + B sinh(ky); each converges with NY
"""
from math import factorial

import numpy as np
import pytest

RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)


def make(O, NX, NY, NZ, **kw):
    p = dict(RE180)
    p.update(kw)
    return O.Oracle(p["nu"], p["dt"], p["dkx"], p["dkz"], p["alpha_mesh"], NX, NY, NZ)


# ---------------------------------------------------------------------------
# transforms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 9, 12, 16, 18, 24, 36, 48, 72, 96, 144, 192, 288, 576])
def test_fft_matches_bruteforce_dft(oracle_mod, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    for sign in (-1, 1):
        F = np.exp(sign * 2j * np.pi * (np.outer(j, j) % n) / n)
        ref = F @ x
        got = oracle_mod.fft(x, sign)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def effective_spectrum(F, NXH, NZH, NZ):
    """Full 2-D spectrum S[i + NXH - 1, m + NZH - 1] of the physical field the
    reference's Fourier2Physical_dealiased produces from wave field F[k][i][y]."""
    NY = F.shape[2]
    S = np.zeros((2 * NXH - 1, 2 * NZH - 1, NY), dtype=complex)
    def k_of(m):
        return m if m >= 0 else m + NZ
    for m in range(-NZH + 1, NZH):
        for i in range(1, NXH):
            S[i + NXH - 1, m + NZH - 1] = F[k_of(m), i]
            S[-i + NXH - 1, -m + NZH - 1] = np.conj(F[k_of(m), i])
        S[NXH - 1, m + NZH - 1] = 0.5 * (F[k_of(m), 0] + np.conj(F[k_of(-m), 0]))
    return S


def conv2(A, B):
    a0, a1 = A.shape[:2]
    b0, b1 = B.shape[:2]
    H = np.zeros((a0 + b0 - 1, a1 + b1 - 1, A.shape[2]), dtype=complex)
    for p in range(a0):
        for q in range(a1):
            H[p:p + b0, q:q + b1] += A[p, q] * B
    return H


@pytest.mark.parametrize("NX,NY,NZ", [(8, 7, 8), (12, 9, 16), (16, 7, 12)])
def test_lamb_dealiasing_identity(oracle_mod, NX, NY, NZ):
    """u x omega through the 3/2-padded transforms equals the direct, truncated
    spectral convolution: exact to roundoff for every retained mode."""
    o = make(oracle_mod, NX, NY, NZ)
    NXH, NZH = NX // 2, NZ // 2
    rng = np.random.default_rng(NX * 100 + NZ)
    six = rng.standard_normal((6, NZ, NXH, NY)) + 1j * rng.standard_normal((6, NZ, NXH, NY))
    got = o.lamb(six)
    S = [effective_spectrum(f, NXH, NZH, NZ) for f in six]
    u, v, w, ox, oy, oz = S
    H = [conv2(v, oz) - conv2(w, oy), conv2(w, ox) - conv2(u, oz), conv2(u, oy) - conv2(v, ox)]
    ci, cm = 2 * NXH - 2, 2 * NZH - 2  # index of (i, m) = (0, 0) in the convolution
    for q in range(3):
        ref = np.zeros((NZ, NXH, NY), dtype=complex)
        for k in range(NZ):
            m = k if k < NZH else k - NZ
            for i in range(NXH):
                ref[k, i] = H[q][ci + i, cm + m]
        err = np.abs(got[q] - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (q, err)


def test_f2p_p2f_roundtrip_and_realness(oracle_mod):
    o = make(oracle_mod, 16, 9, 16)
    rng = np.random.default_rng(5)
    F = rng.standard_normal(o.wave_shape) + 1j * rng.standard_normal(o.wave_shape)
    P = o.f2p(F)
    back = o.p2f(P)
    # modes that survive the round trip: z-Nyquist dropped, kx = 0 Hermitian-projected
    S = effective_spectrum(F, o.NXH, o.NZH, o.NZ)
    expect = np.zeros_like(F)
    for k in range(o.NZ):
        m = k if k < o.NZH else k - o.NZ
        if k == o.NZH:
            continue
        for i in range(o.NXH):
            expect[k, i] = S[i + o.NXH - 1, m + o.NZH - 1]
    assert np.abs(back - expect).max() < 1e-13 * np.abs(expect).max()


# ---------------------------------------------------------------------------
# dense LU (src/linalg.hpp:493-540, 615-689)
# ---------------------------------------------------------------------------
def test_lu_matches_lapack(oracle_mod):
    import scipy.linalg as sla
    rng = np.random.default_rng(0)
    for n in (5, 17, 65, 129):
        A = rng.standard_normal((n, n))
        b = rng.standard_normal((n, 3))
        x = oracle_mod.lu_solve(A, b)
        ref = np.linalg.solve(A, b)
        assert np.abs(x - ref).max() <= 1e-11 * np.abs(ref).max()
        lu, piv, info = oracle_mod.lu_factor(A)
        lu_ref, piv_ref = sla.lu_factor(A)
        assert info == 0
        assert np.array_equal(piv, piv_ref)  # same pivot rows as dgetrf
        assert np.abs(lu - lu_ref).max() <= 1e-12 * np.abs(lu_ref).max()


def test_lu_singular_reports_info(oracle_mod):
    A = np.ones((4, 4))
    with pytest.raises(np.linalg.LinAlgError):
        oracle_mod.lu_solve(A, np.ones(4))


def assemble_Av_Ag(o, k2):
    """Av, Ag with boundary rows as written in src/solver.cpp:577-591."""
    DY1, DY2, DY4 = o.matrix("DY1"), o.matrix("DY2"), o.matrix("DY4")
    N, nu, dt, g0 = o.NY, o.nu, o.dt, 11 / 6
    Av = -nu * DY4 + (g0 / dt + 2 * nu * k2) * DY2 - (nu * k2 ** 2 + g0 * k2 / dt) * np.eye(N)
    Ag = -nu * DY2 + (nu * k2 + g0 / dt) * np.eye(N)
    for M in (Av, Ag):
        M[0] = 0
        M[0, 0] = 1
        M[-1] = 0
        M[-1, -1] = 1
    Av[1] = DY1[0]
    Av[-2] = DY1[-1]
    return Av, Ag


def test_lu_on_step_operators_matches_lapack(oracle_mod):
    o = make(oracle_mod, 8, 65, 8)
    rng = np.random.default_rng(3)
    for k2 in (0.0, 0.25, 17.0, 400.0):
        for M in assemble_Av_Ag(o, k2):
            b = rng.standard_normal(o.NY)
            x = oracle_mod.lu_solve(M, b)
            ref = np.linalg.solve(M, b)
            # both are dgetrf/dgetrs; residual-level agreement on operators with cond ~ 1e8
            assert np.linalg.norm(x - ref) <= 1e-9 * np.linalg.norm(ref)


@pytest.mark.parametrize("NY", [65, 160, 193])
def test_v_operator_mirror_form(oracle_mod, NY):
    """The premise of the mixed step's mirror-form v solves: with row N-2
    negated, Av's part that is odd under the reflection j -> N-1-j is rank 2
    and the same for every k^2 (it comes from the reference's DY4 top-row
    overwrite; DY1 and DY2 are mirror symmetric on the rows Av keeps), so
    X b = Ss b - P (Q b) with Ss the symmetrised inverse of S = R Av - U V^T
    (Woodbury). Checked against LAPACK's solve on smooth right-hand sides, and
    against the full inverse's own error on right-hand sides carrying 1e-7
    noise (the FP32 Lamb vector's rounding)."""
    o = make(oracle_mod, 8, NY, 8, nu=1e-4, alpha_mesh=2.0)
    N = NY
    J = np.eye(N)[::-1]
    R = np.eye(N)
    R[N - 2, N - 2] = -1
    y = o.vector("y")
    rng = np.random.default_rng(NY)
    for k2 in (0.25, 17.0, 400.0, 1e4):
        Av, _ = assemble_Av_Ag(o, k2)
        Ap = R @ Av
        E = (Ap - J @ Ap @ J) / 2
        sv = np.linalg.svd(E, compute_uv=False)
        # two directions of about 1e-3 |Av|, the rest at Av's rounding
        assert sv[1] >= 1e-4 * np.linalg.norm(Ap, 2) and sv[2] <= 1e-14 * np.linalg.norm(Ap, 2), sv[:4]
        U_, s_, Vt = np.linalg.svd(E)
        U, V = U_[:, :2], (s_[:2, None] * Vt[:2]).T
        X = np.linalg.inv(Av)
        Xp = X @ R
        P = Xp @ U
        G = V.T @ Xp
        Q = np.linalg.solve(np.eye(2) - G @ U, G)
        Sinv = Xp + P @ Q
        Ss = (Sinv + J @ Sinv @ J) / 2
        for noise in (0.0, 1e-7):
            c = rng.standard_normal(5)
            v = (1 - y ** 2) ** 2 * sum(c[m] * np.cos(m * np.pi * y / 2 + c[m]) for m in range(5))
            b = Av @ v
            b = b + noise * np.abs(b).max() * rng.standard_normal(N)
            b[[0, 1, N - 2, N - 1]] = 0
            # the solution to well below FP64 rounding: LU refined with long-double residuals
            ref = np.linalg.solve(Av, b)
            for _ in range(3):
                r = (b.astype(np.longdouble) - Av.astype(np.longdouble) @ ref.astype(np.longdouble)).astype(np.float64)
                ref = ref + np.linalg.solve(Av, r)
            e = lambda x: np.linalg.norm(x - ref) / np.linalg.norm(ref)
            xm = Ss @ b - P @ (Q @ b)
            # as accurate as the full inverse and LAPACK's solve, to a factor
            assert e(xm) <= max(1e-12, 10 * max(e(X @ b), e(np.linalg.solve(Av, b)))), (k2, noise, e(xm), e(X @ b))


# ---------------------------------------------------------------------------
# compact schemes and metrics (src/compdiff.hpp, src/solver.cpp:57-216)
# ---------------------------------------------------------------------------
def exactness_degrees(A, B, s, n):
    h = s[1] - s[0]
    Bh = B * h ** n
    out = []
    for i in range(len(s)):
        t = (s - s[i]) / h
        d = -1
        for p in range(0, 14):
            f = t ** p
            fn = np.zeros_like(s) if p < n else factorial(p) / factorial(p - n) * t ** (p - n)
            r = Bh[i] @ f - A[i] @ fn
            scale = np.abs(Bh[i]) @ np.abs(f) + np.abs(A[i]) @ np.abs(fn)
            if abs(r) > 1e-11 * scale:
                break
            d = p
        out.append(d)
    return out


def test_compact_rows_exactness(oracle_mod):
    """Degree p up to which row i satisfies B f = A f^(n) exactly for f = ((s - s_i)/h)^p.
    Interior rows: Lele's pentadiagonal 8th-order schemes for d/ds, d2/ds2, d3/ds3
    (degree 8 + n - 1) and the tridiagonal 6th-order scheme for d4/ds4 (degree 9).
    Near-wall rows: the degrees their coefficient sets give (the reference's comments
    overstate the order of the d3 and d4 closures). The last row of the third
    derivative is not exact even for constants: the overwrite at
    src/compdiff.hpp:257-258 leaves -e0/h^3 at column N-4 and nothing at N-5."""
    o = make(oracle_mod, 8, 33, 8)
    s = o.vector("s")
    expected = {
        1: [4, 4] + [8] * 29 + [4, 4],
        2: [5, 7] + [9] * 29 + [7, 5],
        3: [5, 6, 8] + [10] * 27 + [8, 6, -1],
        4: [5, 6, 7] + [9] * 27 + [7, 6, 5],
    }
    for n in range(1, 5):
        got = exactness_degrees(o.matrix(f"A{n}"), o.matrix(f"B{n}"), s, n)
        assert got == expected[n], (n, got)


def test_b3_top_row_signature(oracle_mod):
    o = make(oracle_mod, 8, 33, 8)
    h = o.vector("s")[1] - o.vector("s")[0]
    B3 = o.matrix("B3") * h ** 3
    N = o.NY
    # bottom closure (8, -26, 30, -14, 2) annihilates constants; the top row as the
    # reference leaves it is (-8, 26, -30, -2, 0) read from the wall inward
    assert abs(B3[0].sum()) < 1e-12
    np.testing.assert_allclose(B3[N - 1, N - 5:][::-1], [-8, 26, -30, -2, 0], atol=1e-12)
    assert abs(B3[N - 1].sum() + 14.0) < 1e-12
    # every other row is consistent
    for i in range(N - 1):
        assert abs(B3[i].sum()) < 1e-9 * np.abs(B3[i]).sum()


def test_metric_matches_tanh_derivatives(oracle_mod):
    o = make(oracle_mod, 8, 65, 8, alpha_mesh=2.0)
    s = o.vector("s")
    a = 2.0
    th = np.tanh(a * s)
    sech2 = 1 - th ** 2
    # derivatives of y = tanh(a s)/tanh(a), differentiated independently
    d1 = a * sech2
    d2 = -2 * a ** 2 * sech2 * th
    d3 = -2 * a ** 3 * sech2 * (sech2 - 2 * th ** 2)
    d4 = 8 * a ** 4 * sech2 * th * (2 * sech2 - th ** 2)
    for name, ref in (("dyds1", d1), ("dyds2", d2), ("dyds3", d3), ("dyds4", d4)):
        got = o.vector(name)
        np.testing.assert_allclose(got, ref / np.tanh(a), rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(o.vector("y"), np.tanh(a * s) / np.tanh(a), rtol=1e-15, atol=1e-15)


def test_dy_operators_converge_on_stretched_grid(oracle_mod):
    """DY_n f -> f^(n) for a smooth f(y); the error falls at high order with N.
    DY4 is checked on the lower half only: its top-wall rows carry the D3 overwrite."""
    f = lambda y: np.sin(2 * y) + y ** 3
    df = [lambda y: 2 * np.cos(2 * y) + 3 * y ** 2, lambda y: -4 * np.sin(2 * y) + 6 * y,
          lambda y: -8 * np.cos(2 * y) + 6, lambda y: 16 * np.sin(2 * y)]
    errs = {}
    for N in (33, 65, 129):
        o = make(oracle_mod, 8, N, 8)
        y = o.vector("y")
        for n in (1, 2, 4):
            D = o.matrix(f"DY{n}")
            rows = slice(0, N) if n < 4 else slice(0, N // 2)
            e = np.abs((D @ f(y) - df[n - 1](y))[rows]).max() / np.abs(df[n - 1](y)).max()
            errs[(n, N)] = e
    # the low-order near-wall closures (degree 4..5) set the rate: better than
    # third order everywhere, and the absolute error shrinks to the closures' level
    assert errs[(1, 129)] < 1e-6 and errs[(2, 129)] < 1e-5 and errs[(4, 129)] < 0.5
    for n in (1, 2, 4):
        assert errs[(n, 129)] < errs[(n, 65)] / 8 and errs[(n, 65)] < errs[(n, 33)] / 8, (n, errs)
    # the overwrite makes DY4 wrong near the top wall
    o = make(oracle_mod, 8, 65, 8)
    y = o.vector("y")
    top = np.abs((o.matrix("DY4") @ f(y) - df[3](y))[-3:]).max() / np.abs(df[3](y)).max()
    assert top > 100 * errs[(4, 65)]


# ---------------------------------------------------------------------------
# step invariants and closed forms
# ---------------------------------------------------------------------------
def laminar_fields(o):
    y = o.vector("y")
    u = np.zeros(o.wave_shape, dtype=complex)
    u[0, 0] = 1.5 * (1 - y ** 2)
    z = np.zeros(o.wave_shape, dtype=complex)
    return u, z, z


def test_laminar_poiseuille_balance(oracle_mod):
    """U = 1.5(1 - y^2): friction force 3 nu balances nu U''; the flow stays put and
    the nonlinear terms vanish (src/solver.cpp:306-327, 428-514)."""
    o = make(oracle_mod, 16, 65, 16)
    u, v, w = laminar_fields(o)
    o.set_velocity(u, v, w)
    fx, bottom, upper = o.wall_shear()
    # DY1 wall rows are 4th order: on this grid the parabola is differentiated to ~1e-5
    assert abs(fx - 3 * o.nu) < 5e-5 * 3 * o.nu
    assert abs(bottom + upper) < 1e-12 and abs(bottom - 3 * o.nu) < 5e-5 * 3 * o.nu
    o.step(20)
    U = o.field("u")[0, 0]
    assert np.abs(U - u[0, 0]).max() < 1e-6
    # zero up to the transforms' roundoff on a field that is constant in x and z
    for name in ("v", "w", "Hv", "Hg", "lmbx", "lmbz", "omx", "omy"):
        assert np.abs(o.field(name)).max() < 1e-14, name
    d = o.diag()
    assert d.shape == (21, 3)
    assert np.all(np.abs(d[:, 0] - 3 * o.nu) < 5e-5 * 3 * o.nu)


def test_seeded_init_properties(oracle_mod):
    o = make(oracle_mod, 32, 33, 32)
    o.init(2, seed=7, amp=0.15)
    u, v, w = o.field("u"), o.field("v"), o.field("w")
    y = o.vector("y")
    up = u.copy()
    up[0, 0] -= 1.5 * (1 - y ** 2)
    P = np.stack([o.f2p(up), o.f2p(v), o.f2p(w)])
    assert abs(np.sqrt((P ** 2).sum(0)).max() - 0.15) < 1e-12
    # v and omega_y vanish at the walls exactly; u and w carry the truncation error
    # of DY1 at the wall rows until the first solve imposes DY1 v = 0 there
    assert np.abs(v[:, :, 0]).max() == 0 and np.abs(v[:, :, -1]).max() == 0
    assert np.abs(o.field("omy")[:, :, [0, -1]]).max() < 1e-15
    for f in (u, w):
        assert np.abs(f[1:, :, [0, -1]]).max() < 1e-3 * 0.15 and np.abs(f[0, 1:, [0, -1]]).max() < 1e-3 * 0.15
    # Hermitian kx = 0 plane
    for m in range(1, o.NZH):
        for f in (u, v, w):
            assert np.abs(f[m, 0] - np.conj(f[o.NZ - m, 0])).max() < 1e-15


def trapezoid_weights(y):
    q = np.empty_like(y)
    q[1:-1] = (y[2:] - y[:-2]) / 2
    q[0], q[-1] = (y[1] - y[0]) / 2, (y[-1] - y[-2]) / 2
    return q


def test_pressure_laminar_closed_form(oracle_mod):
    """Poiseuille flow: L = u x omega = (0, -u omega_z, 0) with omega_z = -u', so
    dP/dy = Ly gives P = u^2/2 up to the gauge constant (zero trapezoid mean);
    every other pencil is zero. The discrete DY1 and DY2 are exact to their
    truncation error only on the stretched grid, so the error falls with NY."""
    errs = []
    for NY in (33, 65):
        o = make(oracle_mod, 16, NY, 16)
        u, v, w = laminar_fields(o)
        o.set_velocity(u, v, w)
        P = o.field("p")
        y = o.vector("y")
        ex = 0.5 * u[0, 0].real ** 2
        q = trapezoid_weights(y)
        ex -= (q * ex).sum() / q.sum()
        errs.append(np.abs(P[0, 0] - ex).max() / np.abs(ex).max())
        assert np.abs(P[1:]).max() < 1e-16 and np.abs(P[0, 1:]).max() < 1e-16
    assert errs[0] < 5e-4 and errs[1] < 5e-5 and errs[1] < errs[0] / 6, errs


def test_pressure_closed_forms(oracle_mod):
    """Two manufactured cases of Oracle.pressure_of, on every pencil:
      L = (i kx phi, phi', i kz phi), v = 0   ->  P = phi (k = 0: phi minus its mean)
      L = 0, v given                          ->  (D^2 - k^2) P = 0 with P' = nu v''
                                                  at each wall: A cosh ky + B sinh ky
                                                  (k = 0: slope nu v''(-1), gauge mean 0)
    phi differs from pencil to pencil (a swapped kx/kz fails), and v'' differs at
    the two walls (swapped or dropped wall rows fail)."""
    e1s, e2s = [], []
    for NY in (33, 65, 129):
        o = make(oracle_mod, 8, NY, 8)
        y, kx, kz = o.vector("y"), o.vector("kx"), o.vector("kz")
        K = np.sqrt(kz[:, None] ** 2 + kx[None, :] ** 2)
        c = (1 + 0.5j) * (1 + np.arange(o.NZ))[:, None] * (1 + 0.3 * np.arange(o.NXH))[None, :]
        f = np.exp(0.7 * y) * np.cos(1.9 * y + 0.4)
        fp = 0.7 * f - 1.9 * np.exp(0.7 * y) * np.sin(1.9 * y + 0.4)
        phi = c[..., None] * f
        z = np.zeros_like(phi)
        P = o.pressure_of(1j * kx[None, :, None] * phi, c[..., None] * fp, 1j * kz[:, None, None] * phi, z)
        q = trapezoid_weights(y)
        ex = phi.copy()
        ex[0, 0] -= (q * phi[0, 0]).sum() / q.sum()
        e1s.append(np.abs(P - ex).max() / np.abs(ex).max())

        g, gp, gpp = (1 - y ** 2) ** 2, -4 * y * (1 - y ** 2), -4 + 12 * y ** 2
        s, cs = np.sin(y + 0.3), np.cos(y + 0.3)
        vpp = gpp * s + 2 * gp * cs - g * s
        P2 = o.pressure_of(z, z, z, c[..., None] * (g * s))
        g0, g1 = o.nu * vpp[0], o.nu * vpp[-1]
        ex2 = np.empty_like(P2)
        for k in range(o.NZ):
            for i in range(o.NXH):
                kk = K[k, i]
                if kk == 0:
                    ex2[k, i] = c[k, i] * g0 * y
                else:
                    A, B = (g1 - g0) / (2 * kk * np.sinh(kk)), (g1 + g0) / (2 * kk * np.cosh(kk))
                    ex2[k, i] = c[k, i] * (A * np.cosh(kk * y) + B * np.sinh(kk * y))
        e2s.append(np.abs(P2 - ex2).max() / np.abs(ex2).max())
    # fourth order or better in the wall spacing
    assert e1s[0] < 1e-3 and e1s[2] < 3e-6 and e1s[1] < e1s[0] / 10 and e1s[2] < e1s[1] / 10, e1s
    assert e2s[0] < 2e-2 and e2s[2] < 6e-5 and e2s[1] < e2s[0] / 10 and e2s[2] < e2s[1] / 10, e2s


def test_step_preserves_hermitian_plane_and_divergence(oracle_mod):
    o = make(oracle_mod, 16, 33, 16)
    o.init(2, seed=3, amp=0.3)
    o.step(3)
    DY1 = o.matrix("DY1")
    kx, kz = o.vector("kx"), o.vector("kz")
    u, v, w = o.field("u"), o.field("v"), o.field("w")
    for name in ("u", "v", "w", "omx", "omy", "omz", "Hv", "Hg"):
        f = o.field(name)
        scale = np.abs(f).max()
        for m in range(1, o.NZH):
            assert np.abs(f[m, 0] - np.conj(f[o.NZ - m, 0])).max() < 1e-12 * scale, name
    # i kx u + DY1 v + i kz w = 0 on every k^2 > 0 pencil (src/solver.cpp:532-533)
    for k in range(o.NZ):
        for i in range(o.NXH):
            if kx[i] ** 2 + kz[k] ** 2 < 1e-8:
                continue
            d1v = DY1 @ v[k, i]
            div = 1j * kx[i] * u[k, i] + d1v + 1j * kz[k] * w[k, i]
            assert np.abs(div).max() <= 1e-12 * max(np.abs(d1v).max(), 1e-30)


def test_seeded_gauss_is_standard_normal(oracle_mod):
    g = np.array([oracle_mod.seeded_gauss(1, 0, i, m, p) for i in range(20) for m in range(-10, 10) for p in range(4)])
    z = np.concatenate([g.real, g.imag])
    assert abs(z.mean()) < 0.06 and abs(z.var() - 1) < 0.08
    assert len(np.unique(np.round(z, 12))) == len(z)


# ---------------------------------------------------------------------------
# Orr-Sommerfeld (linear dynamics of the whole step)
# ---------------------------------------------------------------------------
def lamb_path_vs_textbook(oracle_mod, N):
    import orr_sommerfeld as os_

    a = os_.ALPHA
    o = make(oracle_mod, 4, N, 4, nu=os_.NU, dkx=a)
    y, DY1, DY2, DY4 = o.vector("y"), o.matrix("DY1"), o.matrix("DY2"), o.matrix("DY4")
    shp = o.wave_shape
    base = np.zeros((6,) + shp, dtype=complex)
    U = os_.UC * (1 - y ** 2)
    base[0, 0, 0] = U
    base[5, 0, 0] = -(DY1 @ U)
    H = np.zeros((N, N), dtype=complex)
    for j in range(N):
        e = np.zeros(N, dtype=complex)
        e[j] = 1.0
        p = np.zeros((6,) + shp, dtype=complex)
        p[1, 0, 1] = e
        p[0, 0, 1] = (1j / a) * (DY1 @ e)
        p[5, 0, 1] = -(1j / a) * (DY2 @ e - a * a * e)
        L = (o.lamb(base + 1e-3 * p) - o.lamb(base - 1e-3 * p)) / 2e-3
        H[:, j] = -a * a * L[1, 0, 1] - DY1 @ (1j * a * L[0, 0, 1])
    # the Lamb path forms U'' v as the commutator Omega_z DY1 v - DY1 (Omega_z v), a
    # dense matrix, so the operators agree in their action on smooth fields and in
    # their eigenvalues, not entry by entry
    T = -1j * a * np.diag(U) @ (DY2 - a * a * np.eye(N)) + 1j * a * (-2 * os_.UC) * np.eye(N)
    A, B = os_.textbook_operator(y, DY1, DY2, DY4)
    A2 = H + os_.NU * (DY4 - 2 * a * a * DY2 + a ** 4 * np.eye(N))
    B2 = DY2 - a * a * np.eye(N)
    os_.boundary_rows(A2, B2, DY1)
    l1, phi = os_.least_stable(A, B)
    l2, _ = os_.least_stable(A2, B2)
    r = np.linalg.norm(((H - T) @ phi)[2:N - 2]) / np.linalg.norm((T @ phi)[2:N - 2])
    return r, abs(l1 - l2) / abs(l1)


def test_orr_sommerfeld_lamb_path_is_the_textbook_operator(oracle_mod):
    """The v equation's explicit operator as the oracle computes it (Lamb vector of
    base + perturbation through the dealiased transforms, then Hv), linearized by
    the exact +-eps difference of a bilinear map, against the textbook
    -i a U (DY2 - a^2) + i a U'' on the same DY matrices. The Lamb path forms U'' v
    as the commutator Omega_z DY1 v - DY1 (Omega_z v), so the two differ by that
    commutator's discretisation error, which falls with NY; a sign or a term wrong
    would leave an O(1) difference."""
    r65, e65 = lamb_path_vs_textbook(oracle_mod, 65)
    r129, e129 = lamb_path_vs_textbook(oracle_mod, 129)
    print(f"[os] action {r65:.2e} {r129:.2e}, eigenvalue {e65:.2e} {e129:.2e}")
    assert r65 < 1e-3 and e65 < 1e-6
    assert r129 < r65 / 4 and e129 < e65 / 8


def test_orr_sommerfeld_orszag_eigenvalue(oracle_mod):
    """With DY4 = DY2^2 the operators reproduce Orszag's eigenvalue (the compact
    DY2 on the stretched grid); the reference's own DY4, whose top-wall rows carry
    the B3 overwrite (src/compdiff.hpp:257-258), moves it by about 1 %."""
    import orr_sommerfeld as os_

    o = make(oracle_mod, 4, 193, 4, nu=os_.NU, dkx=os_.ALPHA)
    y, DY1, DY2, DY4 = o.vector("y"), o.matrix("DY1"), o.matrix("DY2"), o.matrix("DY4")
    c_clean = os_.c_of(os_.least_stable(*os_.textbook_operator(y, DY1, DY2, DY2 @ DY2))[0])
    c_ref = os_.c_of(os_.least_stable(*os_.textbook_operator(y, DY1, DY2, DY4))[0])
    ref = os_.ORSZAG_C
    assert abs(c_clean.real - ref.real) < 1e-5 * ref.real, c_clean
    assert abs(c_clean.imag - ref.imag) < 1e-3 * ref.imag, c_clean
    assert abs(c_ref.real - ref.real) > 5e-3 * ref.real, c_ref


def test_orr_sommerfeld_eigenmode_grows_at_the_discrete_rate(oracle_mod):
    """The oracle's time step from the laminar profile plus a 1e-6 eigenmode: the
    mode grows and turns at the discrete eigenvalue to the scheme's O(dt^3)."""
    import orr_sommerfeld as os_

    N, dt = 65, 0.01
    o = make(oracle_mod, 4, N, 4, nu=os_.NU, dkx=os_.ALPHA, dt=dt)
    y, DY1, DY2, DY4 = o.vector("y"), o.matrix("DY1"), o.matrix("DY2"), o.matrix("DY4")
    lam, phi = os_.least_stable(*os_.textbook_operator(y, DY1, DY2, DY4))
    phi = phi / phi[np.argmax(np.abs(phi))]
    o.set_velocity(*os_.eigenmode_fields(o.wave_shape, y, DY1, phi, 1e-6))
    lm = os_.growth(lambda: o.step(1), lambda: o.field("v")[0, 1], phi, dt, 100, 600)
    assert abs(lm.real - lam.real) < 2e-5 * abs(lam.real), (lm, lam)
    assert abs(lm.imag - lam.imag) < 1e-6 * abs(lam.imag), (lm, lam)


# ---------------------------------------------------------------------------
# Pins of the paths the pins above leave open (tests/pins.py): the omega_y /
# Squire path, Hg, the horizontal recovery's omega_y terms, and the mean-flow
# forcing (friction lag, mean-Lamb history). Each is also run with a planted
# mistake in the oracle (Oracle.set_mutation), which it must detect.
# ---------------------------------------------------------------------------
import pins  # noqa: E402


def test_recovery_identities_after_stepping(oracle_mod):
    """After three nonlinear steps every k^2 > 0 pencil satisfies the four exact
    identities of the horizontal recovery to roundoff (and the mean pencil's
    vorticity is DY1 W, -DY1 U); a flipped omega_y term in u breaks them."""
    for mut, expect_ok in ((0, True), (oracle_mod.Oracle.MUT_U_OMY_SIGN, False)):
        o = make(oracle_mod, 16, 33, 16)
        o.set_mutation(mut)
        o.init(2, seed=3, amp=0.3)
        o.step(3)
        kx, kz = o.vector("kx"), o.vector("kz")
        res = pins.recovery_identities(o, kx, kz, o.matrix("DY1"), o.matrix("DY2"))
        print(f"[pin] recovery identities, mutation {mut}: {res}")
        assert (res.max() < 1e-12) == expect_ok, res


def test_hg_is_the_linearized_squire_forcing(oracle_mod):
    """Hg of laminar flow plus an oblique perturbation (v, omega_y at kx = 0.5,
    kz = 2), linearized by the exact +-eps difference of the bilinear step terms:
    Hg' = -i kz (DY1 U) v - i kx U omega_y (Lx' = -U' v, Lz' = U omega_y through the
    dealiased transforms of a mean profile times one mode: exact to roundoff).
    A flipped Hg (src/solver.cpp:257), or a flipped kx term in it, fails."""
    for mut, expect_ok in ((0, True), (oracle_mod.Oracle.MUT_HG_SIGN, False), (oracle_mod.Oracle.MUT_HG_KX_SIGN, False)):
        o = make(oracle_mod, 8, 33, 8)
        o.set_mutation(mut)
        y, DY1 = o.vector("y"), o.matrix("DY1")
        kx, kz = o.vector("kx"), o.vector("kz")
        U = 1.5 * (1 - y ** 2)
        i, k = 1, 2
        a, b = kx[i], kz[k]
        k2 = a * a + b * b
        vp = (1 - y ** 2) ** 2 * (1 + 0.3j * y + 0.2 * y ** 2)
        gp = (1 - y ** 2) * (0.5 - 0.7j * y + 0.1 * y ** 3)
        d1v = DY1 @ vp
        up = (1j * a / k2) * d1v - (1j * b / k2) * gp
        wp = (1j * a / k2) * gp + (1j * b / k2) * d1v
        eps = 1e-3
        res = []
        for sgn in (1, -1):
            u = np.zeros(o.wave_shape, dtype=complex)
            v = np.zeros_like(u)
            w = np.zeros_like(u)
            u[0, 0] = U
            u[k, i], v[k, i], w[k, i] = sgn * eps * up, sgn * eps * vp, sgn * eps * wp
            o.set_velocity(u, v, w)
            res.append((o.field("Hg")[k, i], o.field("omy")[k, i]))
        hg = (res[0][0] - res[1][0]) / (2 * eps)
        omy = (res[0][1] - res[1][1]) / (2 * eps)
        ref = -1j * b * (DY1 @ U) * vp - 1j * a * U * omy
        err = np.abs(hg - ref).max() / np.abs(ref).max()
        print(f"[pin] Hg linearization, mutation {mut}: {err:.2e}")
        assert (err < 1e-12) == expect_ok, err


def test_squire_mode_follows_the_bdf3_recursion(oracle_mod):
    """A kx = 0 omega_y mode without mean flow (Hg = 0): the implicit omega_y solve
    and the scheme's coefficients give the exact BDF3 recursion at the Dirichlet
    eigenvalue of DY2 (src/solver.cpp:548-556, 578, 603)."""
    N, dt, nu = 65, 0.01, 1 / 2800
    o = make(oracle_mod, 4, N, 8, dt=dt)
    dev = pins.squire_sequence(o, o.wave_shape, o.matrix("DY2"), o.vector("kz")[1], nu, dt, 30)
    print(f"[pin] Squire BDF3 recursion: {dev:.2e}")
    assert dev < 1e-12


@pytest.mark.parametrize("mut", [0, 2, 16, 32])
def test_mean_flow_follows_its_recursion(oracle_mod, mut):
    """The mean flow and the spanwise mean driven by a streamwise vortex's mean Lamb
    vector, step by step against the implicit recursion with the friction and
    mean-Lamb histories (src/solver.cpp:428-514, 306-327, 293-303). The planted
    mistakes (friction sign, mean-Lamb history or friction history not rotated)
    each fail."""
    o = make(oracle_mod, 4, 33, 8, nu=0.02, dt=0.01, dkz=2.0)
    o.set_mutation(mut)
    y = o.vector("y")
    dev, fdev = pins.mean_flow_recursion(o, o.wave_shape, y, o.matrix("DY1"), o.matrix("DY2"), 2.0, 0.02, 0.01, 12)
    print(f"[pin] mean-flow recursion, mutation {mut}: {dev:.2e}, friction {fdev:.2e}")
    assert fdev < 1e-12
    assert (dev < 1e-12) == (mut == 0), dev


@pytest.mark.parametrize("mut", [0, 2])
def test_friction_force_holds_the_flow_rate(oracle_mod, mut):
    """Bulk flow rate over 40 nonlinear steps: drifts by ~1e-8 with the friction
    feedback (the quadrature and DY2's wall rows), by ~1e-4 with its sign flipped."""
    o = make(oracle_mod, 16, 33, 16)
    o.set_mutation(mut)
    o.init(2, seed=3, amp=0.3)
    drift = pins.flow_rate_drift(o, o.vector("y"), 40)
    print(f"[pin] flow-rate drift, mutation {mut}: {drift:.2e}")
    assert (drift < 1e-7) == (mut == 0), drift


@pytest.mark.parametrize("mut", [0, 1, 8])
def test_oblique_orr_sommerfeld_squire_mode(oracle_mod, mut):
    """An oblique mode (kx = 1, kz = 0.5, Re 1e4) of the coupled Orr-Sommerfeld /
    Squire operator built on the same DY matrices: the stepped mode grows and turns
    at the discrete eigenvalue to the scheme's O(dt^3) and keeps the eigenvector's
    omega_y / v ratio; with Hg's sign or its kx term flipped the ratio is off by
    O(1)."""
    nu, dt = 1.5e-4, 0.01
    o = make(oracle_mod, 4, 65, 4, nu=nu, dt=dt, dkx=1.0, dkz=0.5)
    o.set_mutation(mut)
    lm, lam, ratio = pins.oblique_mode(o, o.wave_shape, o.vector("y"), o.matrix("DY1"), o.matrix("DY2"),
                                       o.matrix("DY4"), 1.0, 0.5, nu, dt, 100, 600)
    print(f"[pin] oblique mode, mutation {mut}: measured {lm:.10e} discrete {lam:.10e} ratio {ratio:.8f}")
    ok = abs(ratio - 1) < 1e-5 and abs(lm.real - lam.real) < 2e-5 * abs(lam.real) and \
        abs(lm.imag - lam.imag) < 2e-6 * abs(lam.imag)
    assert ok == (mut == 0), (lm, lam, ratio)


@pytest.mark.parametrize("NX,NY,NZ", [(16, 33, 16), (8, 160, 8)])
def test_factor_cache_is_bit_identical(oracle_mod, NX, NY, NZ):
    """The oracle's factor cache (one factorization per distinct k^2, used for
    the full-size reference run) gives the fields and friction series of
    factoring every pencil at every step, bit for bit."""
    outs = []
    for cache in (False, True):
        o = make(oracle_mod, NX, NY, NZ)
        o.set_factor_cache(cache)
        o.init(2, seed=3, amp=0.15)
        o.step(4)
        outs.append({n: o.field(n) for n in ("u", "v", "w", "omy", "Hv", "Hg")} | {"diag": o.diag()})
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
