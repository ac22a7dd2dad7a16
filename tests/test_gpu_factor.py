"""LU factors of the implicit operators computed on the GPU (lu_kernels.hip,
k_factor) against the oracle's dgetrf restatement of the same matrices.

The operators are assembled here from the library's own DY matrices with the
element expressions of src/solver.cpp:577-591 (Av) and 601-603 (Ag), one
IEEE rounding per operation as numpy float64 does. The pivot rows must agree
exactly (LAPACK's rule: first largest |a| in the interchanged order), and so
must the factors, bit for bit: the GPU rounds every product and difference of
the elimination separately and applies each element's updates in ascending
pivot step, as the reference BLAS does. Above 256 rows the factorization runs
by panels (k_factor_panel); forced at 193 rows it gives the register kernel's
factors exactly.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G0 = 11.0 / 6.0
RE180 = dict(nu=1 / 2800, dt=1e-3, dkx=0.5, dkz=1.0, alpha_mesh=1.5)


def assemble(op, k2, nu, dt, DY1, DY2, DY4):
    N = DY2.shape[0]
    A = np.zeros((N, N))
    if op == 0:
        s1 = G0 / dt + (2.0 * nu) * k2
        s2 = nu * (k2 * k2) + (G0 * k2) / dt
        A[2:N - 2] = (-nu) * DY4[2:N - 2] + s1 * DY2[2:N - 2]
        for r in range(2, N - 2):
            A[r, r] = A[r, r] - s2
        A[1] = DY1[0]
        A[N - 2] = DY1[N - 1]
    else:
        sg = nu * k2 + G0 / dt
        A[1:N - 1] = (-nu) * DY2[1:N - 1]
        for r in range(1, N - 1):
            A[r, r] = A[r, r] + sg
    A[0, 0] = 1.0
    A[N - 1, N - 1] = 1.0
    return A


def ipiv_to_perm(ipiv):
    perm = np.arange(len(ipiv))
    for k, p in enumerate(ipiv):
        perm[k], perm[p] = perm[p], perm[k]
    return perm


@pytest.mark.parametrize("shape", [(8, 17, 8), (16, 65, 16), (8, 129, 8), (8, 193, 8), (8, 257, 8), (8, 288, 8)])
def test_factor_matches_oracle(owt_mod, oracle_mod, shape):
    NX, NY, NZ = shape
    p = dict(RE180, NX=NX, NY=NY, NZ=NZ, precision=0, method=1)
    print(f"[gpu] factor {shape}", flush=True)
    ch = owt_mod.Channel(params=p)
    DY1, DY2, DY4 = ch.matrix("DY1"), ch.matrix("DY2"), ch.matrix("DY4")
    NXH, NZH = NX // 2, NZ // 2
    kx = np.arange(NXH) * p["dkx"]
    kz = np.arange(NZH + 1) * p["dkz"]
    kz[NZH] = -kz[NZH]  # the z-Nyquist pencil carries kz = -NZH dkz
    systems = range(NXH * (NZH + 1))
    if NY > 100:
        systems = [0, 1, NXH, NXH + 3, NXH * NZH + NXH - 1]
    for sys in systems:
        mz, i = divmod(sys, NXH)
        k2 = kx[i] * kx[i] + kz[mz] * kz[mz]
        for op in (0, 1):
            A = assemble(op, k2, p["nu"], p["dt"], DY1, DY2, DY4)
            LU, ipiv, info = oracle_mod.lu_factor(A)
            assert info == 0
            lu, perm = ch.factor(sys, op)
            assert np.array_equal(perm, ipiv_to_perm(ipiv)), (sys, op)
            assert np.array_equal(lu, LU), (sys, op, np.abs(lu - LU).max())
            # and the factorization itself: L U = A with rows in pivot order
            L = np.tril(lu, -1) + np.eye(NY)
            U = np.triu(lu)
            res = np.abs(L @ U - A[perm]).max() / np.abs(A).max()
            assert res < 1e-13, (sys, op, res)
    ch.close()


def test_panel_factorization_equals_register_factorization(owt_mod):
    """OWT_FACTOR_PANEL=1 forces the panel kernel; at 193 rows both kernels run."""
    p = dict(RE180, NX=16, NY=193, NZ=16, precision=0, method=1)
    print("[gpu] factor panel vs register 16x193x16", flush=True)
    reg = owt_mod.Channel(params=p)
    os.environ["OWT_FACTOR_PANEL"] = "1"
    try:
        pan = owt_mod.Channel(params=p)
    finally:
        del os.environ["OWT_FACTOR_PANEL"]
    for sys in range(8 * 9):
        for op in (0, 1):
            a, pa = reg.factor(sys, op)
            b, pb = pan.factor(sys, op)
            assert np.array_equal(pa, pb) and np.array_equal(a, b), (sys, op)
    for ch in (reg, pan):
        ch.init_fields(2, seed=1, amp=0.15)
        ch.step(3)
    for name in ("u", "v", "w"):
        assert np.array_equal(reg.field(name), pan.field(name)), name
    reg.close()
    pan.close()
