// Wall-normal discretisation built on the host in FP64 at setup:
//   tanh-stretched grid and chain-rule metrics   (src/solver.cpp:57-106)
//   compact schemes A_n d_n = B_n f              (src/compdiff.hpp:11-357)
//   dense D_n = A_n^{-1} B_n and DY_n            (src/solver.cpp:171-216)
// The schemes are written as row-stencil tables; every row type of the
// reference appears once, including the two top-wall rows that the reference
// overwrites (B2 row N-1, harmless; B3 row N-1, which loses its fifth
// coefficient and keeps -e0/h^3 in place of -d0/h^3 at column N-4).
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "owt_internal.hpp"

namespace owt {

// LU with partial pivoting and the pivot/scaling rules of LAPACK dgetrf
// (src/linalg.hpp:508-520): first maximal |a| as pivot, whole-row swaps,
// multipliers scaled by the reciprocal pivot, updates accumulated in order.
int host_lu_factor(int n, double *A, int *ipiv) {
    int info = 0;
    auto at = [&](int i, int j) -> double & { return A[(size_t)i + (size_t)j * n]; };
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::fabs(at(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(at(i, k)) > best) best = std::fabs(at(i, k)), p = i;
        ipiv[k] = p;
        if (at(p, k) == 0.0) {
            if (!info) info = k + 1;
            continue;
        }
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(at(k, j), at(p, j));
        const double piv = at(k, k);
        if (std::fabs(piv) >= DBL_MIN) {
            const double r = 1.0 / piv;
            for (int i = k + 1; i < n; ++i) at(i, k) *= r;
        } else {
            for (int i = k + 1; i < n; ++i) at(i, k) /= piv;
        }
        for (int j = k + 1; j < n; ++j) {
            const double t = at(k, j);
            if (t == 0.0) continue;
            for (int i = k + 1; i < n; ++i) at(i, j) -= at(i, k) * t;
        }
    }
    return info;
}

// dgetrs 'N' for one right-hand side (src/linalg.hpp:522-532)
void host_lu_solve(int n, const double *LU, const int *ipiv, double *b) {
    for (int k = 0; k < n; ++k)
        if (ipiv[k] != k) std::swap(b[k], b[ipiv[k]]);
    for (int k = 0; k < n; ++k) {
        if (b[k] == 0.0) continue;
        for (int i = k + 1; i < n; ++i) b[i] -= b[k] * LU[i + (size_t)k * n];
    }
    for (int k = n - 1; k >= 0; --k) {
        if (b[k] == 0.0) continue;
        b[k] = b[k] / LU[k + (size_t)k * n];
        for (int i = 0; i < k; ++i) b[i] -= b[k] * LU[i + (size_t)k * n];
    }
}

namespace {

// One stencil entry: matrix column = row + off (or an absolute column when abs_col >= 0).
struct Tap {
    int off;
    double c;
};

struct Builder {
    int N;
    std::vector<double> &M;
    void set(int r, int c, double v) { M[(size_t)r + (size_t)c * N] = v; }
    void row(int r, std::initializer_list<Tap> taps) {
        for (const Tap &t : taps) set(r, r + t.off, t.c);
    }
};

// first derivative (src/compdiff.hpp:11-69)
void scheme1(int N, double h, std::vector<double> &A, std::vector<double> &B) {
    Builder a{N, A}, b{N, B};
    const double al = 4.0 / 9.0, be = 1.0 / 36.0, ca = 40.0 / 27.0, cb = 25.0 / 54.0;
    for (int i = 2; i < N - 2; ++i) {
        a.row(i, {{-2, be}, {-1, al}, {0, 1.0}, {1, al}, {2, be}});
        b.row(i, {{-2, -cb / (4 * h)}, {-1, -ca / (2 * h)}, {1, ca / (2 * h)}, {2, cb / (4 * h)}});
    }
    for (int i : {1, N - 2}) {
        a.row(i, {{-1, 0.25}, {0, 1.0}, {1, 0.25}});
        b.row(i, {{-1, -1.5 / (2 * h)}, {1, 1.5 / (2 * h)}});
    }
    const double a0 = -17.0 / 6.0, b0 = 1.5, c0 = 1.5, d0 = -1.0 / 6.0;
    a.row(0, {{0, 1.0}, {1, 3.0}});
    a.row(N - 1, {{0, 1.0}, {-1, 3.0}});
    b.row(0, {{0, a0 / h}, {1, b0 / h}, {2, c0 / h}, {3, d0 / h}});
    b.row(N - 1, {{0, -a0 / h}, {-1, -b0 / h}, {-2, -c0 / h}, {-3, -d0 / h}});
}

// second derivative (src/compdiff.hpp:71-166); rows 1 and N-2 are centred on
// nodes 2 and N-3
void scheme2(int N, double h, std::vector<double> &A, std::vector<double> &B) {
    Builder a{N, A}, b{N, B};
    const double h2 = std::pow(h, 2);
    const double al = 344.0 / 1179.0, be = 23.0 / 2358.0, ca = 320.0 / 393.0, cb = 310.0 / 393.0;
    for (int i = 2; i < N - 2; ++i) {
        a.row(i, {{-2, be}, {-1, al}, {0, 1.0}, {1, al}, {2, be}});
        b.row(i, {{-2, cb / (4 * h2)}, {-1, ca / h2}, {0, -2 * cb / (4 * h2) - 2 * ca / h2}, {1, ca / h2}, {2, cb / (4 * h2)}});
    }
    const double al1 = 2.0 / 11.0, a1 = 12.0 / 11.0, b1 = 3.0 / 11.0;
    for (int i : {1, N - 2}) {
        int s = (i == 1) ? 1 : -1;  // shift of the stencil centre
        a.row(i, {{s - 1, al1}, {s, 1.0}, {s + 1, al1}});
        b.row(i, {{s - 2, b1 / (4 * h2)}, {s - 1, a1 / h2}, {s, -2 * b1 / (4 * h2) - 2 * a1 / h2}, {s + 1, a1 / h2}, {s + 2, b1 / (4 * h2)}});
    }
    a.row(0, {{0, 1.0}, {1, 10.0}, {2, 1.0}});
    a.row(N - 1, {{0, 1.0}, {-1, 10.0}, {-2, 1.0}});
    b.row(0, {{0, 12.0 / h2}, {1, -24.0 / h2}, {2, 12.0 / h2}, {3, 0.0 / h2}, {4, 0.0 / h2}});
    // top row: the fourth coefficient is written twice in the reference, both zero
    b.row(N - 1, {{0, 12.0 / h2}, {-1, -24.0 / h2}, {-2, 12.0 / h2}, {-3, 0.0 / h2}});
}

// third derivative (src/compdiff.hpp:168-259)
void scheme3(int N, double h, std::vector<double> &A, std::vector<double> &B) {
    Builder a{N, A}, b{N, B};
    const double h3 = std::pow(h, 3);
    const double al = 147.0 / 332.0, be = 1.0 / 166.0, ca = 160.0 / 83.0, cb = -5.0 / 166.0;
    for (int i = 3; i < N - 3; ++i) {
        a.row(i, {{-2, be}, {-1, al}, {0, 1.0}, {1, al}, {2, be}});
        b.row(i, {{-3, -cb / (8 * h3)}, {-2, -ca / (2 * h3)}, {-1, ca / h3 + 3 * cb / (8 * h3)},
                  {1, -ca / h3 - 3 * cb / (8 * h3)}, {2, ca / (2 * h3)}, {3, cb / (8 * h3)}});
    }
    const double al2 = 4.0 / 9.0, be2 = 1.0 / 126.0, a2 = 40.0 / 21.0;
    for (int i : {2, N - 3}) {
        a.row(i, {{-2, be2}, {-1, al2}, {0, 1.0}, {1, al2}, {2, be2}});
        b.row(i, {{-2, -a2 / (2 * h3)}, {-1, a2 / h3}, {1, -a2 / h3}, {2, a2 / (2 * h3)}});
    }
    // row 1 centred on node 2, row N-2 on node N-3
    a.row(1, {{0, 0.5}, {1, 1.0}, {2, 0.5}});
    a.row(N - 2, {{-2, 0.5}, {-1, 1.0}, {0, 0.5}});
    b.row(1, {{-1, -2.0 / (2 * h3)}, {0, 2.0 / h3}, {2, -2.0 / h3}, {3, 2.0 / (2 * h3)}});
    b.row(N - 2, {{-3, -2.0 / (2 * h3)}, {-2, 2.0 / h3}, {0, -2.0 / h3}, {1, 2.0 / (2 * h3)}});
    a.row(0, {{0, 1.0}, {1, -7.0}});
    a.row(N - 1, {{0, 1.0}, {-1, -7.0}});
    b.row(0, {{0, 8.0 / h3}, {1, -26.0 / h3}, {2, 30.0 / h3}, {3, -14.0 / h3}, {4, 2.0 / h3}});
    // top row as the reference leaves it: column N-4 holds -e0/h^3, column N-5 is not set
    b.row(N - 1, {{0, -8.0 / h3}, {-1, 26.0 / h3}, {-2, -30.0 / h3}, {-3, -2.0 / h3}});
}

// fourth derivative (src/compdiff.hpp:261-357)
void scheme4(int N, double h, std::vector<double> &A, std::vector<double> &B) {
    Builder a{N, A}, b{N, B};
    const double h4 = std::pow(h, 4);
    const double al = 7.0 / 26.0, ca = 19.0 / 13.0, cb = 1.0 / 13.0;
    for (int i = 3; i < N - 3; ++i) {
        a.row(i, {{-1, al}, {0, 1.0}, {1, al}});
        b.row(i, {{-3, cb / (6 * h4)}, {-2, ca / h4}, {-1, -4 * ca / h4 - 9 * cb / (6 * h4)}, {0, 6 * ca / h4 + 16 * cb / (6 * h4)},
                  {1, -4 * ca / h4 - 9 * cb / (6 * h4)}, {2, ca / h4}, {3, cb / (6 * h4)}});
    }
    for (int i : {2, N - 3}) {
        a.row(i, {{-1, 0.25}, {0, 1.0}, {1, 0.25}});
        b.row(i, {{-2, 1.5 / h4}, {-1, -4 * 1.5 / h4}, {0, 6 * 1.5 / h4}, {1, -4 * 1.5 / h4}, {2, 1.5 / h4}});
    }
    a.row(1, {{-1, 1.0}, {0, -2.0}, {1, 7.0}});
    b.row(1, {{-1, 6.0 / h4}, {0, -24.0 / h4}, {1, 36.0 / h4}, {2, -24.0 / h4}, {3, 6.0 / h4}});
    a.row(N - 2, {{1, 1.0}, {0, -2.0}, {-1, 7.0}});
    b.row(N - 2, {{1, 6.0 / h4}, {0, -24.0 / h4}, {-1, 36.0 / h4}, {-2, -24.0 / h4}, {-3, 6.0 / h4}});
    a.row(0, {{0, 1.0}, {1, -2.0}});
    b.row(0, {{0, -1.0 / h4}, {1, 4.0 / h4}, {2, -6.0 / h4}, {3, 4.0 / h4}, {4, -1.0 / h4}});
    a.row(N - 1, {{0, 1.0}, {-1, -2.0}});
    b.row(N - 1, {{0, -1.0 / h4}, {-1, 4.0 / h4}, {-2, -6.0 / h4}, {-3, 4.0 / h4}, {-4, -1.0 / h4}});
}

std::vector<double> solve_dense(int N, const std::vector<double> &A, const std::vector<double> &B) {
    std::vector<double> LU(A), X(B);
    std::vector<int> ipiv(N);
    if (host_lu_factor(N, LU.data(), ipiv.data()) != 0)
        throw Error(4, "LU factorization failed. Matrix A may be singular.");
    for (int j = 0; j < N; ++j) host_lu_solve(N, LU.data(), ipiv.data(), X.data() + (size_t)j * N);
    return X;
}

// sum_k diag(c_k) D_k, accumulated left to right as the reference's expressions are
std::vector<double> combine(int N, std::initializer_list<std::pair<const std::vector<double> *, const std::vector<double> *>> terms) {
    std::vector<double> R((size_t)N * N, 0.0);
    bool first = true;
    for (auto &t : terms) {
        const std::vector<double> &c = *t.first, &D = *t.second;
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                double v = c[i] * D[(size_t)i + (size_t)j * N];
                R[(size_t)i + (size_t)j * N] = first ? v : R[(size_t)i + (size_t)j * N] + v;
            }
        first = false;
    }
    return R;
}

}  // namespace

Operators build_operators(int N, double alpha) {
    Operators op;
    op.N = N;
    op.ds = 2.0 / (N - 1.0);
    auto sz = [&](std::vector<double> &v) { v.assign(N, 0.0); };
    for (auto *v : {&op.s, &op.y, &op.dyds1, &op.dyds2, &op.dyds3, &op.dyds4, &op.c11, &op.c21, &op.c22, &op.c31, &op.c32,
                    &op.c33, &op.c41, &op.c42, &op.c43, &op.c44})
        sz(*v);
    // y = tanh(alpha s)/tanh(alpha) and its s-derivatives (src/solver.cpp:71-81)
    const double ap = alpha, ap2 = std::pow(alpha, 2), ap3 = std::pow(alpha, 3), ap4 = std::pow(alpha, 4);
    const double cth = 1.0 / std::tanh(ap);
    for (int i = 0; i < N; ++i) {
        op.s[i] = (double)(2 * (size_t)i) / (N - 1.0) - 1.0;
        const double as = ap * op.s[i];
        const double sech = 1.0 / std::cosh(as);
        op.y[i] = std::tanh(as) / std::tanh(ap);
        op.dyds1[i] = ap * cth * std::pow(sech, 2);
        op.dyds2[i] = -2 * ap2 * cth * std::pow(sech, 2) * std::tanh(as);
        op.dyds3[i] = 2 * ap3 * (-2 + std::cosh(2 * as)) * cth * std::pow(sech, 4);
        op.dyds4[i] = -2 * ap4 * cth * std::pow(sech, 5) * (-11 * std::sinh(as) + std::sinh(3 * as));
    }
    // chain rule d^n/dy^n = sum_k c_nk d^k/ds^k (src/solver.cpp:94-106)
    for (int i = 0; i < N; ++i) {
        const double d1 = op.dyds1[i], d2 = op.dyds2[i], d3 = op.dyds3[i], d4 = op.dyds4[i];
        auto p = [&](int e) { return std::pow(d1, (double)e); };
        op.c11[i] = p(-1);
        op.c21[i] = (-d2) * p(-3);
        op.c22[i] = p(-2);
        op.c31[i] = (-d3) * p(-4) + (3.0 * std::pow(d2, 2.0)) * p(-5);
        op.c32[i] = (-3.0 * d2) * p(-4);
        op.c33[i] = p(-3);
        op.c41[i] = ((-d4) * p(-5) + ((10.0 * d3) * d2) * p(-6)) - (15.0 * std::pow(d2, 3.0)) * p(-7);
        op.c42[i] = (-4.0 * d3) * p(-5) + (15.0 * std::pow(d2, 2.0)) * p(-6);
        op.c43[i] = (-6.0 * d2) * p(-5);
        op.c44[i] = p(-4);
    }
    for (int n = 0; n < 4; ++n) {
        op.A[n].assign((size_t)N * N, 0.0);
        op.B[n].assign((size_t)N * N, 0.0);
    }
    scheme1(N, op.ds, op.A[0], op.B[0]);
    scheme2(N, op.ds, op.A[1], op.B[1]);
    scheme3(N, op.ds, op.A[2], op.B[2]);
    scheme4(N, op.ds, op.A[3], op.B[3]);
    for (int n = 0; n < 4; ++n) op.D[n] = solve_dense(N, op.A[n], op.B[n]);
    op.DY1 = combine(N, {{&op.c11, &op.D[0]}});
    op.DY2 = combine(N, {{&op.c21, &op.D[0]}, {&op.c22, &op.D[1]}});
    op.DY3 = combine(N, {{&op.c31, &op.D[0]}, {&op.c32, &op.D[1]}, {&op.c33, &op.D[2]}});
    op.DY4 = combine(N, {{&op.c41, &op.D[0]}, {&op.c42, &op.D[1]}, {&op.c43, &op.D[2]}, {&op.c44, &op.D[3]}});
    return op;
}

std::vector<double> mirror_halves(const std::vector<double> &M, int N, int s) {
    const int h = (N + 1) / 2;
    auto at = [&](int i, int j) { return (long double)M[(size_t)j * N + i]; };
    // rows of the symmetrised operator
    auto ms = [&](int i, int j) { return (at(i, j) + (long double)s * at(N - 1 - i, N - 1 - j)) / 2; };
    std::vector<double> out(2 * (size_t)h * h);
    double *A = out.data(), *B = out.data() + (size_t)h * h;
    for (int j = 0; j < h; ++j)
        for (int i = 0; i < h; ++i) {
            const size_t e = (size_t)j * h + i;
            if (2 * j == N - 1) {
                A[e] = (double)(ms(i, j) / 2);
                B[e] = 0.0;
            } else {
                A[e] = (double)((ms(i, j) + ms(i, N - 1 - j)) / 2);
                B[e] = (double)((ms(i, j) - ms(i, N - 1 - j)) / 2);
            }
        }
    return out;
}

}  // namespace owt

namespace owt {

// E = -nu (DY4 - J DY4 J) / 2 with rows 0, 1, N-2, N-1 zero (those rows of Av
// are the boundary rows, mirror images of each other once row N-2 changes
// sign). The reference's DY4 top-row overwrite makes E rank 2; its factors
// come from two steps of Gram-Schmidt with column pivoting in long double.
bool mirror_lowrank(const std::vector<double> &DY4, int N, double nu, std::vector<double> &UV) {
    if (N < 8) return false;
    auto at = [&](int i, int j) { return (long double)DY4[(size_t)j * N + i]; };
    std::vector<long double> E((size_t)N * N, 0.0L);
    long double emax = 0;
    for (int j = 0; j < N; ++j)
        for (int i = 2; i < N - 2; ++i) {
            const long double e = -(long double)nu * (at(i, j) - at(N - 1 - i, N - 1 - j)) / 2;
            E[(size_t)j * N + i] = e;
            emax = std::max(emax, std::fabs(e));
        }
    if (emax == 0) return false;
    std::vector<long double> R = E, Q(2 * (size_t)N);
    for (int m = 0; m < 2; ++m) {
        int best = -1;
        long double bn = 0;
        for (int j = 0; j < N; ++j) {
            long double s = 0;
            for (int i = 0; i < N; ++i) s += R[(size_t)j * N + i] * R[(size_t)j * N + i];
            if (s > bn) bn = s, best = j;
        }
        if (best < 0) return false;
        const long double nrm = std::sqrt(bn);
        for (int i = 0; i < N; ++i) Q[(size_t)m * N + i] = R[(size_t)best * N + i] / nrm;
        for (int j = 0; j < N; ++j) {
            long double d = 0;
            for (int i = 0; i < N; ++i) d += Q[(size_t)m * N + i] * R[(size_t)j * N + i];
            for (int i = 0; i < N; ++i) R[(size_t)j * N + i] -= d * Q[(size_t)m * N + i];
        }
    }
    // what the two directions leave must be rounding
    long double rmax = 0;
    for (long double r : R) rmax = std::max(rmax, std::fabs(r));
    if (rmax > 1e-12L * emax) return false;
    // U = Q (orthonormal columns), V^T = U^T E
    UV.assign(4 * (size_t)N, 0.0);
    for (int m = 0; m < 2; ++m) {
        for (int i = 0; i < N; ++i) UV[(size_t)m * N + i] = (double)Q[(size_t)m * N + i];
        for (int j = 0; j < N; ++j) {
            long double d = 0;
            for (int i = 0; i < N; ++i) d += Q[(size_t)m * N + i] * E[(size_t)j * N + i];
            UV[(size_t)(2 + m) * N + j] = (double)d;
        }
    }
    return true;
}

}  // namespace owt
