// Eigendecomposition of the interior block of DY2, used to solve the omega_y
// Helmholtz systems of every k^2 with two shared-matrix products:
//
//   Ag(k^2) = -nu DY2 + (nu k^2 + gamma0/dt) I with Dirichlet rows 0, N-1
//   (src/solver.cpp:578, 601-603). Rows 0 and N-1 fix omega_y = 0 there, so the
//   interior unknowns satisfy (c I - nu D) x = r, D = DY2[1..N-2, 1..N-2],
//   c = nu k^2 + gamma0/dt. With D = Q B Q^-1, B block diagonal,
//   x = Q (c I - nu B)^-1 Q^-1 r.
//
// B has 1x1 blocks for real eigenvalues and 2x2 blocks [[a, b], [-b, a]] for the
// pairs a +- i b (columns x_r, x_i of the eigenvector x_r + i x_i). The pairs come
// first, at rows (2m, 2m+1). Computed in long double: balancing (Parlett and
// Reinsch, powers of two), reduction to Hessenberg form by Householder
// reflections, eigenvalues by the Francis double-shift QR iteration with
// deflation (Golub and Van Loan, Matrix Computations, Algorithms 7.4.2 and 7.5.1-2),
// eigenvectors by inverse iteration on the Hessenberg form, Q^-1 by LU with
// partial pivoting. The
// result is checked: |D Q - Q B| and |Q Q^-1 - I| must be at rounding level.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "owt_internal.hpp"

namespace owt {

namespace {

using ld = long double;
using cld = std::complex<long double>;

// A(i, j) of an n x n row-major long double matrix
struct Mat {
    int n;
    std::vector<ld> a;
    explicit Mat(int n_) : n(n_), a((size_t)n_ * n_, 0.0L) {}
    ld &operator()(int i, int j) { return a[(size_t)i * n + j]; }
    ld operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

// Parlett-Reinsch balancing by powers of two: A <- S^-1 A S, returns S
std::vector<ld> balance(Mat &A) {
    const int n = A.n;
    std::vector<ld> s(n, 1.0L);
    bool done = false;
    while (!done) {
        done = true;
        for (int i = 0; i < n; ++i) {
            ld c = 0, r = 0;
            for (int j = 0; j < n; ++j)
                if (j != i) {
                    c += std::fabs(A(j, i));
                    r += std::fabs(A(i, j));
                }
            if (c == 0 || r == 0) continue;
            const ld tot = c + r;
            ld f = 1;
            while (c < r / 2) {
                c *= 4;
                f *= 2;
            }
            while (c > r * 2) {
                c /= 4;
                f /= 2;
            }
            if ((c + r) / f < 0.95L * tot) {
                done = false;
                s[i] *= f;
                for (int j = 0; j < n; ++j) A(i, j) /= f;
                for (int j = 0; j < n; ++j) A(j, i) *= f;
            }
        }
    }
    return s;
}

// Householder reduction to upper Hessenberg form: H = Z^T A Z with Z orthogonal
// (accumulated, row-major), so that an eigenvector y of H gives Z y of A
void hessenberg(Mat &H, Mat &Z) {
    const int n = H.n;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) Z(i, j) = i == j ? 1 : 0;
    std::vector<ld> v(n);
    for (int k = 0; k + 2 < n; ++k) {
        ld alpha = 0;
        for (int i = k + 1; i < n; ++i) alpha += H(i, k) * H(i, k);
        alpha = std::sqrt(alpha);
        if (alpha == 0) continue;
        if (H(k + 1, k) > 0) alpha = -alpha;
        for (int i = 0; i < n; ++i) v[i] = 0;
        v[k + 1] = H(k + 1, k) - alpha;
        for (int i = k + 2; i < n; ++i) v[i] = H(i, k);
        ld vv = 0;
        for (int i = k + 1; i < n; ++i) vv += v[i] * v[i];
        if (vv == 0) continue;
        // H <- P H P, Z <- Z P with P = I - 2 v v^T / vv
        for (int j = 0; j < n; ++j) {
            ld d = 0;
            for (int i = k + 1; i < n; ++i) d += v[i] * H(i, j);
            d = 2 * d / vv;
            for (int i = k + 1; i < n; ++i) H(i, j) -= d * v[i];
        }
        for (int i = 0; i < n; ++i) {
            ld d = 0, e = 0;
            for (int j = k + 1; j < n; ++j) {
                d += H(i, j) * v[j];
                e += Z(i, j) * v[j];
            }
            d = 2 * d / vv;
            e = 2 * e / vv;
            for (int j = k + 1; j < n; ++j) {
                H(i, j) -= d * v[j];
                Z(i, j) -= e * v[j];
            }
        }
        for (int i = k + 2; i < n; ++i) H(i, k) = 0;
    }
}

// (H - s I) y = b for an upper Hessenberg H: LU with partial pivoting between
// neighbouring rows (the only candidates), O(n^2)
void hess_solve(const Mat &H, cld s, std::vector<cld> &b) {
    const int n = H.n;
    std::vector<cld> a((size_t)n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = (j >= i - 1) ? H(i, j) - (i == j ? s : cld(0)) : cld(0);
    const ld tiny = std::numeric_limits<ld>::epsilon() * std::numeric_limits<ld>::epsilon();
    for (int k = 0; k + 1 < n; ++k) {
        cld *rk = &a[(size_t)k * n], *rn = &a[(size_t)(k + 1) * n];
        if (std::abs(rn[k]) > std::abs(rk[k])) {
            for (int j = k; j < n; ++j) std::swap(rk[j], rn[j]);
            std::swap(b[k], b[k + 1]);
        }
        if (std::abs(rk[k]) < tiny) rk[k] = tiny;
        const cld l = rn[k] / rk[k];
        for (int j = k + 1; j < n; ++j) rn[j] -= l * rk[j];
        b[k + 1] -= l * b[k];
    }
    if (std::abs(a[(size_t)(n - 1) * n + n - 1]) < tiny) a[(size_t)(n - 1) * n + n - 1] = tiny;
    for (int i = n - 1; i >= 0; --i) {
        cld t = b[i];
        for (int j = i + 1; j < n; ++j) t -= a[(size_t)i * n + j] * b[j];
        b[i] = t / a[(size_t)i * n + i];
    }
}

// Householder vector of (x, y, z) (m = 3) or (x, y) (m = 2): P = I - beta u u^T
// maps it to a multiple of e1; u[0] = 1
void house(int m, ld x, ld y, ld z, ld u[3], ld &beta) {
    const ld sigma = y * y + (m == 3 ? z * z : 0);
    u[0] = 1;
    u[1] = y;
    u[2] = m == 3 ? z : 0;
    if (sigma == 0) {
        beta = 0;
        return;
    }
    const ld mu = std::sqrt(x * x + sigma);
    const ld v0 = x <= 0 ? x - mu : -sigma / (x + mu);
    beta = 2 * v0 * v0 / (sigma + v0 * v0);
    u[1] = y / v0;
    u[2] = m == 3 ? z / v0 : 0;
}

// eigenvalues of the trailing 2 x 2 block [[p, q], [r, s]]
void eig2(ld p, ld q, ld r, ld s, cld &e1, cld &e2) {
    const ld tr = (p + s) / 2, dd = (p - s) / 2;
    const ld disc = dd * dd + q * r;
    if (disc >= 0) {
        const ld sq = std::sqrt(disc);
        const ld a = tr + (tr >= 0 ? sq : -sq);
        const ld det = p * s - q * r;
        e1 = a;
        e2 = a != 0 ? det / a : tr - (tr >= 0 ? sq : -sq);
    } else {
        const ld sq = std::sqrt(-disc);
        e1 = cld(tr, sq);
        e2 = cld(tr, -sq);
    }
}

// Francis double-shift QR on an upper Hessenberg matrix: all eigenvalues
std::vector<cld> francis(Mat H) {
    const int n = H.n;
    std::vector<cld> ev;
    const ld eps = std::numeric_limits<ld>::epsilon();
    int hi = n - 1, its = 0;
    while (hi >= 0) {
        // deflation: the last negligible subdiagonal at or below hi
        int lo = hi;
        while (lo > 0) {
            ld s = std::fabs(H(lo - 1, lo - 1)) + std::fabs(H(lo, lo));
            if (std::fabs(H(lo, lo - 1)) <= eps * s) {
                H(lo, lo - 1) = 0;
                break;
            }
            --lo;
        }
        if (lo == hi) {
            ev.push_back(H(hi, hi));
            --hi;
            its = 0;
            continue;
        }
        if (lo == hi - 1) {
            cld e1, e2;
            eig2(H(hi - 1, hi - 1), H(hi - 1, hi), H(hi, hi - 1), H(hi, hi), e1, e2);
            ev.push_back(e1);
            ev.push_back(e2);
            hi -= 2;
            its = 0;
            continue;
        }
        if (++its > 60) throw Error(2, "eigendecomposition of DY2: QR iteration did not converge");
        // shifts: the eigenvalues of the trailing 2 x 2 block (exceptional shifts
        // every 10 iterations without deflation)
        ld s, t;
        if (its % 10 == 0) {
            const ld w = std::fabs(H(hi, hi - 1)) + std::fabs(H(hi - 1, hi - 2));
            s = 1.5L * w;
            t = w * w;
        } else {
            s = H(hi - 1, hi - 1) + H(hi, hi);
            t = H(hi - 1, hi - 1) * H(hi, hi) - H(hi - 1, hi) * H(hi, hi - 1);
        }
        // first column of (H - s1)(H - s2) restricted to the active block
        ld x = H(lo, lo) * H(lo, lo) + H(lo, lo + 1) * H(lo + 1, lo) - s * H(lo, lo) + t;
        ld y = H(lo + 1, lo) * (H(lo, lo) + H(lo + 1, lo + 1) - s);
        ld z = H(lo + 1, lo) * H(lo + 2, lo + 1);
        for (int k = lo - 1; k <= hi - 3; ++k) {
            ld u[3], beta;
            house(3, x, y, z, u, beta);
            const int r0 = k + 1;
            const int c0 = std::max(lo, k);
            for (int j = c0; j <= hi; ++j) {
                const ld d = beta * (H(r0, j) + u[1] * H(r0 + 1, j) + u[2] * H(r0 + 2, j));
                H(r0, j) -= d;
                H(r0 + 1, j) -= d * u[1];
                H(r0 + 2, j) -= d * u[2];
            }
            const int rend = std::min(k + 4, hi);
            for (int i = lo; i <= rend; ++i) {
                const ld d = beta * (H(i, r0) + u[1] * H(i, r0 + 1) + u[2] * H(i, r0 + 2));
                H(i, r0) -= d;
                H(i, r0 + 1) -= d * u[1];
                H(i, r0 + 2) -= d * u[2];
            }
            x = H(k + 2, k + 1);
            y = H(k + 3, k + 1);
            z = k < hi - 3 ? H(k + 4, k + 1) : 0;
        }
        ld u[3], beta;
        house(2, x, y, 0, u, beta);
        const int r0 = hi - 1;
        for (int j = std::max(lo, hi - 2); j <= hi; ++j) {
            const ld d = beta * (H(r0, j) + u[1] * H(r0 + 1, j));
            H(r0, j) -= d;
            H(r0 + 1, j) -= d * u[1];
        }
        for (int i = lo; i <= hi; ++i) {
            const ld d = beta * (H(i, r0) + u[1] * H(i, r0 + 1));
            H(i, r0) -= d;
            H(i, r0 + 1) -= d * u[1];
        }
        for (int i = lo + 2; i <= hi; ++i)
            for (int j = lo; j < i - 1; ++j) H(i, j) = 0;
    }
    return ev;
}

// complex LU with partial pivoting, in place; solve for one right-hand side
struct CLU {
    int n;
    std::vector<cld> a;
    std::vector<int> piv;
    CLU(const Mat &A, cld shift) : n(A.n), a((size_t)A.n * A.n), piv(A.n) {
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = A(i, j) - (i == j ? shift : cld(0));
        for (int k = 0; k < n; ++k) {
            int p = k;
            for (int i = k + 1; i < n; ++i)
                if (std::abs(a[(size_t)i * n + k]) > std::abs(a[(size_t)p * n + k])) p = i;
            piv[k] = p;
            if (p != k)
                for (int j = 0; j < n; ++j) std::swap(a[(size_t)k * n + j], a[(size_t)p * n + j]);
            cld d = a[(size_t)k * n + k];
            if (d == cld(0)) d = a[(size_t)k * n + k] = cld(std::numeric_limits<ld>::epsilon());
            for (int i = k + 1; i < n; ++i) {
                const cld l = a[(size_t)i * n + k] / d;
                a[(size_t)i * n + k] = l;
                for (int j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
            }
        }
    }
    void solve(std::vector<cld> &b) const {
        for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < i; ++k) b[i] -= a[(size_t)i * n + k] * b[k];
        for (int i = n - 1; i >= 0; --i) {
            for (int k = i + 1; k < n; ++k) b[i] -= a[(size_t)i * n + k] * b[k];
            b[i] /= a[(size_t)i * n + i];
        }
    }
};

}  // namespace

// D = Q B Q^-1 of one M x M block (see the file comment); `what` names the block
// in errors
static HelmholtzBasis eigen_basis(const Mat &D, const char *what) {
    const int M = D.n;
    Mat Bal = D;
    const std::vector<ld> sc = balance(Bal);
    Mat H = Bal, Z(M);
    hessenberg(H, Z);
    std::vector<cld> ev = francis(H);
    const Mat &H0 = H;
    if ((int)ev.size() != M) throw Error(2, std::string("eigendecomposition of ") + what + ": eigenvalue count");
    // pairs first (positive imaginary part kept), then the real eigenvalues
    ld norm = 0;
    for (int i = 0; i < M; ++i) {
        ld r = 0;
        for (int j = 0; j < M; ++j) r += std::fabs(D(i, j));
        norm = std::max(norm, r);
    }
    std::vector<cld> pairs, reals;
    for (const cld &e : ev) {
        if (std::fabs(e.imag()) > 1e-12L * norm) {
            if (e.imag() > 0) pairs.push_back(e);
        } else {
            reals.push_back(cld(e.real(), 0));
        }
    }
    if ((int)(2 * pairs.size() + reals.size()) != M)
        throw Error(2, std::string("eigendecomposition of ") + what + ": unpaired complex eigenvalue");
    std::sort(reals.begin(), reals.end(), [](const cld &a, const cld &b) { return a.real() > b.real(); });
    HelmholtzBasis hb;
    hb.M = M;
    hb.npair = (int)pairs.size();
    std::vector<ld> Q((size_t)M * M);  // column-major
    std::vector<ld> lre(M), lim(M);
    int col = 0;
    // eigenvector by inverse iteration on the Hessenberg form, back to D's basis
    // (x = S Z y), normalised to unit largest component (real where the
    // eigenvalue is real)
    auto vec = [&](cld lam) {
        const ld del = std::max(std::abs(lam), (ld)1) * 64 * std::numeric_limits<ld>::epsilon();
        std::vector<cld> y(M);
        for (int i = 0; i < M; ++i) y[i] = cld(1.0L + 0.01L * (ld)((i * 7919) % 101), 0.001L * (ld)(i % 13));
        for (int it = 0; it < 3; ++it) {
            hess_solve(H0, lam + cld(del, 0), y);
            ld mx = 0;
            for (const cld &v : y) mx = std::max(mx, std::abs(v));
            for (cld &v : y) v /= mx;
        }
        std::vector<cld> x(M, cld(0));
        for (int i = 0; i < M; ++i) {
            cld t = 0;
            for (int k = 0; k < M; ++k) t += Z(i, k) * y[k];
            x[i] = t * sc[i];
        }
        int im = 0;
        for (int i = 1; i < M; ++i)
            if (std::abs(x[i]) > std::abs(x[im])) im = i;
        const cld ph = x[im];
        for (cld &v : x) v /= ph;
        return x;
    };
    for (const cld &lam : pairs) {
        const std::vector<cld> x = vec(lam);
        for (int i = 0; i < M; ++i) {
            Q[(size_t)col * M + i] = x[i].real();
            Q[(size_t)(col + 1) * M + i] = x[i].imag();
        }
        lre[col] = lre[col + 1] = lam.real();
        lim[col] = lim[col + 1] = lam.imag();
        col += 2;
    }
    for (const cld &lam : reals) {
        const std::vector<cld> x = vec(lam);
        for (int i = 0; i < M; ++i) Q[(size_t)col * M + i] = x[i].real();
        lre[col] = lam.real();
        lim[col] = 0;
        ++col;
    }
    // Q^-1 by LU with partial pivoting (long double)
    Mat Qr(M);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) Qr(i, j) = Q[(size_t)j * M + i];
    CLU qlu(Qr, cld(0));
    std::vector<ld> Qi((size_t)M * M);
    for (int j = 0; j < M; ++j) {
        std::vector<cld> e(M, cld(0));
        e[j] = 1;
        qlu.solve(e);
        for (int i = 0; i < M; ++i) Qi[(size_t)j * M + i] = e[i].real();
    }
    // checks: D Q = Q B and Q Q^-1 = I
    ld qn = 0, res = 0, ide = 0;
    for (int j = 0; j < M; ++j)
        for (int i = 0; i < M; ++i) qn = std::max(qn, std::fabs(Q[(size_t)j * M + i]));
    for (int j = 0; j < M; ++j) {
        for (int i = 0; i < M; ++i) {
            ld dq = 0;
            for (int k = 0; k < M; ++k) dq += D(i, k) * Q[(size_t)j * M + k];
            ld qb;
            if (lim[j] == 0) qb = Q[(size_t)j * M + i] * lre[j];
            else if (j % 2 == 0 && j < 2 * hb.npair)  // column x_r: D x_r = a x_r - b x_i
                qb = lre[j] * Q[(size_t)j * M + i] - lim[j] * Q[(size_t)(j + 1) * M + i];
            else  // column x_i: D x_i = b x_r + a x_i
                qb = lim[j] * Q[(size_t)(j - 1) * M + i] + lre[j] * Q[(size_t)j * M + i];
            res = std::max(res, std::fabs(dq - qb));
            ld qq = 0;
            for (int k = 0; k < M; ++k) qq += Q[(size_t)k * M + i] * Qi[(size_t)j * M + k];
            ide = std::max(ide, std::fabs(qq - (i == j ? 1 : 0)));
        }
    }
    hb.residual = (double)(res / (norm * qn));
    hb.identity = (double)ide;
    if (!(hb.residual < 1e-15) || !(hb.identity < 1e-12))
        throw Error(2, std::string("eigendecomposition of ") + what + " failed its check (residual " +
                           std::to_string(hb.residual) + ", identity " + std::to_string(hb.identity) + ")");
    hb.Q.resize((size_t)M * M);
    hb.Qinv.resize((size_t)M * M);
    for (size_t e = 0; e < (size_t)M * M; ++e) {
        hb.Q[e] = (double)Q[e];
        hb.Qinv[e] = (double)Qi[e];
    }
    hb.lre.assign(lre.begin(), lre.end());
    hb.lim.assign(lim.begin(), lim.end());
    return hb;
}

HelmholtzBasis helmholtz_basis(const std::vector<double> &DY2, int N) {
    const int M = N - 2;
    Mat D(M);
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) D(i, j) = DY2[(size_t)(j + 1) * N + (i + 1)];
    return eigen_basis(D, "DY2");
}

// The interior block D symmetrised, Ds = (D + J D J) / 2, maps even vectors
// (x_j = x_{M-1-j}) to even and odd to odd. On the half coordinates j < hm
// (even) and j < ho (odd; the middle of odd M is 0 there) it acts as
//   DE_ij = Ds_ij + Ds_i,M-1-j  (j != middle),  DE_i,mid = Ds_i,mid,
//   DO_ij = Ds_ij - Ds_i,M-1-j,
// and each block is decomposed on its own.
std::vector<double> helmholtz_mirror_basis(const std::vector<double> &DY2, int N) {
    const int M = N - 2, hm = (M + 1) / 2, ho = M / 2;
    auto d = [&](int i, int j) { return (ld)DY2[(size_t)(j + 1) * N + (i + 1)]; };
    auto ds = [&](int i, int j) { return (d(i, j) + d(M - 1 - i, M - 1 - j)) / 2; };
    Mat DE(hm), DO(ho);
    for (int i = 0; i < hm; ++i)
        for (int j = 0; j < hm; ++j) DE(i, j) = 2 * j == M - 1 ? ds(i, j) : ds(i, j) + ds(i, M - 1 - j);
    for (int i = 0; i < ho; ++i)
        for (int j = 0; j < ho; ++j) DO(i, j) = ds(i, j) - ds(i, M - 1 - j);
    const HelmholtzBasis be = eigen_basis(DE, "DY2 (even block)"), bo = eigen_basis(DO, "DY2 (odd block)");
    const size_t hh = (size_t)hm * hm;
    std::vector<double> out(4 * hh + 4 * (size_t)hm, 0.0);
    // the inverses take the fold's factor 1/2: e_j = r_j + r_{M-1-j}, o_j = r_j - r_{M-1-j}
    for (int j = 0; j < hm; ++j)
        for (int i = 0; i < hm; ++i) {
            out[(size_t)j * hm + i] = 0.5 * be.Qinv[(size_t)j * hm + i];
            out[hh + (size_t)j * hm + i] = be.Q[(size_t)j * hm + i];
        }
    for (int j = 0; j < ho; ++j)
        for (int i = 0; i < ho; ++i) {
            out[2 * hh + (size_t)j * hm + i] = 0.5 * bo.Qinv[(size_t)j * ho + i];
            out[3 * hh + (size_t)j * hm + i] = bo.Q[(size_t)j * ho + i];
        }
    double *lr = out.data() + 4 * hh;
    for (int i = 0; i < hm; ++i) {
        lr[i] = be.lre[i];
        lr[hm + i] = be.lim[i];
    }
    for (int i = 0; i < ho; ++i) {
        lr[2 * hm + i] = bo.lre[i];
        lr[3 * hm + i] = bo.lim[i];
    }
    return out;
}

}  // namespace owt
