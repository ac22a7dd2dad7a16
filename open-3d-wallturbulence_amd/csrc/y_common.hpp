// Device helpers shared by the wall-normal kernels (y_kernels.hip, lu_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace owt {

// KIO stiffly stable third-order scheme (src/solver.cpp:548-556)
constexpr double G0 = 11.0 / 6.0, AL0 = 3.0, AL1 = -1.5, AL2 = 1.0 / 3.0, BE0 = 3.0, BE1 = -3.0, BE2 = 1.0;

// exactly rounded elementwise arithmetic, as the reference's separate vector
// operations round each product and sum. HIP's __dmul_rn / __dadd_rn are plain
// operators that hipcc contracts into fused multiply-adds; contraction is
// switched off inside these helpers, so a product and the sum it feeds stay two
// roundings (v_mul_f64 + v_add_f64).
__device__ __forceinline__ double mul(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ double sub(double a, double b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ double2 c2(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cscale(double2 v, double s) { return c2(mul(v.x, s), mul(v.y, s)); }
__device__ __forceinline__ double2 caddc(double2 a, double2 b) { return c2(add(a.x, b.x), add(a.y, b.y)); }
__device__ __forceinline__ double2 csubc(double2 a, double2 b) { return c2(sub(a.x, b.x), sub(a.y, b.y)); }
__device__ __forceinline__ double2 cdiv(double2 v, double s) { return c2(__ddiv_rn(v.x, s), __ddiv_rn(v.y, s)); }
// (0 + i s) * v
__device__ __forceinline__ double2 imul(double s, double2 v) { return c2(-mul(s, v.y), mul(s, v.x)); }

__device__ __forceinline__ double2 *hist(const Chan &c, int f, int slot) { return c.hist + (size_t)(f * 3 + slot) * c.W; }

__device__ __forceinline__ double k2_of(const Chan &c, int k, int i) {
    const double a = c.kx[i], b = c.kz[k];
    return add(mul(a, a), mul(b, b));
}

// The stored wall-normal term of a history level (Chan::d2v) from DY2 v and v
// of one row: DY2 v in FP64, DY2 v - k^2 v in mixed precision (the combination
// the v right-hand side takes, so that its levels need no reads of v)
__device__ __forceinline__ double2 d2v_store(const Chan &c, double2 d2v, double2 v, double k2) {
    return c.precision == 1 ? csubc(d2v, cscale(v, k2)) : d2v;
}

template <typename T>
__device__ __forceinline__ double2 ld2(const T *p) {
    T v = *p;
    return c2((double)v.x, (double)v.y);
}

// value of v in lane `lane` (wave-uniform lane index)
__device__ __forceinline__ double readlane_d(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Four real right-hand sides held by one wave, row l + 64q in lane l, register q.
// Gather them in pivot order (dlaswp) from the wave's LDS copy S.
template <int NQ>
__device__ __forceinline__ void lu_gather(const double (*S)[64 * NQ], const int *Po, int N, int lane, double (&t)[4][NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int row = lane + 64 * q;
        int pr = row < N ? Po[row] : 0;
        if ((unsigned)pr >= (unsigned)N) pr = 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) t[r][q] = row < N ? S[r][pr] : 0.0;
    }
}

// Forward substitution with the unit lower factor, then back substitution with
// the upper factor (dtrsm as called by dgetrs, src/linalg.hpp:660-680), on a
// packed factor. EXACT rounds every product and difference separately and
// divides by the diagonal, as the reference's BLAS does (b_i - b_k l_ik, then
// b_k / u_kk); otherwise fused multiply-adds and one reciprocal per column.
//  - Steps run block by block: step k belongs to row block qk = k / 64, a
//    compile-time index, so the pivot value is read from a fixed register of
//    lane k % 64 and only the row blocks that meet the triangle are touched.
//  - Columns come in chunks of CK, and the loads of the next chunk are issued
//    before the current one is used (two buffers), so a wave always has factor
//    loads in flight. Every lane of a touched block loads: a lane whose row lies
//    outside the triangle reads an element of the column's own run instead (one
//    cache line for all such lanes) and its value is masked, so no load sits
//    behind a branch and only the triangle is fetched.
//  - Without EXACT the back substitution scales by the reciprocal of the
//    diagonal, as optimized trsm kernels do (one division per column).
template <int NQ, int CK>
__device__ __forceinline__ void lu_load_lower(const double *__restrict__ Fo, int N, int lane, int qk, int kb, double (&cb)[CK][NQ]) {
#pragma unroll
    for (int c = 0; c < CK; ++c) {
        const int kc = kb + c < N ? kb + c : N - 1;
        // rows kc+1..N-1 of column kc (for kc = N-1 an empty run: the masked
        // lanes read the first upper element, still inside the factor)
        const double *col = Fo + lu_lower_at(N, kc);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q < qk) continue;
            const int row = lane + 64 * q;
            cb[c][q] = col[(row > kc && row < N) ? row - kc - 1 : 0];
        }
    }
}

template <int NR, int NQ, int CK, bool EXACT>
__device__ __forceinline__ void lu_apply_lower(int N, int lane, int qk, int kb, double (&cb)[CK][NQ], double (&t)[NR][NQ]) {
#pragma unroll
    for (int c = 0; c < CK; ++c) {
        const int k = kb + c;
        if (k >= N) break;
        double yk[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) yk[r] = readlane_d(t[r][qk], k & 63);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            if (q < qk) continue;
            const int row = lane + 64 * q;
            const double m = (row > k && row < N) ? cb[c][q] : 0.0;
#pragma unroll
            for (int r = 0; r < NR; ++r) t[r][q] = EXACT ? sub(t[r][q], mul(yk[r], m)) : fma(-m, yk[r], t[r][q]);
        }
    }
}

template <int NQ, int NR = 4, int CK = 4, bool EXACT = false>
__device__ __forceinline__ void lu_substitute(const double *__restrict__ Fo, int N, int lane, double (&t)[NR][NQ]) {
    // forward: chunks of CK columns inside each row block, next chunk in flight
#pragma unroll
    for (int qk = 0; qk < NQ; ++qk) {
        const int kend = 64 * qk + 64 < N ? 64 * qk + 64 : N;
        int kb = 64 * qk;
        if (kb >= kend) break;
        double ca[CK][NQ], cc[CK][NQ];
        lu_load_lower<NQ, CK>(Fo, N, lane, qk, kb, ca);
        while (kb < kend) {
            const bool more = kb + CK < kend;
            if (more) lu_load_lower<NQ, CK>(Fo, N, lane, qk, kb + CK, cc);
            lu_apply_lower<NR, NQ, CK, EXACT>(N, lane, qk, kb, ca, t);
            kb += CK;
            if (!more) break;
            const bool more2 = kb + CK < kend;
            if (more2) lu_load_lower<NQ, CK>(Fo, N, lane, qk, kb + CK, ca);
            lu_apply_lower<NR, NQ, CK, EXACT>(N, lane, qk, kb, cc, t);
            kb += CK;
            if (!more2) break;
        }
    }
    // backward: the same, columns from the top of each row block down
#pragma unroll
    for (int qk = NQ - 1; qk >= 0; --qk) {
        const int lo = 64 * qk;
        int kb = (lo + 63 < N - 1 ? lo + 63 : N - 1);
        if (kb < lo) continue;
        auto load = [&](int k0, double (&cb)[CK][NQ]) {
#pragma unroll
            for (int c = 0; c < CK; ++c) {
                const int kc = k0 - c >= lo ? k0 - c : lo;
                const double *col = Fo + lu_upper_at(N, kc);  // rows 0..kc
#pragma unroll
                for (int q = 0; q <= qk; ++q) {
                    const int row = lane + 64 * q;
                    cb[c][q] = col[row <= kc ? row : kc];
                }
            }
        };
        auto apply = [&](int k0, double (&cb)[CK][NQ]) {
#pragma unroll
            for (int c = 0; c < CK; ++c) {
                const int k = k0 - c;
                if (k < lo) break;
                const int lk = k & 63;
                const double dk = readlane_d(cb[c][qk], lk);
                const double rinv = EXACT ? 0.0 : 1.0 / dk;
                double xk[NR];
#pragma unroll
                for (int r = 0; r < NR; ++r) xk[r] = EXACT ? __ddiv_rn(readlane_d(t[r][qk], lk), dk) : readlane_d(t[r][qk], lk) * rinv;
                if (lane == lk) {
#pragma unroll
                    for (int r = 0; r < NR; ++r) t[r][qk] = xk[r];
                }
#pragma unroll
                for (int q = 0; q <= qk; ++q) {
                    const int row = lane + 64 * q;
                    const double m = row < k ? cb[c][q] : 0.0;
#pragma unroll
                    for (int r = 0; r < NR; ++r) t[r][q] = EXACT ? sub(t[r][q], mul(xk[r], m)) : fma(-m, xk[r], t[r][q]);
                }
            }
        };
        double ca[CK][NQ], cc[CK][NQ];
        load(kb, ca);
        while (kb >= lo) {
            const bool more = kb - CK >= lo;
            if (more) load(kb - CK, cc);
            apply(kb, ca);
            kb -= CK;
            if (!more) break;
            const bool more2 = kb - CK >= lo;
            if (more2) load(kb - CK, ca);
            apply(kb, cc);
            kb -= CK;
            if (!more2) break;
        }
    }
}

// register blocks of 64 rows per lane set, by NY
#define OWT_NQ_DISPATCH(N, BODY)                                      \
    do {                                                              \
        const int nq_ = ((N) + 63) / 64;                              \
        if (nq_ <= 1) { constexpr int NQ = 1; BODY; }                 \
        else if (nq_ == 2) { constexpr int NQ = 2; BODY; }            \
        else if (nq_ == 3) { constexpr int NQ = 3; BODY; }            \
        else if (nq_ == 4) { constexpr int NQ = 4; BODY; }            \
        else { constexpr int NQ = 5; BODY; }                          \
    } while (0)

}  // namespace owt
