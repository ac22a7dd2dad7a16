// Device helpers shared by the wall-normal kernels (y_kernels.hip, lu_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "device.hpp"

namespace owt {

// KIO stiffly stable third-order scheme (src/solver.cpp:548-556)
constexpr double G0 = 11.0 / 6.0, AL0 = 3.0, AL1 = -1.5, AL2 = 1.0 / 3.0, BE0 = 3.0, BE1 = -3.0, BE2 = 1.0;

// exactly rounded elementwise arithmetic, as the reference's separate vector
// operations round each product and sum
__device__ __forceinline__ double mul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double add(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double sub(double a, double b) { return __dsub_rn(a, b); }
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
// the upper factor (dtrsm as called by dgetrs, src/linalg.hpp:660-680). The
// factor is N x N column-major. Columns are loaded CK at a time (coalesced,
// all CK loads in flight before the first is used). Every lane loads: a lane
// whose row lies outside the triangle being applied reads the column's
// diagonal element instead (one cache line shared by all such lanes), and its
// value is masked afterwards, so that no load sits behind a branch and only
// the triangle is fetched. The value of each step is broadcast from the lane
// that holds it.
template <int NQ, int CK = 8>
__device__ __forceinline__ void lu_substitute(const double *__restrict__ Fo, int N, int lane, double (&t)[4][NQ]) {
    for (int k0 = 0; k0 < N; k0 += CK) {
        double cb[CK][NQ];
#pragma unroll
        for (int c = 0; c < CK; ++c) {
            const int k = k0 + c;
            const double *col = Fo + (size_t)(k < N ? k : N - 1) * N;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q;
                const int kc = k < N ? k : N - 1;
                cb[c][q] = col[(row > kc && row < N) ? row : kc];
            }
        }
#pragma unroll
        for (int c = 0; c < CK; ++c)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q, k = k0 + c;
                if (!(k < N && row > k && row < N)) cb[c][q] = 0.0;
            }
#pragma unroll
        for (int c = 0; c < CK; ++c) {
            const int k = k0 + c;
            if (k >= N) break;
            const int qk = k >> 6, lk = k & 63;
            double yk[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double v = t[r][0];
#pragma unroll
                for (int q = 1; q < NQ; ++q)
                    if (q == qk) v = t[r][q];
                yk[r] = readlane_d(v, lk);
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q;
                if (row > k && row < N) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) t[r][q] = fma(-cb[c][q], yk[r], t[r][q]);
                }
            }
        }
    }
    for (int k0 = N - 1; k0 >= 0; k0 -= CK) {
        double cb[CK][NQ];
#pragma unroll
        for (int c = 0; c < CK; ++c) {
            const int k = k0 - c;
            const double *col = Fo + (size_t)(k >= 0 ? k : 0) * N;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q;
                const int kc = k >= 0 ? k : 0;
                cb[c][q] = col[row <= kc ? row : kc];
            }
        }
#pragma unroll
        for (int c = 0; c < CK; ++c)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q, k = k0 - c;
                if (!(k >= 0 && row <= k)) cb[c][q] = 0.0;
            }
#pragma unroll
        for (int c = 0; c < CK; ++c) {
            const int k = k0 - c;
            if (k < 0) break;
            const int qk = k >> 6, lk = k & 63;
            // the diagonal U(k,k) sits in lane lk, register qk
            double dv = cb[c][0];
#pragma unroll
            for (int q = 1; q < NQ; ++q)
                if (q == qk) dv = cb[c][q];
            const double d = readlane_d(dv, lk);
            double xk[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                double v = t[r][0];
#pragma unroll
                for (int q = 1; q < NQ; ++q)
                    if (q == qk) v = t[r][q];
                xk[r] = __ddiv_rn(readlane_d(v, lk), d);
            }
            if (lane == lk) {
#pragma unroll
                for (int q = 0; q < NQ; ++q)
                    if (q == qk) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) t[r][q] = xk[r];
                    }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int row = lane + 64 * q;
                if (row < k) {
#pragma unroll
                    for (int r = 0; r < 4; ++r) t[r][q] = fma(-cb[c][q], xk[r], t[r][q]);
                }
            }
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
        else { constexpr int NQ = 4; BODY; }                          \
    } while (0)

}  // namespace owt
