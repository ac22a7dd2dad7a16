// Wall-normal (y-pencil) phases of the time step.
//
//   y_rhs       right-hand sides of the v and omega_y systems from three history
//               levels (src/solver.cpp:568-599)
//   lu_*        the implicit solves of Av and Ag (src/solver.cpp:576-603) as a
//               batched dense LU with partial pivoting and LAPACK's pivot rule:
//                 plain  : assemble + factor + solve every pencil every step
//                          (the reference method)
//                 store  : assemble + factor once per distinct k^2 (the operators
//                          depend on k^2 only; +kz and -kz share a factor)
//                 apply  : persistent forward/back substitution with stored
//                          factors, four real right-hand sides per factor
//   horizontal  u, w, omega_x, omega_z from v, omega_y (src/solver.cpp:516-539)
//   mean_*      mean flow, friction force, mean vorticity (src/solver.cpp:306-327,
//               428-514) on the single k = 0 pencil
//   nonlinear   Hv, Hg from the Lamb vector (src/solver.cpp:240-260)
//
// Pencil p = k*NXH + i holds NY contiguous complex values (y fastest). Mat-vecs
// with the dense DY_n run one row per thread with the pencil vectors staged in
// LDS and the operator columns read coalesced from L2.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "device.hpp"

namespace owt {

namespace {

constexpr double G0 = 11.0 / 6.0, AL0 = 3.0, AL1 = -1.5, AL2 = 1.0 / 3.0, BE0 = 3.0, BE1 = -3.0, BE2 = 1.0;

__device__ __forceinline__ int slot_of(int ring, int level) { return (ring - level + 3) % 3; }
__device__ __forceinline__ int new_slot(int ring) { return (ring + 1) % 3; }
__device__ __forceinline__ double2 *hist(const Chan &c, int f, int slot) { return c.hist + (size_t)(f * 3 + slot) * c.W; }

// exactly rounded elementwise arithmetic (no contraction), as the reference's
// separate vector operations round each product and sum
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

__device__ __forceinline__ double k2_of(const Chan &c, int k, int i) {
    const double a = c.kx[i], b = c.kz[k];
    return add(mul(a, a), mul(b, b));
}

template <typename T>
__device__ __forceinline__ double2 ld2(const T *p) {
    T v = *p;
    return c2((double)v.x, (double)v.y);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

}  // namespace

// ---------------------------------------------------------------------------
// right-hand sides (src/solver.cpp:568-574, 593-599)
// ---------------------------------------------------------------------------
template <int TP>
__global__ __launch_bounds__(256) void k_y_rhs(Chan c) {
    __shared__ double2 sa[TP][kMaxNY];
    __shared__ double2 sv[TP][kMaxNY];
    const int N = c.NY, r = c.ctrl[C_RING];
    const double dt = c.dt;
    const double2 *v0 = hist(c, H_V, slot_of(r, 0)), *v1 = hist(c, H_V, slot_of(r, 1)), *v2 = hist(c, H_V, slot_of(r, 2));
    const double2 *g0 = hist(c, H_G, slot_of(r, 0)), *g1 = hist(c, H_G, slot_of(r, 1)), *g2 = hist(c, H_G, slot_of(r, 2));
    const double2 *h0 = hist(c, H_HV, slot_of(r, 0)), *h1 = hist(c, H_HV, slot_of(r, 1)), *h2 = hist(c, H_HV, slot_of(r, 2));
    const double2 *q0 = hist(c, H_HG, slot_of(r, 0)), *q1 = hist(c, H_HG, slot_of(r, 1)), *q2 = hist(c, H_HG, slot_of(r, 2));
    double2 *srcv = c.src, *srcg = c.src + c.W;
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 a = c2(0, 0), s = c2(0, 0);
        if (p < c.npencil) {
            const int k = p / c.NXH, i = p % c.NXH;
            const double k2 = k2_of(c, k, i);
            const size_t o = (size_t)p * N + j;
            a = caddc(caddc(cscale(v0[o], AL0), cscale(v1[o], AL1)), cscale(v2[o], AL2));
            double2 hv = caddc(caddc(cscale(h0[o], BE0), cscale(h1[o], BE1)), cscale(h2[o], BE2));
            s = csubc(hv, cscale(a, __ddiv_rn(k2, dt)));
            double2 hg = caddc(caddc(cscale(q0[o], BE0), cscale(q1[o], BE1)), cscale(q2[o], BE2));
            double2 va = caddc(caddc(cscale(g0[o], AL0), cscale(g1[o], AL1)), cscale(g2[o], AL2));
            double2 sg = caddc(hg, cdiv(va, dt));
            if (j == 0 || j == N - 1) sg = c2(0, 0);
            srcg[o] = sg;
        }
        sa[pp][j] = a;
        sv[pp][j] = s;
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row >= N) return;
    double2 acc[TP];
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(0, 0);
    for (int j = 0; j < N; ++j) {
        const double d = c.DY2[(size_t)j * N + row];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            acc[pp].x = fma(d, sa[pp][j].x, acc[pp].x);
            acc[pp].y = fma(d, sa[pp][j].y, acc[pp].y);
        }
    }
    const bool bc = row <= 1 || row >= N - 2;
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) {
        const int p = p0 + pp;
        if (p >= c.npencil) break;
        double2 v = caddc(sv[pp][row], cdiv(acc[pp], dt));
        if (bc) v = c2(0, 0);
        srcv[(size_t)p * N + row] = v;
    }
}

// ---------------------------------------------------------------------------
// Dense LU with partial pivoting on one matrix per block.
// 32 x TC threads; thread (tr, tc) keeps rows tr + 32a and columns tc + TC*b
// in registers (RPT x CPT). Rows are never moved: the pivot of step k is
// recorded and its row retires; the rank of a row is its pivot step, which is
// its row index in LAPACK's permuted factors. Ties in |a| go to the row with
// the smaller position in LAPACK's interchanged order.
// Right-hand sides ride along as extra columns (forward substitution for free),
// followed by a column-oriented back substitution.
// ---------------------------------------------------------------------------
enum LuMode { LU_PLAIN = 0, LU_MEAN = 1, LU_STORE = 2 };

constexpr int LTR = 32;

template <int RPT, int CPT, int LTC>
__global__ __launch_bounds__(LTR *LTC) void k_lu(Chan c, int mode) {
    constexpr int LNT = LTR * LTC;
    __shared__ double colval[2][kMaxNY];
    __shared__ double colmag[2][kMaxNY];
    __shared__ double rowbuf[2][kMaxNY + 8];
    __shared__ int pos[kMaxNY], rowAt[kMaxNY], perm[kMaxNY];
    __shared__ int pivrow[2];
    __shared__ double pivval[2];
    __shared__ double xval[2][4];

    const int N = c.NY;
    const int t = threadIdx.x, tr = t % LTR, tc = t / LTR;
    int op, nrhs = 0;
    double k2 = 0.0;
    const double2 *rhs[2] = {nullptr, nullptr};
    double2 *outv[2] = {nullptr, nullptr};
    int sys = 0;
    if (mode == LU_PLAIN) {
        const int p = blockIdx.x;
        op = blockIdx.y;
        const int k = p / c.NXH, i = p % c.NXH;
        k2 = k2_of(c, k, i);
        nrhs = 2;
        rhs[0] = c.src + (size_t)op * c.W + (size_t)p * N;
        outv[0] = hist(c, op == 0 ? H_V : H_G, new_slot(c.ctrl[C_RING])) + (size_t)p * N;
    } else if (mode == LU_MEAN) {
        op = 1;  // -nu DY2 + gamma0/dt I with Dirichlet rows, i.e. Ag at k^2 = 0
        nrhs = 4;
        rhs[0] = c.mrhs;
        rhs[1] = c.mrhs + N;
        outv[0] = c.msol;
        outv[1] = c.msol + N;
    } else {
        sys = blockIdx.x;
        op = blockIdx.y;
        const int mz = sys / c.NXH, i = sys % c.NXH;
        k2 = k2_of(c, mz, i);
    }
    const int ncols = N + nrhs;
    const double nu = c.nu, dt = c.dt;
    const double k4 = mul(k2, k2);
    const double s1 = add(__ddiv_rn(G0, dt), mul(mul(2.0, nu), k2));
    const double s2 = add(mul(nu, k4), __ddiv_rn(mul(G0, k2), dt));
    const double sg = add(mul(nu, k2), __ddiv_rn(G0, dt));

    // assemble Av or Ag with boundary rows (src/solver.cpp:577-591), RHS columns appended
    double M[RPT][CPT];
#pragma unroll
    for (int a = 0; a < RPT; ++a) {
        const int r = tr + LTR * a;
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
            const int col = tc + LTC * b;
            double v = 0.0;
            if (r < N && col < N) {
                const size_t e = (size_t)col * N + r;
                if (r == 0 || r == N - 1) {
                    v = (r == col) ? 1.0 : 0.0;
                } else if (op == 0) {
                    if (r == 1) v = c.DY1[(size_t)col * N];
                    else if (r == N - 2) v = c.DY1[(size_t)col * N + N - 1];
                    else {
                        v = add(mul(-nu, c.DY4[e]), mul(s1, c.DY2[e]));
                        if (r == col) v = sub(v, s2);
                    }
                } else {
                    v = mul(-nu, c.DY2[e]);
                    if (r == col) v = add(v, sg);
                }
            } else if (r < N && col < ncols) {
                const int q = col - N;
                const double2 z = rhs[q >> 1][r];
                v = (q & 1) ? z.y : z.x;
            }
            M[a][b] = v;
        }
    }
    for (int r = t; r < N; r += LNT) {
        pos[r] = r;
        rowAt[r] = r;
    }
    unsigned active = 0;
#pragma unroll
    for (int a = 0; a < RPT; ++a)
        if (tr + LTR * a < N) active |= 1u << a;
    int rank[RPT];
#pragma unroll
    for (int a = 0; a < RPT; ++a) rank[a] = INT_MAX;
    __syncthreads();

    bool singular = false;
    for (int k = 0; k < N; ++k) {
        const int bsel = k & 1;
        const int kb = k / LTC;
        const bool owns_k = (tc == (k % LTC));
        // 1. publish column k of the active rows
        if (owns_k) {
#pragma unroll
            for (int a = 0; a < RPT; ++a) {
                const int r = tr + LTR * a;
                if (r < N) {
                    double v = M[a][0];
#pragma unroll
                    for (int b = 1; b < CPT; ++b)
                        if (b == kb) v = M[a][b];
                    colval[bsel][r] = v;
                    colmag[bsel][r] = ((active >> a) & 1u) ? fabs(v) : -1.0;
                }
            }
        }
        __syncthreads();
        // 2. pivot: largest |a|, first in LAPACK's current row order
        if (t < 64) {
            double best = -2.0;
            int bpos = INT_MAX, brow = -1;
            for (int r = t; r < N; r += 64) {
                const double m = colmag[bsel][r];
                const int ps = pos[r];
                if (m > best || (m == best && ps < bpos)) {
                    best = m;
                    bpos = ps;
                    brow = r;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int op2 = __shfl_xor(bpos, off), orow = __shfl_xor(brow, off);
                if (ob > best || (ob == best && op2 < bpos)) {
                    best = ob;
                    bpos = op2;
                    brow = orow;
                }
            }
            if (t == 0) {
                const int p = brow;
                const int q = pos[p], rk = rowAt[k];
                rowAt[q] = rk;
                pos[rk] = q;
                rowAt[k] = p;
                pos[p] = k;
                perm[k] = p;
                pivrow[bsel] = p;
                pivval[bsel] = colval[bsel][p];
            }
        }
        __syncthreads();
        const int p = pivrow[bsel];
        const double pv = pivval[bsel];
        if (pv == 0.0) singular = true;
        // 3. publish the pivot row, retire it
        const int pa = p / LTR;
        if (tr == (p % LTR)) {
#pragma unroll
            for (int b = 0; b < CPT; ++b) {
                const int col = tc + LTC * b;
                if (col > k && col < ncols) {
                    double v = M[0][b];
#pragma unroll
                    for (int a = 1; a < RPT; ++a)
                        if (a == pa) v = M[a][b];
                    rowbuf[bsel][col] = v;
                }
            }
            active &= ~(1u << pa);
#pragma unroll
            for (int a = 0; a < RPT; ++a)
                if (a == pa) rank[a] = k;
        }
        __syncthreads();
        // 4. multipliers and rank-1 update of the active rows
        const bool recip = fabs(pv) >= DBL_MIN;
        const double rinv = 1.0 / pv;
        double l[RPT];
#pragma unroll
        for (int a = 0; a < RPT; ++a) {
            const int r = tr + LTR * a;
            l[a] = 0.0;
            if ((active >> a) & 1u) l[a] = recip ? colval[bsel][r] * rinv : colval[bsel][r] / pv;
        }
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
            const int col = tc + LTC * b;
            if (col > k && col < ncols) {
                const double u = rowbuf[bsel][col];
#pragma unroll
                for (int a = 0; a < RPT; ++a)
                    if ((active >> a) & 1u) M[a][b] = fma(-l[a], u, M[a][b]);
            } else if (col == k) {
#pragma unroll
                for (int a = 0; a < RPT; ++a)
                    if ((active >> a) & 1u) M[a][b] = l[a];
            }
        }
    }
    if (singular) {
        if (t == 0) atomicOr(c.ctrl + C_ERR, E_SINGULAR);
        return;
    }
    __syncthreads();

    if (mode == LU_STORE) {
        double *LU = (op == 0 ? c.LUv : c.LUg) + (size_t)sys * N * N;
        int *pm = (op == 0 ? c.permv : c.permg) + (size_t)sys * N;
#pragma unroll
        for (int a = 0; a < RPT; ++a) {
            const int r = tr + LTR * a;
            if (r >= N) continue;
#pragma unroll
            for (int b = 0; b < CPT; ++b) {
                const int col = tc + LTC * b;
                if (col < N) LU[(size_t)col * N + rank[a]] = M[a][b];
            }
        }
        for (int k = t; k < N; k += LNT) pm[k] = perm[k];
        return;
    }

    // back substitution with U (dtrsm upper, non-unit; src/linalg.hpp:670-680)
    for (int k = N - 1; k >= 0; --k) {
        const int bsel = k & 1;
        const int kb = k / LTC;
        if (tc == (k % LTC)) {
#pragma unroll
            for (int a = 0; a < RPT; ++a) {
                const int r = tr + LTR * a;
                if (r < N) {
                    double v = M[a][0];
#pragma unroll
                    for (int b = 1; b < CPT; ++b)
                        if (b == kb) v = M[a][b];
                    colval[bsel][r] = v;
                }
            }
        }
        __syncthreads();
        const int p = perm[k];
        if (tr == (p % LTR)) {
            const int pa = p / LTR;
#pragma unroll
            for (int b = 0; b < CPT; ++b) {
                const int col = tc + LTC * b;
                if (col >= N && col < ncols) {
                    double y = M[0][b];
#pragma unroll
                    for (int a = 1; a < RPT; ++a)
                        if (a == pa) y = M[a][b];
                    const double x = __ddiv_rn(y, colval[bsel][p]);
                    const int q = col - N;
                    xval[bsel][q] = x;
                    double *o = reinterpret_cast<double *>(outv[q >> 1] + k);
                    o[q & 1] = x;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < CPT; ++b) {
            const int col = tc + LTC * b;
            if (col >= N && col < ncols) {
                const double x = xval[bsel][col - N];
#pragma unroll
                for (int a = 0; a < RPT; ++a)
                    if (rank[a] < k) M[a][b] = fma(-colval[bsel][tr + LTR * a], x, M[a][b]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Solve with stored factors: one wavefront per (distinct k^2, operator), four
// real right-hand sides (Re/Im of the +kz and -kz pencils). Lane l holds rows
// l + 64q. Persistent grid-stride over all (system, operator) pairs.
// ---------------------------------------------------------------------------
template <int NQ>
__device__ __forceinline__ void lu_apply_wave(const double *__restrict__ M, const int *__restrict__ pm, int N, const double2 *in0,
                                              const double2 *in1, double2 *out0, double2 *out1, int lane) {
    double t[4][NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int row = lane + 64 * q;
        double2 a = c2(0, 0), b = c2(0, 0);
        if (row < N) {
            const int pr = pm[row];
            a = in0[pr];
            if (in1) b = in1[pr];
        }
        t[0][q] = a.x;
        t[1][q] = a.y;
        t[2][q] = b.x;
        t[3][q] = b.y;
    }
    // forward substitution with unit-lower L (dtrsm lower, unit)
    for (int k = 0; k < N; ++k) {
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
        const double *col = M + (size_t)k * N;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int row = lane + 64 * q;
            if (row > k && row < N) {
                const double m = col[row];
#pragma unroll
                for (int r = 0; r < 4; ++r) t[r][q] = fma(-m, yk[r], t[r][q]);
            }
        }
    }
    // back substitution with U
    for (int k = N - 1; k >= 0; --k) {
        const int qk = k >> 6, lk = k & 63;
        const double *col = M + (size_t)k * N;
        const double d = col[k];
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
                const double m = col[row];
#pragma unroll
                for (int r = 0; r < 4; ++r) t[r][q] = fma(-m, xk[r], t[r][q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int row = lane + 64 * q;
        if (row < N) {
            out0[row] = c2(t[0][q], t[1][q]);
            if (out1) out1[row] = c2(t[2][q], t[3][q]);
        }
    }
}

template <int NQ>
__global__ __launch_bounds__(256) void k_lu_apply(Chan c) {
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int nw = gridDim.x * (blockDim.x >> 6);
    const int N = c.NY, ns = new_slot(c.ctrl[C_RING]);
    const int total = 2 * c.nsys;
    for (int w = wid; w < total; w += nw) {
        const int op = w & 1, sys = w >> 1;
        const int mz = sys / c.NXH, i = sys % c.NXH;
        const int kp = mz;                                            // +kz pencil (or the z-Nyquist)
        const int km = (mz > 0 && mz < c.NZH) ? c.NZ - mz : -1;       // -kz pencil
        const double2 *src = c.src + (size_t)op * c.W;
        double2 *dst = hist(c, op == 0 ? H_V : H_G, ns);
        const size_t op_ = (size_t)(kp * c.NXH + i) * N;
        const size_t om_ = km >= 0 ? (size_t)(km * c.NXH + i) * N : 0;
        const double *M = (op == 0 ? c.LUv : c.LUg) + (size_t)sys * N * N;
        const int *pm = (op == 0 ? c.permv : c.permg) + (size_t)sys * N;
        lu_apply_wave<NQ>(M, pm, N, src + op_, km >= 0 ? src + om_ : nullptr, dst + op_, km >= 0 ? dst + om_ : nullptr, lane);
    }
}

template <int NQ>
__global__ __launch_bounds__(64) void k_lu_apply_mean(Chan c) {
    // the mean-flow operator is Ag at k^2 = 0: system 0 of the omega_y factors
    lu_apply_wave<NQ>(c.LUg, c.permg, c.NY, c.mrhs, c.mrhs + c.NY, c.msol, c.msol + c.NY, threadIdx.x);
}

// ---------------------------------------------------------------------------
// horizontal velocity and vorticity (src/solver.cpp:516-539)
// ---------------------------------------------------------------------------
template <int TP>
__global__ __launch_bounds__(256) void k_horizontal(Chan c) {
    __shared__ double2 sv[TP][kMaxNY];
    __shared__ double2 sg[TP][kMaxNY];
    const int N = c.NY, ns = new_slot(c.ctrl[C_RING]);
    const double2 *v = hist(c, H_V, ns), *g = hist(c, H_G, ns);
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 a = c2(0, 0), b = c2(0, 0);
        if (p < c.npencil) {
            a = v[(size_t)p * N + j];
            b = g[(size_t)p * N + j];
        }
        sv[pp][j] = a;
        sg[pp][j] = b;
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row >= N) return;
    double2 d1v[TP], d2v[TP], d1g[TP];
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) d1v[pp] = d2v[pp] = d1g[pp] = c2(0, 0);
    for (int j = 0; j < N; ++j) {
        const double a = c.DY1[(size_t)j * N + row], b = c.DY2[(size_t)j * N + row];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const double2 x = sv[pp][j], y = sg[pp][j];
            d1v[pp].x = fma(a, x.x, d1v[pp].x);
            d1v[pp].y = fma(a, x.y, d1v[pp].y);
            d2v[pp].x = fma(b, x.x, d2v[pp].x);
            d2v[pp].y = fma(b, x.y, d2v[pp].y);
            d1g[pp].x = fma(a, y.x, d1g[pp].x);
            d1g[pp].y = fma(a, y.y, d1g[pp].y);
        }
    }
    double2 *u = c.cur, *w = c.cur + c.W, *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W;
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) {
        const int p = p0 + pp;
        if (p >= c.npencil) break;
        const int k = p / c.NXH, i = p % c.NXH;
        const double kx = c.kx[i], kz = c.kz[k];
        const double k2 = k2_of(c, k, i);
        if (k2 < 1.e-8) continue;  // the mean pencil is set by the mean-flow kernels
        const double ix = __ddiv_rn(kx, k2), iz = __ddiv_rn(kz, k2), ir = __ddiv_rn(1.0, k2);
        const double2 vv = sv[pp][row], gg = sg[pp][row];
        const size_t o = (size_t)p * N + row;
        u[o] = csubc(imul(ix, d1v[pp]), imul(iz, gg));
        w[o] = caddc(imul(ix, gg), imul(iz, d1v[pp]));
        const double2 X = csubc(caddc(cscale(d1g[pp], kx), cscale(d2v[pp], kz)), cscale(vv, mul(kz, k2)));
        const double2 Y = caddc(csubc(cscale(d1g[pp], kz), cscale(d2v[pp], kx)), cscale(vv, mul(kx, k2)));
        ox[o] = imul(ir, X);
        oz[o] = imul(ir, Y);
    }
}

// ---------------------------------------------------------------------------
// mean flow (src/solver.cpp:428-514) and friction force (src/solver.cpp:306-327)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mean_rhs(Chan c) {
    const int N = c.NY, r = c.ctrl[C_RING];
    const double dt = c.dt;
    int sl[3];
    for (int l = 0; l < 3; ++l) sl[l] = slot_of(r, l);
    const double BE[3] = {BE0, BE1, BE2}, AL[3] = {AL0, AL1, AL2};
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double2 su, sw;
        for (int comp = 0; comp < 2; ++comp) {
            const double2 *lm = c.means + (size_t)((comp == 0 ? M_LMBX : M_LMBZ) * 3) * N;
            const double2 *vm = c.means + (size_t)((comp == 0 ? M_VELX : M_VELZ) * 3) * N;
            const double *fr = c.fric + comp * 3;
            double2 t[3], va[3];
            for (int l = 0; l < 3; ++l) {
                const double2 m = lm[(size_t)sl[l] * N + j];
                t[l] = cscale(c2(add(m.x, fr[sl[l]]), m.y), BE[l]);
                va[l] = cscale(vm[(size_t)sl[l] * N + j], AL[l]);
            }
            double2 s = caddc(caddc(caddc(t[0], t[1]), t[2]), cdiv(caddc(caddc(va[0], va[1]), va[2]), dt));
            if (j == 0 || j == N - 1) s = c2(0, 0);
            if (comp == 0) su = s;
            else sw = s;
        }
        c.mrhs[j] = su;
        c.mrhs[N + j] = sw;
    }
}

// dudy rows 0 and N-1, the friction force, the new mean history and the mean
// vorticity; `init` evaluates the friction of the initial field only
__device__ void mean_friction(const Chan &c, const double2 *U, const double2 *Wm, bool init, int ns, int diag_row) {
    __shared__ double2 sU[kMaxNY], sW[kMaxNY];
    __shared__ double wall[4];
    const int N = c.NY;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        sU[j] = U[j];
        sW[j] = Wm[j];
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row < N) {
        double2 du = c2(0, 0), dw = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double a = c.DY1[(size_t)j * N + row];
            du.x = fma(a, sU[j].x, du.x);
            du.y = fma(a, sU[j].y, du.y);
            dw.x = fma(a, sW[j].x, dw.x);
            dw.y = fma(a, sW[j].y, dw.y);
        }
        if (!init) {
            c.cur[row] = sU[row];                                 // u(0,0)
            c.cur[c.W + row] = sW[row];                           // w(0,0)
            c.cur[2 * c.W + row] = dw;                            // omega_x(0,0) = DY1 W
            c.cur[3 * c.W + row] = c2(-du.x, -du.y);              // omega_z(0,0) = -DY1 U
            c.means[(size_t)(M_VELX * 3 + ns) * N + row] = sU[row];
            c.means[(size_t)(M_VELZ * 3 + ns) * N + row] = sW[row];
        }
        if (row == 0) {
            wall[0] = du.x;
            wall[2] = dw.x;
        }
        if (row == N - 1) {
            wall[1] = du.x;
            wall[3] = dw.x;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double nu = c.nu, Ly = c.Ly;
        const double fx = -nu / Ly * (wall[1] - wall[0]);
        const double fz = -nu / Ly * (wall[3] - wall[2]);
        double *d = c.diag + (size_t)(diag_row % c.diag_cap) * 3;
        d[0] = fx;
        d[1] = 2 * wall[0] * nu / Ly;
        d[2] = 2 * wall[1] * nu / Ly;
        if (init) {
            for (int l = 0; l < 3; ++l) {
                c.fric[l] = fx;
                c.fric[3 + l] = fz;
            }
        } else {
            c.fric[ns] = fx;
            c.fric[3 + ns] = fz;
        }
    }
}

__global__ __launch_bounds__(256) void k_mean_post(Chan c) {
    const int ns = new_slot(c.ctrl[C_RING]);
    mean_friction(c, c.msol, c.msol + c.NY, false, ns, c.ctrl[C_STEP] + 1);
}

// ---------------------------------------------------------------------------
// nonlinear source terms (src/solver.cpp:240-260); the k = 0 pencil of the
// Lamb vector also enters the mean-flow history (src/solver.cpp:293-303)
// ---------------------------------------------------------------------------
template <int TP, typename LC>
__global__ __launch_bounds__(256) void k_nonlinear(Chan c) {
    __shared__ double2 st[TP][kMaxNY];
    const int N = c.NY, ns = new_slot(c.ctrl[C_RING]);
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    double2 *Hv = hist(c, H_HV, ns), *Hg = hist(c, H_HG, ns);
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 tt = c2(0, 0);
        if (p < c.npencil) {
            const int k = p / c.NXH, i = p % c.NXH;
            const double kx = c.kx[i], kz = c.kz[k];
            const size_t o = (size_t)p * N + j;
            const double2 a = ld2(lx + o), b = ld2(lz + o);
            tt = caddc(imul(kx, a), imul(kz, b));
            Hg[o] = csubc(imul(kz, a), imul(kx, b));
            if (p == 0) {
                c.means[(size_t)(M_LMBX * 3 + ns) * N + j] = a;
                c.means[(size_t)(M_LMBZ * 3 + ns) * N + j] = b;
            }
        }
        st[pp][j] = tt;
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row >= N) return;
    double2 acc[TP];
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(0, 0);
    for (int j = 0; j < N; ++j) {
        const double d = c.DY1[(size_t)j * N + row];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            acc[pp].x = fma(d, st[pp][j].x, acc[pp].x);
            acc[pp].y = fma(d, st[pp][j].y, acc[pp].y);
        }
    }
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) {
        const int p = p0 + pp;
        if (p >= c.npencil) break;
        const int k = p / c.NXH, i = p % c.NXH;
        const double k2 = k2_of(c, k, i);
        const size_t o = (size_t)p * N + row;
        Hv[o] = csubc(cscale(ld2(ly + o), -k2), acc[pp]);
    }
}

__global__ void k_advance(Chan c) {
    if (threadIdx.x == 0) {
        c.ctrl[C_RING] = (c.ctrl[C_RING] + 1) % 3;
        c.ctrl[C_STEP] += 1;
    }
}

// ---------------------------------------------------------------------------
// initialisation (src/solver.cpp:218-238, 379-425); ring = 0, current level in slot 0
// ---------------------------------------------------------------------------
template <int TP>
__global__ __launch_bounds__(256) void k_vor_from_vel(Chan c) {
    __shared__ double2 su[TP][kMaxNY];
    __shared__ double2 sw[TP][kMaxNY];
    const int N = c.NY;
    const double2 *u = c.cur, *w = c.cur + c.W, *v = hist(c, H_V, 0);
    double2 *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W, *oy = hist(c, H_G, 0);
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 a = c2(0, 0), b = c2(0, 0);
        if (p < c.npencil) {
            a = u[(size_t)p * N + j];
            b = w[(size_t)p * N + j];
        }
        su[pp][j] = a;
        sw[pp][j] = b;
    }
    __syncthreads();
    const int row = threadIdx.x;
    if (row >= N) return;
    double2 d1u[TP], d1w[TP];
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) d1u[pp] = d1w[pp] = c2(0, 0);
    for (int j = 0; j < N; ++j) {
        const double a = c.DY1[(size_t)j * N + row];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            d1u[pp].x = fma(a, su[pp][j].x, d1u[pp].x);
            d1u[pp].y = fma(a, su[pp][j].y, d1u[pp].y);
            d1w[pp].x = fma(a, sw[pp][j].x, d1w[pp].x);
            d1w[pp].y = fma(a, sw[pp][j].y, d1w[pp].y);
        }
    }
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) {
        const int p = p0 + pp;
        if (p >= c.npencil) break;
        const int k = p / c.NXH, i = p % c.NXH;
        const double kx = c.kx[i], kz = c.kz[k];
        const size_t o = (size_t)p * N + row;
        const double2 vv = v[o];
        ox[o] = csubc(d1w[pp], imul(kz, vv));
        oy[o] = csubc(imul(kz, su[pp][row]), imul(kx, sw[pp][row]));
        oz[o] = csubc(imul(kx, vv), d1u[pp]);
    }
}

// v, omega_y sit in slot 0 and Hv, Hg in slot 1 (written with ring 0): copy each to all levels
__global__ void k_init_history(Chan c) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < c.W; e += (size_t)gridDim.x * blockDim.x) {
        const double2 v = hist(c, H_V, 0)[e], g = hist(c, H_G, 0)[e];
        const double2 hv = hist(c, H_HV, 1)[e], hg = hist(c, H_HG, 1)[e];
        for (int s = 0; s < 3; ++s) {
            hist(c, H_V, s)[e] = v;
            hist(c, H_G, s)[e] = g;
            hist(c, H_HV, s)[e] = hv;
            hist(c, H_HG, s)[e] = hg;
        }
    }
}

template <typename LC>
__global__ __launch_bounds__(256) void k_init_mean(Chan c) {
    const int N = c.NY;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb), *lz = lx + 2 * c.W;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double2 a = ld2(lx + j), b = ld2(lz + j), u = c.cur[j], w = c.cur[c.W + j];
        for (int s = 0; s < 3; ++s) {
            c.means[(size_t)(M_LMBX * 3 + s) * N + j] = a;
            c.means[(size_t)(M_LMBZ * 3 + s) * N + j] = b;
            c.means[(size_t)(M_VELX * 3 + s) * N + j] = u;
            c.means[(size_t)(M_VELZ * 3 + s) * N + j] = w;
        }
    }
    __syncthreads();
    mean_friction(c, c.cur, c.cur + c.W, true, 0, 0);
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kTP = 8;

static int blocks_tp(const Chan &c) { return (c.npencil + kTP - 1) / kTP; }

void launch_y_rhs(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_y_rhs<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c); }

// (rows per thread, columns per thread, column groups) by NY: the tile holds
// NY rows and NY + 4 columns (right-hand sides) in registers
#define OWT_LU_DISPATCH(N, BODY)                                                                   \
    do {                                                                                           \
        const int n_ = (N);                                                                        \
        if (n_ <= 32) { constexpr int RPT = 1, CPT = 3, TC = 16; BODY; }                           \
        else if (n_ <= 64) { constexpr int RPT = 2, CPT = 5, TC = 16; BODY; }                      \
        else if (n_ <= 96) { constexpr int RPT = 3, CPT = 7, TC = 16; BODY; }                      \
        else if (n_ <= 128) { constexpr int RPT = 4, CPT = 9, TC = 16; BODY; }                     \
        else if (n_ <= 160) { constexpr int RPT = 5, CPT = 11, TC = 16; BODY; }                    \
        else if (n_ <= 192) { constexpr int RPT = 6, CPT = 13, TC = 16; BODY; }                    \
        else if (n_ <= 212) { constexpr int RPT = 7, CPT = 9, TC = 24; BODY; }                     \
        else if (n_ <= 224) { constexpr int RPT = 7, CPT = 10, TC = 24; BODY; }                    \
        else { constexpr int RPT = 8, CPT = 9, TC = 32; BODY; }                                    \
    } while (0)

void launch_lu_plain(const Chan &c, hipStream_t s) {
    OWT_LU_DISPATCH(c.NY, hipLaunchKernelGGL((k_lu<RPT, CPT, TC>), dim3(c.npencil, 2), dim3(32 * TC), 0, s, c, (int)LU_PLAIN));
}

void launch_lu_factor_store(const Chan &c, hipStream_t s) {
    OWT_LU_DISPATCH(c.NY, hipLaunchKernelGGL((k_lu<RPT, CPT, TC>), dim3(c.nsys, 2), dim3(32 * TC), 0, s, c, (int)LU_STORE));
}

#define OWT_NQ_DISPATCH(N, BODY)                                      \
    do {                                                              \
        const int nq_ = ((N) + 63) / 64;                              \
        if (nq_ <= 1) { constexpr int NQ = 1; BODY; }                 \
        else if (nq_ == 2) { constexpr int NQ = 2; BODY; }            \
        else if (nq_ == 3) { constexpr int NQ = 3; BODY; }            \
        else { constexpr int NQ = 4; BODY; }                          \
    } while (0)

void launch_lu_apply(const Chan &c, int grid, hipStream_t s) {
    const int waves = 2 * c.nsys;
    int blocks = (waves + 3) / 4;
    if (grid > 0 && blocks > grid) blocks = grid;
    OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_lu_apply<NQ>, dim3(blocks), dim3(256), 0, s, c));
}

void launch_horizontal(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_horizontal<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c); }

void launch_mean_rhs(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_mean_rhs, dim3(1), dim3(256), 0, s, c); }

void launch_mean_solve(const Chan &c, bool plain, hipStream_t s) {
    if (plain)
        OWT_LU_DISPATCH(c.NY, hipLaunchKernelGGL((k_lu<RPT, CPT, TC>), dim3(1), dim3(32 * TC), 0, s, c, (int)LU_MEAN));
    else
        OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_lu_apply_mean<NQ>, dim3(1), dim3(64), 0, s, c));
}

void launch_mean_post(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_mean_post, dim3(1), dim3(256), 0, s, c); }

void launch_nonlinear(const Chan &c, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL((k_nonlinear<kTP, double2>), dim3(blocks_tp(c)), dim3(256), 0, s, c);
    else
        hipLaunchKernelGGL((k_nonlinear<kTP, float2>), dim3(blocks_tp(c)), dim3(256), 0, s, c);
}

void launch_advance(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, c); }

void launch_vor_from_vel(const Chan &c, hipStream_t s) {
    hipLaunchKernelGGL(k_vor_from_vel<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c);
}

void launch_init_history(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_init_history, dim3(1024), dim3(256), 0, s, c); }

void launch_init_mean(const Chan &c, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL(k_init_mean<double2>, dim3(1), dim3(256), 0, s, c);
    else
        hipLaunchKernelGGL(k_init_mean<float2>, dim3(1), dim3(256), 0, s, c);
}

}  // namespace owt
