// Wall-normal (y-pencil) phases of the time step other than the implicit
// solves (lu_kernels.hip):
//
//   horizontal  u, w, omega_x, omega_z from v, omega_y (src/solver.cpp:516-539)
//   mean        mean flow, friction force and mean vorticity on the single
//               k = 0 pencil (src/solver.cpp:306-327, 428-514)
//   nonlinear   Hv, Hg from the Lamb vector (src/solver.cpp:240-260)
//   pressure    right-hand sides of the dynamic pressure (no reference counterpart)
//   init_*      the post-processing of the initial field (src/solver.cpp:379-425)
//
// Pencil p = k*NXH + i holds NY contiguous complex values (y fastest). Mat-vecs
// with the dense DY_n run one row per thread with the pencil vectors staged in
// LDS and the operator columns read coalesced from L2.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "device.hpp"
#include "y_common.hpp"

namespace owt {

// operator columns fetched per batch in the row-per-thread mat-vecs
constexpr int JC = 8;

typedef double owt_d4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Dense wall-normal derivatives on the matrix cores (both precisions, every
// NY <= kMaxNY). Across pencils the mat-vecs form one GEMM,
// Y (N x 2P) = DY (N x N) X (N x 2P), the columns of X being the pencils' real
// and imaginary parts. A block takes kMP = 16 pencils (32 columns) and every
// row, so each DY element read from L2 serves 32 columns; the pencils are staged
// in LDS kKC rows at a time. v_mfma_f64_16x16x4f64 per wave: A = DY rows x 4 k
// (lane l: row l%16, k l/16), B = 4 k x 16 columns (lane l: k l/16, column
// l%16), D = 16 x 16 (lane l, register r: row 4r + l/16, column l%16). Wave w
// of eight owns the row tiles w and w + 8; column c of a column tile is pencil
// c/2, real (c even) or imaginary part, and the two parts meet by a lane swap.
// ---------------------------------------------------------------------------
constexpr int kMP = 16, kKC = 32;  // kMP * kKC = 512, the block size of both kernels
static_assert(kMP * kKC == 512, "one staged element per thread");

__device__ __forceinline__ owt_d4 mfma4(double a, double b, owt_d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }

// horizontal velocity and vorticity (src/solver.cpp:516-539) with the three
// products DY1 v, DY2 v, DY1 omega_y on the matrix cores. Eight waves; wave w
// owns row tiles w, w + 8, ... (RT of them: 2 up to NY = 256, 3 above) and
// keeps the DY operands of the next four k steps in flight.
template <int RT>
__device__ __forceinline__ void horizontal_mfma_body(const Chan &c, const StepPtrs &sp, const int bid) {
    constexpr int NW = 8;
    // the block's whole columns of v and omega_y (2 x 16 x 289 x 16 B = 148 KB
    // of the 160 KB per CU), staged chunk by chunk and read again by the epilogue
    __shared__ double2 sv[kMP][kMaxNY + 1], sg[kMP][kMaxNY + 1];
    const int N = c.NY, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p0 = bid * kMP, ntile = (N + 15) / 16;
    if (p0 >= c.npencil) return;
    owt_d4 acc[RT][2][3];
    // staging: this thread's element of a chunk (kMP * kKC == blockDim)
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, sp_ = p0 + sq;
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[rt][ct][q] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    // DY operands of k step k for this lane's row. Indices are clamped into the
    // matrix instead of masked: staged columns past N are zero, so a clamped k
    // adds nothing, and rows past N are never stored. The loads are therefore
    // unconditional, which keeps the count of loads in flight static.
    auto loadA = [&](int k, double (&a1)[RT], double (&a2)[RT]) {
        const size_t kr = (size_t)(k < N ? k : N - 1) * N;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int row = (w + NW * rt) * 16 + am;
            const size_t e = kr + (row < N ? row : N - 1);
            a1[rt] = c.DY1[e];
            a2[rt] = c.DY2[e];
        }
    };
    // the next chunk's staged element, fetched into registers a chunk ahead of
    // its LDS write so that the write waits only for loads older than the ring's
    const size_t so = (size_t)(sp_ < c.npencil ? sp_ : c.npencil - 1) * N;
    double2 nv, ng;
    auto fetch = [&](int k0) {
        const int k = k0 + skk < N ? k0 + skk : N - 1;
        nv = sp.vn[so + k];
        ng = sp.gn[so + k];
    };
    // a ring of PF slots of DY operands: k step t uses slot t % PF, loaded PF
    // steps earlier, and reloads it for step t + PF. A chunk's kKC / 4 steps
    // (a multiple of PF) are unrolled, so every slot is a fixed register set
    // (a register move of an in-flight load would wait for it).
    constexpr int PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    double q1[PF][RT], q2[PF][RT];
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);  // the ring's loads stay younger than the fetch
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(4 * d + ak, q1[d], q2[d]);
    // the chunk loop is unrolled too (at most kMaxNY / kKC chunks): across a
    // loop back edge the compiler's wait counting would drain the ring
#pragma unroll
    for (int ci = 0; ci < kMaxNY / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= N) break;
        const bool live = sp_ < c.npencil && k0 + skk < N;
        sv[sq][k0 + skk] = live ? nv : c2(0, 0);
        sg[sq][k0 + skk] = live ? ng : c2(0, 0);
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((N - k0 < kKC ? N - k0 : kKC) + 3) / 4;  // k steps in this chunk
#pragma unroll
        for (int st = 0; st < kKC / 4; ++st) {
            const int d = st % PF, kk = 4 * st;
            if (st < nst) {
                double bv[2], bg[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 x = sv[ct * 8 + bq][k0 + kk + ak], y = sg[ct * 8 + bq][k0 + kk + ak];
                    bv[ct] = bc ? x.y : x.x;
                    bg[ct] = bc ? y.y : y.x;
                }
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    if (w + NW * rt >= ntile) break;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        acc[rt][ct][0] = mfma4(q1[d][rt], bv[ct], acc[rt][ct][0]);
                        acc[rt][ct][1] = mfma4(q2[d][rt], bv[ct], acc[rt][ct][1]);
                        acc[rt][ct][2] = mfma4(q1[d][rt], bg[ct], acc[rt][ct][2]);
                    }
                }
            }
            loadA(k0 + kk + 4 * PF + ak, q1[d], q2[d]);
        }
        // no barrier here: the next chunk is staged into rows no wave reads yet
    }
    double2 *u = c.cur, *wv = c.cur + c.W, *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W;
    // a lane stores rows of two pencils only (one per column tile): their
    // wavenumbers and reciprocals once, not per row
    double pkx[2], pkz[2], pk2[2], pix[2], piz[2], pir[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
        const int k = pq / c.NXH, i = pq % c.NXH;
        pkx[ct] = c.kx[i];
        pkz[ct] = c.kz[k];
        pk2[ct] = k2_of(c, k, i);
        const double d = pk2[ct] < 1.e-8 ? 1.0 : pk2[ct];
        pix[ct] = __ddiv_rn(pkx[ct], d);
        piz[ct] = __ddiv_rn(pkz[ct], d);
        pir[ct] = __ddiv_rn(1.0, d);
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int tile = w + NW * rt;
        if (tile >= ntile) break;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double m1 = acc[rt][ct][0][r], m2 = acc[rt][ct][1][r], m3 = acc[rt][ct][2][r];
                const double o1 = __shfl_xor(m1, 1), o2 = __shfl_xor(m2, 1), o3 = __shfl_xor(m3, 1);
                const int row = tile * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
                if (bc != 0 || row >= N || p >= c.npencil) continue;
                const double2 d1v = c2(m1, o1), d2v = c2(m2, o2), d1g = c2(m3, o3);
                const size_t o = (size_t)p * N + row;
                sp.dvn[o] = d2v_store(c, d2v, sv[ct * 8 + bq][row], pk2[ct]);
                const double kx = pkx[ct], kz = pkz[ct], k2 = pk2[ct];
                if (k2 < 1.e-8) continue;  // the mean pencil is set by the mean-flow kernels
                const double ix = pix[ct], iz = piz[ct], ir = pir[ct];
                const double2 vv = sv[ct * 8 + bq][row], gg = sg[ct * 8 + bq][row];
                u[o] = csubc(imul(ix, d1v), imul(iz, gg));
                wv[o] = caddc(imul(ix, gg), imul(iz, d1v));
                const double2 X = csubc(caddc(cscale(d1g, kx), cscale(d2v, kz)), cscale(vv, mul(kz, k2)));
                const double2 Yv = caddc(csubc(cscale(d1g, kz), cscale(d2v, kx)), cscale(vv, mul(kx, k2)));
                ox[o] = imul(ir, X);
                oz[o] = imul(ir, Yv);
            }
        }
    }
}
template <int RT>
__global__ __launch_bounds__(512) void k_horizontal_mfma(Chan c, StepPtrs sp) {
    horizontal_mfma_body<RT>(c, sp, blockIdx.x);
}
template <int RT>
__global__ __launch_bounds__(512) void k_horizontal_mfma_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    horizontal_mfma_body<RT>(e.c, e.sp, bid);
}

// ---------------------------------------------------------------------------
// Mixed precision: the products with DY1 (odd under the reflection
// j -> N-1-j) and DY2 (even) as two half-size products each. With h = (N+1)/2,
// e_j = x_j + x_{N-1-j}, o_j = x_j - x_{N-1-j} (j < h) and the operator's
// mirror halves A, B (h x h, from the symmetrised operator; operators.cpp):
//   P = A e, Q = B o on the rows i < h,
//   y_i = P_i + Q_i,  y_{N-1-i} = s (P_i - Q_i)   (s = +1 for DY2, -1 for DY1),
// half the matrix-core work of the full products. The operators' departure from
// mirror symmetry is the rounding of their construction (a few hundred FP64
// ulps), far below the FP32 rounding of the transform phase. NY <= 256
// (h <= 128): wave w owns row tile w of the half range with both parities, so P
// and Q of a row meet in registers. The pencils' whole columns are staged in
// LDS, each chunk from both ends of the column (k step j reads rows j and
// N-1-j), and folded as the B operands are read.
// ---------------------------------------------------------------------------
constexpr int kMirrorMaxH = 128;

// out32 null: u, w, omega_x, omega_z into Chan::cur (FP64). Otherwise the six
// fields the z inverse transforms (u, v, w, omega_x, omega_y, omega_z of the new
// level, W each) go to out32 in FP32 instead, each value the FP32 rounding the z
// inverse applies to the FP64 one as it loads it, and cur is left as it was
// (the mean pencil's values, set by the mean-flow kernel, are copied over).
__device__ __forceinline__ void horizontal_mirror_body(const Chan &c, const StepPtrs &sp, const int bid, float2 *out32) {
    __shared__ double2 sv[kMP][kMaxNY + 1], sg[kMP][kMaxNY + 1];
    const int N = c.NY, h = (N + 1) / 2, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p0 = bid * kMP;
    if (p0 >= c.npencil) return;
    const bool busy = w < (h + 15) / 16;
    // [parity: 0 = A e, 1 = B o][column tile][product: DY1 v, DY2 v, DY1 omega_y]
    owt_d4 acc[2][2][3];
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[par][ct][q] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, sp_ = p0 + sq;
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    // operands of k step k for this lane's row: A1, B1 (DY1), A2, B2 (DY2);
    // clamped indices as in the full kernel (the B operands of k >= h are zero)
    const int ra = w * 16 + am < h ? w * 16 + am : h - 1;
    const size_t hh = (size_t)h * h;
    auto loadA = [&](int k, double (&a)[4]) {
        const size_t e = (size_t)(k < h ? k : h - 1) * h + ra;
        a[0] = c.DY1m[e];
        a[1] = c.DY1m[hh + e];
        a[2] = c.DY2m[e];
        a[3] = c.DY2m[hh + e];
    };
    const size_t so = (size_t)(sp_ < c.npencil ? sp_ : c.npencil - 1) * N;
    double2 nvt, nvb, ngt, ngb;
    auto fetch = [&](int k0) {
        const int t = k0 + skk < h ? k0 + skk : h - 1;
        nvt = sp.vn[so + t];
        nvb = sp.vn[so + N - 1 - t];
        ngt = sp.gn[so + t];
        ngb = sp.gn[so + N - 1 - t];
    };
    constexpr int PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    double q[PF][4];
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(4 * d + ak, q[d]);
#pragma unroll
    for (int ci = 0; ci < kMirrorMaxH / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= h) break;
        const int t = k0 + skk;
        if (t < h) {
            const bool live = sp_ < c.npencil;
            sv[sq][t] = live ? nvt : c2(0, 0);
            sv[sq][N - 1 - t] = live ? nvb : c2(0, 0);
            sg[sq][t] = live ? ngt : c2(0, 0);
            sg[sq][N - 1 - t] = live ? ngb : c2(0, 0);
            if (out32 && live) {
                // v and omega_y as staged: runs of consecutive rows
                out32[c.W + so + t] = make_float2((float)nvt.x, (float)nvt.y);
                out32[c.W + so + N - 1 - t] = make_float2((float)nvb.x, (float)nvb.y);
                out32[4 * c.W + so + t] = make_float2((float)ngt.x, (float)ngt.y);
                out32[4 * c.W + so + N - 1 - t] = make_float2((float)ngb.x, (float)ngb.y);
            }
        }
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((h - k0 < kKC ? h - k0 : kKC) + 3) / 4;
#pragma unroll
        for (int st = 0; st < kKC / 4; ++st) {
            const int d = st % PF, kk = 4 * st;
            if (st < nst) {
                // rows j and N-1-j, staged by this chunk (rows j >= h read row 0, a
                // staged row, and are replaced by zeros)
                const int j = k0 + kk + ak;
                const bool in = j < h;
                const int jt = in ? j : 0, jb = in ? N - 1 - j : 0;
                double ev[2], ov[2], eg[2], og[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 xt = sv[ct * 8 + bq][jt], xb = sv[ct * 8 + bq][jb];
                    const double2 yt = sg[ct * 8 + bq][jt], yb = sg[ct * 8 + bq][jb];
                    const double a = bc ? xt.y : xt.x, b = bc ? xb.y : xb.x;
                    const double ga = bc ? yt.y : yt.x, gb = bc ? yb.y : yb.x;
                    ev[ct] = in ? add(a, b) : 0.0;
                    ov[ct] = in ? sub(a, b) : 0.0;
                    eg[ct] = in ? add(ga, gb) : 0.0;
                    og[ct] = in ? sub(ga, gb) : 0.0;
                }
                if (busy) {
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        acc[0][ct][0] = mfma4(q[d][0], ev[ct], acc[0][ct][0]);
                        acc[0][ct][1] = mfma4(q[d][2], ev[ct], acc[0][ct][1]);
                        acc[0][ct][2] = mfma4(q[d][0], eg[ct], acc[0][ct][2]);
                        acc[1][ct][0] = mfma4(q[d][1], ov[ct], acc[1][ct][0]);
                        acc[1][ct][1] = mfma4(q[d][3], ov[ct], acc[1][ct][1]);
                        acc[1][ct][2] = mfma4(q[d][1], og[ct], acc[1][ct][2]);
                    }
                }
            }
            loadA(k0 + kk + 4 * PF + ak, q[d]);
        }
        // no barrier here: the next chunk's rows (t and N-1-t for the next t
        // range) are disjoint from the rows this chunk's k steps read
    }
    if (!busy) return;
    double2 *u = c.cur, *wv = c.cur + c.W, *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W;
    double pkx[2], pkz[2], pk2[2], pix[2], piz[2], pir[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
        const int k = pq / c.NXH, i = pq % c.NXH;
        pkx[ct] = c.kx[i];
        pkz[ct] = c.kz[k];
        pk2[ct] = k2_of(c, k, i);
        const double d = pk2[ct] < 1.e-8 ? 1.0 : pk2[ct];
        pix[ct] = __ddiv_rn(pkx[ct], d);
        piz[ct] = __ddiv_rn(pkz[ct], d);
        pir[ct] = __ddiv_rn(1.0, d);
    }
    auto f32 = [](double2 a) { return make_float2((float)a.x, (float)a.y); };
    auto store = [&](int ct, int row, double2 d1v, double2 d2v, double2 d1g) {
        const int p = p0 + ct * 8 + bq;
        const size_t o = (size_t)p * N + row;
        sp.dvn[o] = d2v_store(c, d2v, sv[ct * 8 + bq][row], pk2[ct]);
        const double kx = pkx[ct], kz = pkz[ct], k2 = pk2[ct];
        const double2 vv = sv[ct * 8 + bq][row], gg = sg[ct * 8 + bq][row];
        if (k2 < 1.e-8) {
            // the mean pencil is set by the mean-flow kernels
            if (out32) {
                out32[o] = f32(u[o]);
                out32[2 * c.W + o] = f32(wv[o]);
                out32[3 * c.W + o] = f32(ox[o]);
                out32[5 * c.W + o] = f32(oz[o]);
            }
            return;
        }
        const double ix = pix[ct], iz = piz[ct], ir = pir[ct];
        const double2 uu = csubc(imul(ix, d1v), imul(iz, gg));
        const double2 ww = caddc(imul(ix, gg), imul(iz, d1v));
        const double2 X = csubc(caddc(cscale(d1g, kx), cscale(d2v, kz)), cscale(vv, mul(kz, k2)));
        const double2 Yv = caddc(csubc(cscale(d1g, kz), cscale(d2v, kx)), cscale(vv, mul(kx, k2)));
        const double2 xx = imul(ir, X), zz = imul(ir, Yv);
        if (out32) {
            // (v and omega_y went out as they were staged)
            out32[o] = f32(uu);
            out32[2 * c.W + o] = f32(ww);
            out32[3 * c.W + o] = f32(xx);
            out32[5 * c.W + o] = f32(zz);
        } else {
            u[o] = uu;
            wv[o] = ww;
            ox[o] = xx;
            oz[o] = zz;
        }
    };
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double P1 = acc[0][ct][0][r], Q1 = acc[1][ct][0][r];
            const double P2 = acc[0][ct][1][r], Q2 = acc[1][ct][1][r];
            const double P3 = acc[0][ct][2][r], Q3 = acc[1][ct][2][r];
            // row i: P + Q; row N-1-i: Q - P (DY1, odd), P - Q (DY2, even)
            const double t1 = add(P1, Q1), t2 = add(P2, Q2), t3 = add(P3, Q3);
            const double b1 = sub(Q1, P1), b2 = sub(P2, Q2), b3 = sub(Q3, P3);
            const double ot1 = __shfl_xor(t1, 1), ot2 = __shfl_xor(t2, 1), ot3 = __shfl_xor(t3, 1);
            const double ob1 = __shfl_xor(b1, 1), ob2 = __shfl_xor(b2, 1), ob3 = __shfl_xor(b3, 1);
            const int row = w * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
            if (bc != 0 || row >= h || p >= c.npencil) continue;
            store(ct, row, c2(t1, ot1), c2(t2, ot2), c2(t3, ot3));
            if (N - 1 - row != row) store(ct, N - 1 - row, c2(b1, ob1), c2(b2, ob2), c2(b3, ob3));
        }
    }
}
__global__ __launch_bounds__(512) void k_horizontal_mirror(Chan c, StepPtrs sp, float2 *out32) {
    horizontal_mirror_body(c, sp, blockIdx.x, out32);
}
__global__ __launch_bounds__(512) void k_horizontal_mirror_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    horizontal_mirror_body(e.c, e.sp, bid, nullptr);
}

// nonlinear source terms (src/solver.cpp:240-260) with DY1 (i kx Lx + i kz Lz)
// on the matrix cores; Hg and the mean Lamb history (the k = 0 pencil of the
// Lamb vector, src/solver.cpp:293-303) are written while staging.
// Eight waves, wave w owning row tiles w and w + 8, DY1 prefetched as in
// k_horizontal_mfma.
template <int RT, typename LC>
__device__ __forceinline__ void nonlinear_mfma_body(const Chan &c, const StepPtrs &sp, const int bid) {
    constexpr int NW = 8, PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    __shared__ double2 st[kMP][kKC + 1];
    const int N = c.NY, ns = sp.sn, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    const int p0 = bid * kMP, ntile = (N + 15) / 16;
    if (p0 >= c.npencil) return;
    // staging: this thread's element of a chunk (kMP * kKC == blockDim)
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, spn = p0 + sq;
    double skx = 0.0, skz = 0.0;
    if (spn < c.npencil) {
        skx = c.kx[spn % c.NXH];
        skz = c.kz[spn / c.NXH];
    }
    // the next chunk's Lamb components, fetched into registers a chunk ahead of
    // staging (indices clamped; the staging discards what lies outside)
    const size_t so = (size_t)(spn < c.npencil ? spn : c.npencil - 1) * N;
    LC na, nb;
    auto fetch = [&](int k0) {
        const int k = k0 + skk < N ? k0 + skk : N - 1;
        na = lx[so + k];
        nb = lz[so + k];
    };
    auto stage = [&](int k0) {
        const int k = k0 + skk;
        double2 t = c2(0, 0);
        if (spn < c.npencil && k < N) {
            const size_t o = so + k;
            const double2 a = c2(na.x, na.y), b = c2(nb.x, nb.y);
            t = caddc(imul(skx, a), imul(skz, b));
            sp.hgn[o] = csubc(imul(skz, a), imul(skx, b));
            if (spn == 0) {
                c.means[(size_t)(M_LMBX * 3 + ns) * N + k] = a;
                c.means[(size_t)(M_LMBZ * 3 + ns) * N + k] = b;
            }
        }
        st[sq][skk] = t;
    };
    owt_d4 acc[RT][2];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    // clamped, unconditional operand loads as in k_horizontal_mfma
    auto loadA = [&](int k, double (&a1)[RT]) {
        const size_t kr = (size_t)(k < N ? k : N - 1) * N;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int row = (w + NW * rt) * 16 + am;
            a1[rt] = c.DY1[kr + (row < N ? row : N - 1)];
        }
    };
    double q1[PF][RT];
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(4 * d + ak, q1[d]);
#pragma unroll
    for (int ci = 0; ci < kMaxNY / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= N) break;
        stage(k0);
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((N - k0 < kKC ? N - k0 : kKC) + 3) / 4;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = 4 * stp;
            if (stp < nst) {
                double bt[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 x = st[ct * 8 + bq][kk + ak];
                    bt[ct] = bc ? x.y : x.x;
                }
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    if (w + NW * rt >= ntile) break;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = mfma4(q1[d][rt], bt[ct], acc[rt][ct]);
                }
            }
            loadA(k0 + kk + 4 * PF + ak, q1[d]);
        }
        __syncthreads();
    }
    double pk2[2];  // k^2 of this lane's two pencils
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
        pk2[ct] = k2_of(c, pq / c.NXH, pq % c.NXH);
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int tile = w + NW * rt;
        if (tile >= ntile) break;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double m = acc[rt][ct][r], o1 = __shfl_xor(m, 1);
                const int row = tile * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
                if (bc != 0 || row >= N || p >= c.npencil) continue;
                const double k2 = pk2[ct];
                const size_t o = (size_t)p * N + row;
                sp.hvn[o] = csubc(cscale(ld2(ly + o), -k2), c2(m, o1));
            }
        }
    }
}
template <int RT, typename LC>
__global__ __launch_bounds__(512) void k_nonlinear_mfma(Chan c, StepPtrs sp) {
    nonlinear_mfma_body<RT, LC>(c, sp, blockIdx.x);
}
template <int RT, typename LC>
__global__ __launch_bounds__(512) void k_nonlinear_mfma_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    nonlinear_mfma_body<RT, LC>(e.c, e.sp, bid);
}

// The same stage with DY1 as mirror halves (mixed precision, NY <= 256; see
// k_horizontal_mirror): each chunk stages rows t and N-1-t of
// i kx Lx + i kz Lz (writing Hg and the mean Lamb history of both), and
// DY1 (.) = P + Q on row i, Q - P on row N-1-i, with P = A1 e and Q = B1 o.
template <typename LC>
__device__ __forceinline__ void nonlinear_mirror_body(const Chan &c, const StepPtrs &sp, const int bid) {
    constexpr int PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    __shared__ double2 st[2][kMP][kKC + 1];  // rows t (top) and N-1-t (bottom) of a chunk
    const int N = c.NY, h = (N + 1) / 2, ns = sp.sn, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    const int p0 = bid * kMP;
    if (p0 >= c.npencil) return;
    const bool busy = w < (h + 15) / 16;
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, spn = p0 + sq;
    double skx = 0.0, skz = 0.0;
    if (spn < c.npencil) {
        skx = c.kx[spn % c.NXH];
        skz = c.kz[spn / c.NXH];
    }
    const size_t so = (size_t)(spn < c.npencil ? spn : c.npencil - 1) * N;
    LC nat, nbt, nab, nbb;
    auto fetch = [&](int k0) {
        const int t = k0 + skk < h ? k0 + skk : h - 1;
        nat = lx[so + t];
        nbt = lz[so + t];
        nab = lx[so + N - 1 - t];
        nbb = lz[so + N - 1 - t];
    };
    // one row's staged value, with Hg and the mean Lamb history written
    auto row_term = [&](int k, LC la, LC lb) {
        const size_t o = so + k;
        const double2 a = c2(la.x, la.y), b = c2(lb.x, lb.y);
        sp.hgn[o] = csubc(imul(skz, a), imul(skx, b));
        if (spn == 0) {
            c.means[(size_t)(M_LMBX * 3 + ns) * N + k] = a;
            c.means[(size_t)(M_LMBZ * 3 + ns) * N + k] = b;
        }
        return caddc(imul(skx, a), imul(skz, b));
    };
    auto stage = [&](int k0) {
        const int t = k0 + skk;
        double2 tt = c2(0, 0), tb = c2(0, 0);
        if (spn < c.npencil && t < h) {
            tt = row_term(t, nat, nbt);
            tb = row_term(N - 1 - t, nab, nbb);  // the middle row of odd N: tb = tt
        }
        st[0][sq][skk] = tt;
        st[1][sq][skk] = tb;
    };
    owt_d4 acc[2][2];  // [parity: A1 e, B1 o][column tile]
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[par][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    const int ra = w * 16 + am < h ? w * 16 + am : h - 1;
    const size_t hh = (size_t)h * h;
    auto loadA = [&](int k, double (&a)[2]) {
        const size_t e = (size_t)(k < h ? k : h - 1) * h + ra;
        a[0] = c.DY1m[e];
        a[1] = c.DY1m[hh + e];
    };
    double q[PF][2];
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(4 * d + ak, q[d]);
#pragma unroll
    for (int ci = 0; ci < kMirrorMaxH / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= h) break;
        stage(k0);
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((h - k0 < kKC ? h - k0 : kKC) + 3) / 4;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = 4 * stp;
            if (stp < nst) {
                // rows past h were staged as zeros
                double e[2], o[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 xt = st[0][ct * 8 + bq][kk + ak], xb = st[1][ct * 8 + bq][kk + ak];
                    const double a = bc ? xt.y : xt.x, b = bc ? xb.y : xb.x;
                    e[ct] = add(a, b);
                    o[ct] = sub(a, b);
                }
                if (busy) {
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        acc[0][ct] = mfma4(q[d][0], e[ct], acc[0][ct]);
                        acc[1][ct] = mfma4(q[d][1], o[ct], acc[1][ct]);
                    }
                }
            }
            loadA(k0 + kk + 4 * PF + ak, q[d]);
        }
        __syncthreads();
    }
    if (!busy) return;
    double pk2[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
        pk2[ct] = k2_of(c, pq / c.NXH, pq % c.NXH);
    }
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const double P = acc[0][ct][r], Q = acc[1][ct][r];
            const double top = add(P, Q), bot = sub(Q, P);
            const double otop = __shfl_xor(top, 1), obot = __shfl_xor(bot, 1);
            const int row = w * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
            if (bc != 0 || row >= h || p >= c.npencil) continue;
            const double k2 = pk2[ct];
            size_t o = (size_t)p * N + row;
            sp.hvn[o] = csubc(cscale(ld2(ly + o), -k2), c2(top, otop));
            if (N - 1 - row != row) {
                o = (size_t)p * N + N - 1 - row;
                sp.hvn[o] = csubc(cscale(ld2(ly + o), -k2), c2(bot, obot));
            }
        }
    }
}
template <typename LC>
__global__ __launch_bounds__(512) void k_nonlinear_mirror(Chan c, StepPtrs sp) {
    nonlinear_mirror_body<LC>(c, sp, blockIdx.x);
}
template <typename LC>
__global__ __launch_bounds__(512) void k_nonlinear_mirror_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    nonlinear_mirror_body<LC>(e.c, e.sp, bid);
}

// ---------------------------------------------------------------------------
// omega_y of the new level, batched method (src/solver.cpp:574-603 for the
// omega_y equation): rows 0 and N-1 are omega_y = 0, and on the interior
//   (c I - nu D) x = r,  c = nu k^2 + gamma0/dt,  r = sum beta Hg + sum alpha omega_y / dt,
// with D = Q B Q^-1 (eigen.cpp): x = Q (c I - nu B)^-1 Q^-1 r. Per block of 16
// pencils (32 real columns) two GEMMs on the matrix cores:
//   Z = Q^-1 R    R staged chunk by chunk in LDS as it is formed from the six
//                 history fields; then each Z row scaled by its block of
//                 (c I - nu B)^-1 (a 2 x 2 pair's partner row is the lane 16
//                 apart) and the whole Z kept in LDS
//   X = Q Z       and stored with the zero wall rows.
// Q and Q^-1 are shared by every pencil and read from L2, like the DY matrices
// of the derivative kernels; the operands of the next four k steps are in flight.
// ---------------------------------------------------------------------------
template <int RT>
__device__ __forceinline__ void helmholtz_g_body(const Chan &c, const StepPtrs &sp, const int bid) {
    constexpr int NW = 8, PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    __shared__ double2 st[kMP][kKC + 1];
    extern __shared__ double szs[];  // Z: 32 columns of ZS = 16 ntile + 1 rows
    const int N = c.NY, M = N - 2, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p0 = bid * kMP, ntile = (M + 15) / 16, ZS = 16 * ntile + 1;
    if (p0 >= c.npencil) return;
    const double dt = c.dt, nu = c.nu;
    // staging: this thread's element of a chunk (kMP * kKC == blockDim)
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, spn = p0 + sq;
    const size_t so = (size_t)(spn < c.npencil ? spn : c.npencil - 1) * N;
    double2 ng[3], nh[3];
    auto fetch = [&](int k0) {
        const int j = (k0 + skk < M ? k0 + skk : M - 1) + 1;  // interior row j of y
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            ng[l] = sp.g[l][so + j];
            nh[l] = sp.hg[l][so + j];
        }
    };
    auto stage = [&](int k0) {
        double2 t = c2(0, 0);
        if (spn < c.npencil && k0 + skk < M) {
            const double2 a = caddc(caddc(cscale(ng[0], AL0), cscale(ng[1], AL1)), cscale(ng[2], AL2));
            const double2 h = caddc(caddc(cscale(nh[0], BE0), cscale(nh[1], BE1)), cscale(nh[2], BE2));
            t = caddc(h, cdiv(a, dt));
        }
        st[sq][skk] = t;
    };
    owt_d4 acc[RT][2];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    // clamped, unconditional operand loads of an M x M column-major matrix
    auto loadA = [&](const double *A, int k, double (&a1)[RT]) {
        const size_t kr = (size_t)(k < M ? k : M - 1) * M;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
            const int row = (w + NW * rt) * 16 + am;
            a1[rt] = A[kr + (row < M ? row : M - 1)];
        }
    };
    double q1[PF][RT];
    // ---- Z = Q^-1 R ----
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(c.hQi, 4 * d + ak, q1[d]);
#pragma unroll
    for (int ci = 0; ci < kMaxNY / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= M) break;
        stage(k0);
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((M - k0 < kKC ? M - k0 : kKC) + 3) / 4;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = 4 * stp;
            if (stp < nst) {
                double bt[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 x = st[ct * 8 + bq][kk + ak];
                    bt[ct] = bc ? x.y : x.x;
                }
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    if (w + NW * rt >= ntile) break;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = mfma4(q1[d][rt], bt[ct], acc[rt][ct]);
                }
            }
            loadA(c.hQi, k0 + kk + 4 * PF + ak, q1[d]);
        }
        __syncthreads();
    }
    // operands of the second product start loading under the scaling
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(c.hQ, 4 * d + ak, q1[d]);
    // ---- Z <- (c I - nu B)^-1 Z, into LDS ----
    double cp[2];  // c of this lane's two pencils
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
        const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
        cp[ct] = add(mul(nu, k2_of(c, pq / c.NXH, pq % c.NXH)), __ddiv_rn(G0, dt));
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int tile = w + NW * rt;
        if (tile >= ntile) break;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = tile * 16 + 4 * r + (lane >> 4);
            const int ic = i < M ? i : M - 1;
            const double a = c.hlre[ic], b = c.hlim[ic];
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
                const double y = acc[rt][ct][r];
                const double yo = __shfl_xor(y, 16);  // row i ^ 1 of the same column
                const double cc = sub(cp[ct], mul(nu, a));
                double z;
                if (b == 0.0) {
                    z = __ddiv_rn(y, cc);
                } else {
                    const double nb = mul(nu, b), det = add(mul(cc, cc), mul(nb, nb));
                    z = (i & 1) ? __ddiv_rn(sub(mul(cc, y), mul(nb, yo)), det) : __ddiv_rn(add(mul(cc, y), mul(nb, yo)), det);
                }
                szs[(ct * 16 + am) * ZS + i] = i < M ? z : 0.0;
            }
        }
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    // ---- X = Q Z ----
    const int nk = (M + 3) / 4;  // k steps; Z rows M..4 nk - 1 are zero
#pragma unroll
    for (int ci = 0; ci < kMaxNY / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= M) break;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = k0 + 4 * stp;
            if (kk < 4 * nk) {
                double bt[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) bt[ct] = szs[(ct * 16 + am) * ZS + kk + ak];
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) {
                    if (w + NW * rt >= ntile) break;
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = mfma4(q1[d][rt], bt[ct], acc[rt][ct]);
                }
            }
            loadA(c.hQ, kk + 4 * PF + ak, q1[d]);
        }
    }
    // store omega_y: interior rows from X, wall rows 0
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int tile = w + NW * rt;
        if (tile >= ntile) break;
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double m = acc[rt][ct][r], o1 = __shfl_xor(m, 1);
                const int i = tile * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
                if (bc != 0 || i >= M || p >= c.npencil) continue;
                sp.gn[(size_t)p * N + i + 1] = c2(m, o1);
            }
        }
    }
    if (threadIdx.x < 2 * kMP) {
        const int p = p0 + (threadIdx.x >> 1);
        if (p < c.npencil) sp.gn[(size_t)p * N + ((threadIdx.x & 1) ? N - 1 : 0)] = c2(0, 0);
    }
}
template <int RT>
__global__ __launch_bounds__(512) void k_helmholtz_g(Chan c, StepPtrs sp) { helmholtz_g_body<RT>(c, sp, blockIdx.x); }
template <int RT>
__global__ __launch_bounds__(512) void k_helmholtz_g_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    helmholtz_g_body<RT>(e.c, e.sp, bid);
}

// omega_y of the mixed-precision step by the mirror-parity blocks of the
// symmetrised interior DY2 (helmholtz_mirror_basis, eigen.cpp). M = N - 2,
// hm = (M+1)/2; interior index j is row j + 1 of y. The right-hand side is
// folded as e_j = r_j + r_{M-1-j}, o_j = r_j - r_{M-1-j} (j < hm), then
//   Z_E = (Q_E^-1 / 2) e,  Z_O = (Q_O^-1 / 2) o,
//   scaled by (c I - nu B_E)^-1 and (c I - nu B_O)^-1,
//   X_E = Q_E Z_E,  X_O = Q_O Z_O,  x_j = X_E,j + X_O,j,  x_{M-1-j} = X_E,j - X_O,j:
// four half-size products instead of two full ones. NY <= 256: wave w owns row
// tile w of the half range with both parities; each chunk stages rows t and
// M-1-t of the right-hand side and keeps their sum and difference.
__device__ __forceinline__ void helmholtz_mirror_body(const Chan &c, const StepPtrs &sp, const int bid) {
    constexpr int PF = 4;
    static_assert((kKC / 4) % PF == 0, "chunks hold whole rounds of the operand ring");
    __shared__ double2 st[2][kMP][kKC + 1];  // e, o of a chunk
    extern __shared__ double szs[];          // Z_E, Z_O: 2 x 32 columns of ZS = 16 ntile + 1 rows
    const int N = c.NY, M = N - 2, hm = (M + 1) / 2, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p0 = bid * kMP, ntile = (hm + 15) / 16, ZS = 16 * ntile + 1;
    if (p0 >= c.npencil) return;
    const bool busy = w < ntile;
    const double dt = c.dt, nu = c.nu;
    const size_t hh = (size_t)hm * hm;
    const double *QEi = c.hmir, *QE = QEi + hh, *QOi = QE + hh, *QO = QOi + hh, *LR = QO + hh;
    const int sq = threadIdx.x / kKC, skk = threadIdx.x % kKC, spn = p0 + sq;
    const size_t so = (size_t)(spn < c.npencil ? spn : c.npencil - 1) * N;
    double2 ngt[3], nht[3], ngb[3], nhb[3];
    auto fetch = [&](int k0) {
        const int t = k0 + skk < hm ? k0 + skk : hm - 1;
        const size_t jt = so + t + 1, jb = so + M - t;  // rows of interior t and M-1-t
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            ngt[l] = sp.g[l][jt];
            nht[l] = sp.hg[l][jt];
            ngb[l] = sp.g[l][jb];
            nhb[l] = sp.hg[l][jb];
        }
    };
    auto rhs = [&](const double2 (&g)[3], const double2 (&hg)[3]) {
        const double2 a = caddc(caddc(cscale(g[0], AL0), cscale(g[1], AL1)), cscale(g[2], AL2));
        const double2 h = caddc(caddc(cscale(hg[0], BE0), cscale(hg[1], BE1)), cscale(hg[2], BE2));
        return caddc(h, cdiv(a, dt));
    };
    auto stage = [&](int k0) {
        const int t = k0 + skk;
        double2 e = c2(0, 0), o = c2(0, 0);
        if (spn < c.npencil && t < hm) {
            const double2 rt = rhs(ngt, nht), rb = rhs(ngb, nhb);  // the middle of odd M: o = 0
            e = caddc(rt, rb);
            o = csubc(rt, rb);
        }
        st[0][sq][skk] = e;
        st[1][sq][skk] = o;
    };
    owt_d4 acc[2][2];  // [parity][column tile]
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[par][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    const int am = lane & 15, ak = lane >> 4, bq = (lane & 15) >> 1, bc = lane & 1;
    const int ra = w * 16 + am < hm ? w * 16 + am : hm - 1;
    auto loadA = [&](const double *AE, const double *AO, int k, double (&a)[2]) {
        const size_t e = (size_t)(k < hm ? k : hm - 1) * hm + ra;
        a[0] = AE[e];
        a[1] = AO[e];
    };
    double q[PF][2];
    // ---- Z = Q^-1 (e, o) / 2 ----
    fetch(0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(QEi, QOi, 4 * d + ak, q[d]);
#pragma unroll
    for (int ci = 0; ci < kMirrorMaxH / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= hm) break;
        stage(k0);
        __syncthreads();
        fetch(k0 + kKC);
        const int nst = ((hm - k0 < kKC ? hm - k0 : kKC) + 3) / 4;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = 4 * stp;
            if (stp < nst) {
                double be[2], bo[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double2 xe = st[0][ct * 8 + bq][kk + ak], xo = st[1][ct * 8 + bq][kk + ak];
                    be[ct] = bc ? xe.y : xe.x;
                    bo[ct] = bc ? xo.y : xo.x;
                }
                if (busy) {
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        acc[0][ct] = mfma4(q[d][0], be[ct], acc[0][ct]);
                        acc[1][ct] = mfma4(q[d][1], bo[ct], acc[1][ct]);
                    }
                }
            }
            loadA(QEi, QOi, k0 + kk + 4 * PF + ak, q[d]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int d = 0; d < PF; ++d) loadA(QE, QO, 4 * d + ak, q[d]);
    // ---- Z <- (c I - nu B)^-1 Z per parity, into LDS ----
    if (busy) {
        double cp[2];
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
            const int p = p0 + ct * 8 + bq, pq = p < c.npencil ? p : 0;
            cp[ct] = add(mul(nu, k2_of(c, pq / c.NXH, pq % c.NXH)), __ddiv_rn(G0, dt));
        }
#pragma unroll
        for (int par = 0; par < 2; ++par) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = w * 16 + 4 * r + (lane >> 4);
                const int ic = i < hm ? i : hm - 1;
                const double a = LR[2 * par * hm + ic], b = LR[(2 * par + 1) * hm + ic];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    const double y = acc[par][ct][r];
                    const double yo = __shfl_xor(y, 16);  // row i ^ 1 of the same column
                    const double cc = sub(cp[ct], mul(nu, a));
                    double z;
                    if (b == 0.0) {
                        z = __ddiv_rn(y, cc);
                    } else {
                        const double nb = mul(nu, b), det = add(mul(cc, cc), mul(nb, nb));
                        z = (i & 1) ? __ddiv_rn(sub(mul(cc, y), mul(nb, yo)), det) : __ddiv_rn(add(mul(cc, y), mul(nb, yo)), det);
                    }
                    szs[(par * 32 + ct * 16 + am) * ZS + i] = i < hm ? z : 0.0;
                }
            }
        }
    }
#pragma unroll
    for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[par][ct] = owt_d4{0.0, 0.0, 0.0, 0.0};
    __syncthreads();
    // ---- X = Q Z ----
    const int nk = (hm + 3) / 4;  // k steps; Z rows hm..4 nk - 1 are zero
#pragma unroll
    for (int ci = 0; ci < kMirrorMaxH / kKC; ++ci) {
        const int k0 = ci * kKC;
        if (k0 >= hm) break;
#pragma unroll
        for (int stp = 0; stp < kKC / 4; ++stp) {
            const int d = stp % PF, kk = k0 + 4 * stp;
            if (kk < 4 * nk) {
                double be[2], bo[2];
#pragma unroll
                for (int ct = 0; ct < 2; ++ct) {
                    be[ct] = szs[(ct * 16 + am) * ZS + kk + ak];
                    bo[ct] = szs[(32 + ct * 16 + am) * ZS + kk + ak];
                }
                if (busy) {
#pragma unroll
                    for (int ct = 0; ct < 2; ++ct) {
                        acc[0][ct] = mfma4(q[d][0], be[ct], acc[0][ct]);
                        acc[1][ct] = mfma4(q[d][1], bo[ct], acc[1][ct]);
                    }
                }
            }
            loadA(QE, QO, kk + 4 * PF + ak, q[d]);
        }
    }
    // store omega_y: interior rows j + 1 and M - j from X, wall rows 0
    if (busy) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double XE = acc[0][ct][r], XO = acc[1][ct][r];
                const double top = add(XE, XO), bot = sub(XE, XO);
                const double otop = __shfl_xor(top, 1), obot = __shfl_xor(bot, 1);
                const int i = w * 16 + 4 * r + (lane >> 4), p = p0 + ct * 8 + bq;
                if (bc != 0 || i >= hm || p >= c.npencil) continue;
                sp.gn[(size_t)p * N + i + 1] = c2(top, otop);
                if (M - 1 - i != i) sp.gn[(size_t)p * N + M - i] = c2(bot, obot);
            }
        }
    }
    if (threadIdx.x < 2 * kMP) {
        const int p = p0 + (threadIdx.x >> 1);
        if (p < c.npencil) sp.gn[(size_t)p * N + ((threadIdx.x & 1) ? N - 1 : 0)] = c2(0, 0);
    }
}
__global__ __launch_bounds__(512) void k_helmholtz_mirror(Chan c, StepPtrs sp) { helmholtz_mirror_body(c, sp, blockIdx.x); }
__global__ __launch_bounds__(512) void k_helmholtz_mirror_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    helmholtz_mirror_body(e.c, e.sp, bid);
}
// Z of the mirror kernel: 2 x 32 columns of 16 ntile + 1 rows
static size_t helmholtz_mirror_lds(int ny) {
    const int hm = (ny - 2 + 1) / 2, ntile = (hm + 15) / 16;
    return sizeof(double) * 64 * (size_t)(16 * ntile + 1);
}

// ---------------------------------------------------------------------------
// The plain method's derivative stages in the reference's arithmetic: each
// mat-vec in column order (src/linalg.hpp:434-453) with every product and sum
// rounded separately, TP pencils per block staged in LDS, rows looped over the
// block's threads.
// ---------------------------------------------------------------------------
// solve_horizontal (src/solver.cpp:516-539) and DY2 v of the new level
template <int TP>
__global__ __launch_bounds__(256) void k_horizontal_exact(Chan c, StepPtrs sp) {
    __shared__ double2 sv[TP][kMaxNY], sg[TP][kMaxNY];
    const int N = c.NY, p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        const bool in = p < c.npencil;
        sv[pp][j] = in ? sp.vn[(size_t)p * N + j] : c2(0, 0);
        sg[pp][j] = in ? sp.gn[(size_t)p * N + j] : c2(0, 0);
    }
    __syncthreads();
    double2 *u = c.cur, *wv = c.cur + c.W, *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W;
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 d1v[TP], d2v[TP], d1g[TP];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) d1v[pp] = d2v[pp] = d1g[pp] = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double a1 = c.DY1[(size_t)j * N + row], a2 = c.DY2[(size_t)j * N + row];
#pragma unroll
            for (int pp = 0; pp < TP; ++pp) {
                const double2 x = sv[pp][j], g = sg[pp][j];
                d1v[pp] = c2(add(d1v[pp].x, mul(a1, x.x)), add(d1v[pp].y, mul(a1, x.y)));
                d2v[pp] = c2(add(d2v[pp].x, mul(a2, x.x)), add(d2v[pp].y, mul(a2, x.y)));
                d1g[pp] = c2(add(d1g[pp].x, mul(a1, g.x)), add(d1g[pp].y, mul(a1, g.y)));
            }
        }
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const size_t o = (size_t)p * N + row;
            const int k = p / c.NXH, i = p % c.NXH;
            const double kx = c.kx[i], kz = c.kz[k], k2 = k2_of(c, k, i);
            sp.dvn[o] = d2v_store(c, d2v[pp], sv[pp][row], k2);
            if (k2 < 1.e-8) continue;  // the mean pencil is set by the mean-flow kernel
            const double ix = __ddiv_rn(kx, k2), iz = __ddiv_rn(kz, k2), ir = __ddiv_rn(1.0, k2);
            const double2 vv = sv[pp][row], gg = sg[pp][row];
            u[o] = csubc(imul(ix, d1v[pp]), imul(iz, gg));
            wv[o] = caddc(imul(ix, gg), imul(iz, d1v[pp]));
            const double2 X = csubc(caddc(cscale(d1g[pp], kx), cscale(d2v[pp], kz)), cscale(vv, mul(kz, k2)));
            const double2 Yv = caddc(csubc(cscale(d1g[pp], kz), cscale(d2v[pp], kx)), cscale(vv, mul(kx, k2)));
            ox[o] = imul(ir, X);
            oz[o] = imul(ir, Yv);
        }
    }
}

// nonlinear_source (src/solver.cpp:240-260) and the mean Lamb history
// (src/solver.cpp:293-303)
template <int TP, typename LC>
__global__ __launch_bounds__(256) void k_nonlinear_exact(Chan c, StepPtrs sp) {
    __shared__ double2 st[TP][kMaxNY];
    const int N = c.NY, ns = sp.sn, p0 = blockIdx.x * TP;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 t = c2(0, 0);
        if (p < c.npencil) {
            const size_t o = (size_t)p * N + j;
            const double kx = c.kx[p % c.NXH], kz = c.kz[p / c.NXH];
            const double2 a = ld2(lx + o), b = ld2(lz + o);
            t = caddc(imul(kx, a), imul(kz, b));
            sp.hgn[o] = csubc(imul(kz, a), imul(kx, b));
            if (p == 0) {
                c.means[(size_t)(M_LMBX * 3 + ns) * N + j] = a;
                c.means[(size_t)(M_LMBZ * 3 + ns) * N + j] = b;
            }
        }
        st[pp][j] = t;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 d[TP];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) d[pp] = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double a = c.DY1[(size_t)j * N + row];
#pragma unroll
            for (int pp = 0; pp < TP; ++pp) d[pp] = c2(add(d[pp].x, mul(a, st[pp][j].x)), add(d[pp].y, mul(a, st[pp][j].y)));
        }
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const size_t o = (size_t)p * N + row;
            const double k2 = k2_of(c, p / c.NXH, p % c.NXH);
            sp.hvn[o] = csubc(cscale(ld2(ly + o), -k2), d[pp]);
        }
    }
}

// ---------------------------------------------------------------------------
// mean flow (src/solver.cpp:428-514) and friction force (src/solver.cpp:306-327)
// ---------------------------------------------------------------------------
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
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 du = c2(0, 0), dw = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double a = c.DY1[(size_t)j * N + row];
            du = c2(add(du.x, mul(a, sU[j].x)), add(du.y, mul(a, sU[j].y)));
            dw = c2(add(dw.x, mul(a, sW[j].x)), add(dw.y, mul(a, sW[j].y)));
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
        double *d = c.diag + (size_t)((unsigned)diag_row % (unsigned)c.diag_cap) * 3;
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

// One block: right-hand sides of the mean streamwise and spanwise momentum
// equations with the friction forcing (src/solver.cpp:452-470), the solve with
// Ag(0) (factor Fg, pivot rows Pg) in wave 0, then the friction and mean
// vorticity.
template <int NQ>
__device__ __forceinline__ void mean_body(const Chan &c, const StepPtrs &sp, const double *Fg, const int *Pg, const int bid) {
    __shared__ double S[4][64 * NQ];
    if (bid != 0) return;
    const int N = c.NY, lane = threadIdx.x & 63;
    const double dt = c.dt;
    const int step = c.ctrl[C_STEP];
    const double BE[3] = {BE0, BE1, BE2}, AL[3] = {AL0, AL1, AL2};
    for (int j = threadIdx.x; j < 64 * NQ; j += blockDim.x) {
        double2 su = c2(0, 0), sw = c2(0, 0);
        if (j > 0 && j < N - 1) {
#pragma unroll
            for (int comp = 0; comp < 2; ++comp) {
                const double2 *lm = c.means + (size_t)((comp == 0 ? M_LMBX : M_LMBZ) * 3) * N;
                const double2 *vm = c.means + (size_t)((comp == 0 ? M_VELX : M_VELZ) * 3) * N;
                const double *fr = c.fric + comp * 3;
                double2 t[3], va[3];
#pragma unroll
                for (int l = 0; l < 3; ++l) {
                    const int sl = sp.s[l];
                    const double2 m = lm[(size_t)sl * N + j];
                    t[l] = cscale(c2(add(m.x, fr[sl]), m.y), BE[l]);
                    va[l] = cscale(vm[(size_t)sl * N + j], AL[l]);
                }
                const double2 s = caddc(caddc(caddc(t[0], t[1]), t[2]), cdiv(caddc(caddc(va[0], va[1]), va[2]), dt));
                if (comp == 0) su = s;
                else sw = s;
            }
        }
        S[0][j] = su.x;
        S[1][j] = su.y;
        S[2][j] = sw.x;
        S[3][j] = sw.y;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        double t[4][NQ];
        lu_gather<NQ>(S, Pg, N, lane, t);
        lu_substitute<NQ, 4, 4, true>(Fg, N, lane, t);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int row = lane + 64 * q;
            if (row < N) {
                c.msol[row] = c2(t[0][q], t[1][q]);
                c.msol[N + row] = c2(t[2][q], t[3][q]);
            }
        }
    }
    __syncthreads();
    mean_friction(c, c.msol, c.msol + N, false, sp.sn, step + 1);
    __syncthreads();
    if (threadIdx.x == 0) c.ctrl[C_STEP] = step + 1;
}
template <int NQ>
__global__ __launch_bounds__(256) void k_mean(Chan c, StepPtrs sp, const double *Fg, const int *Pg) {
    mean_body<NQ>(c, sp, Fg, Pg, blockIdx.x);
}
template <int NQ>
__global__ __launch_bounds__(256) void k_mean_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    mean_body<NQ>(e.c, e.sp, e.Fg, e.Pg, bid);
}

// ---------------------------------------------------------------------------
// dynamic pressure P = p + |u|^2/2 of the current level (the reference has no
// pressure). The divergence of the rotational momentum equation
// du/dt = L - grad P + nu lap u (L the Lamb vector, u x omega) gives
//   rows 1..N-2:  (DY2 - k^2) P = i kx Lx + DY1 Ly + i kz Lz
// and its y component at the walls, where v = 0:
//   rows 0, N-1:  DY1 P = Ly + nu DY2 v
// The k = 0 pencil's row N-1 is its gauge row (right-hand side 0). This kernel
// writes the right-hand sides; lu_kernels.hip solves. dv is the level's stored
// wall-normal term (d2v_store), v its v: DY2 v = dv + k^2 v in mixed precision.
// ---------------------------------------------------------------------------
template <int TP, typename LC>
__global__ __launch_bounds__(256) void k_pressure_rhs(Chan c, const double2 *dv, const double2 *v, double2 *R) {
    __shared__ double2 st[TP][kMaxNY];
    const int N = c.NY;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        st[pp][j] = p < c.npencil ? ld2(ly + (size_t)p * N + j) : c2(0, 0);
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 acc[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(0, 0);
        for (int j0 = 0; j0 < N; j0 += JC) {
            double dj[JC];
    #pragma unroll
            for (int cc = 0; cc < JC; ++cc) dj[cc] = c.DY1[(size_t)(j0 + cc < N ? j0 + cc : N - 1) * N + row];
    #pragma unroll
            for (int cc = 0; cc < JC; ++cc) {
                const int j = j0 + cc;
                if (j >= N) break;
    #pragma unroll
                for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(add(acc[pp].x, mul(dj[cc], st[pp][j].x)), add(acc[pp].y, mul(dj[cc], st[pp][j].y)));
            }
        }
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const int k = p / c.NXH, i = p % c.NXH;
            const size_t o = (size_t)p * N + row;
            double2 r;
            if (row == 0 || row == N - 1) {
                const double2 d2 = c.precision == 1 ? caddc(dv[o], cscale(v[o], k2_of(c, k, i))) : dv[o];
                r = caddc(st[pp][row], cscale(d2, c.nu));
                if (row == N - 1 && k2_of(c, k, i) < 1.e-8) r = c2(0, 0);
            } else {
                r = caddc(caddc(imul(c.kx[i], ld2(lx + o)), imul(c.kz[k], ld2(lz + o))), acc[pp]);
            }
            R[o] = r;
        }
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
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 d1u[TP], d1w[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) d1u[pp] = d1w[pp] = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double a = c.DY1[(size_t)j * N + row];
    #pragma unroll
            for (int pp = 0; pp < TP; ++pp) {
                d1u[pp] = c2(add(d1u[pp].x, mul(a, su[pp][j].x)), add(d1u[pp].y, mul(a, su[pp][j].y)));
                d1w[pp] = c2(add(d1w[pp].x, mul(a, sw[pp][j].x)), add(d1w[pp].y, mul(a, sw[pp][j].y)));
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
}

// DY2 v of the initial field (slot 0), with the mat-vec of the horizontal kernel
// (stored as d2v_store does: minus k^2 v in mixed precision)
template <int TP>
__global__ __launch_bounds__(256) void k_init_d2v(Chan c) {
    __shared__ double2 sv[TP][kMaxNY];
    const int N = c.NY;
    const double2 *v = hist(c, H_V, 0);
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        sv[pp][j] = p < c.npencil ? v[(size_t)p * N + j] : c2(0, 0);
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 acc[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(0, 0);
        for (int j = 0; j < N; ++j) {
            const double b = c.DY2[(size_t)j * N + row];
    #pragma unroll
            for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(add(acc[pp].x, mul(b, sv[pp][j].x)), add(acc[pp].y, mul(b, sv[pp][j].y)));
        }
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp)
            if (p0 + pp < c.npencil)
                c.d2v[(size_t)(p0 + pp) * N + row] =
                    d2v_store(c, acc[pp], sv[pp][row], k2_of(c, (p0 + pp) / c.NXH, (p0 + pp) % c.NXH));
    }
}

// v, omega_y and DY2 v sit in slot 0 and Hv, Hg in slot 1 (written with ring 0):
// copy each to all levels
__global__ void k_init_history(Chan c) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < c.W; e += (size_t)gridDim.x * blockDim.x) {
        const double2 v = hist(c, H_V, 0)[e], g = hist(c, H_G, 0)[e], dv = c.d2v[e];
        const double2 hv = hist(c, H_HV, 1)[e], hg = hist(c, H_HG, 1)[e];
        for (int s = 0; s < 3; ++s) {
            hist(c, H_V, s)[e] = v;
            hist(c, H_G, s)[e] = g;
            hist(c, H_HV, s)[e] = hv;
            hist(c, H_HG, s)[e] = hg;
            c.d2v[(size_t)s * c.W + e] = dv;
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
// Adjoints (transposes) of the wall-normal stages, FP64. The channel c is an
// adjoint state: each buffer holds the adjoint of the forward buffer of the same
// name and slot. A stage that overwrites a forward buffer consumes that buffer's
// adjoint (reads it, then zeroes it) and adds to the adjoints of its inputs.
// Complex scalars enter conjugated: the transpose of v -> i s v is v -> -i s v.
// ---------------------------------------------------------------------------

// (DY^T x)_row = sum_j DY(j, row) x_j: column `row` of the column-major DY
template <int TP>
__device__ __forceinline__ void matvec_T(const double *DY, const double2 (*sx)[kMaxNY], int N, int row, double2 (&acc)[TP]) {
    const double *col = DY + (size_t)row * N;
    for (int j = 0; j < N; ++j) {
        const double a = col[j];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            acc[pp].x = fma(a, sx[pp][j].x, acc[pp].x);
            acc[pp].y = fma(a, sx[pp][j].y, acc[pp].y);
        }
    }
}

// transpose of k_horizontal: adjoints of u, w, omega_x, omega_z (all pencils
// but k = 0) and of DY2 v (slot sn) into the adjoints of v and omega_y (slot sn)
template <int TP>
__global__ __launch_bounds__(256) void k_horizontal_adj(Chan c, StepPtrs sp) {
    __shared__ double2 s1v[TP][kMaxNY], s2v[TP][kMaxNY], s1g[TP][kMaxNY];
    const int N = c.NY;
    double2 *au = c.cur, *aw = c.cur + c.W, *ax = c.cur + 2 * c.W, *az = c.cur + 3 * c.W;
    double2 *av = sp.vn, *ag = sp.gn, *ad = sp.dvn;
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        double2 d1v = c2(0, 0), d2v = c2(0, 0), d1g = c2(0, 0);
        if (p < c.npencil) {
            const int k = p / c.NXH, i = p % c.NXH;
            const size_t o = (size_t)p * N + j;
            d2v = ad[o];
            const double k2 = k2_of(c, k, i);
            if (k2 >= 1.e-8) {
                const double kx = c.kx[i], kz = c.kz[k];
                const double ix = __ddiv_rn(kx, k2), iz = __ddiv_rn(kz, k2), ir = __ddiv_rn(1.0, k2);
                const double2 U = au[o], Wv = aw[o], X = ax[o], Z = az[o];
                d1v = caddc(imul(-ix, U), imul(-iz, Wv));
                d1g = caddc(imul(-mul(ir, kx), X), imul(-mul(ir, kz), Z));
                d2v = caddc(d2v, caddc(imul(-mul(ir, kz), X), imul(mul(ir, kx), Z)));
            }
        }
        s1v[pp][j] = d1v;
        s2v[pp][j] = d2v;
        s1g[pp][j] = d1g;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 gv[TP], gg[TP], tmp[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) gv[pp] = gg[pp] = tmp[pp] = c2(0, 0);
        matvec_T<TP>(c.DY1, s1v, N, row, gv);
        matvec_T<TP>(c.DY2, s2v, N, row, tmp);
        matvec_T<TP>(c.DY1, s1g, N, row, gg);
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const int k = p / c.NXH, i = p % c.NXH;
            const size_t o = (size_t)p * N + row;
            double2 v = caddc(gv[pp], tmp[pp]), g = gg[pp];
            const double k2 = k2_of(c, k, i);
            if (k2 >= 1.e-8) {
                const double kx = c.kx[i], kz = c.kz[k];
                const double ix = __ddiv_rn(kx, k2), iz = __ddiv_rn(kz, k2), ir = __ddiv_rn(1.0, k2);
                const double2 U = au[o], Wv = aw[o], X = ax[o], Z = az[o];
                v = caddc(v, caddc(imul(mul(ir, mul(kz, k2)), X), imul(-mul(ir, mul(kx, k2)), Z)));
                g = caddc(g, caddc(imul(iz, U), imul(-ix, Wv)));
                au[o] = aw[o] = ax[o] = az[o] = c2(0, 0);
            }
            av[o] = caddc(av[o], v);
            ag[o] = caddc(ag[o], g);
            ad[o] = c2(0, 0);
        }
    }
}

// transpose of k_nonlinear: adjoints of Hv, Hg (slot sn) and of the mean Lamb
// history (slot sn) added into the adjoint Lamb vector lmb (FP64)
template <int TP>
__global__ __launch_bounds__(256) void k_nonlinear_adj(Chan c, StepPtrs sp) {
    __shared__ double2 sh[TP][kMaxNY];
    const int N = c.NY, ns = sp.sn;
    double2 *lx = reinterpret_cast<double2 *>(c.lmb), *ly = lx + c.W, *lz = lx + 2 * c.W;
    double2 *Hv = sp.hvn, *Hg = sp.hgn;
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        sh[pp][j] = p < c.npencil ? Hv[(size_t)p * N + j] : c2(0, 0);
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 t[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) t[pp] = c2(0, 0);
        matvec_T<TP>(c.DY1, sh, N, row, t);  // DY1^T Hv*; the adjoint of t is its negative
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const int k = p / c.NXH, i = p % c.NXH;
            const double kx = c.kx[i], kz = c.kz[k];
            const double k2 = k2_of(c, k, i);
            const size_t o = (size_t)p * N + row;
            const double2 at = c2(-t[pp].x, -t[pp].y), hg = Hg[o];
            double2 ax = caddc(imul(-kx, at), imul(-kz, hg));
            double2 az = caddc(imul(-kz, at), imul(kx, hg));
            if (p == 0) {
                double2 *mx = c.means + (size_t)(M_LMBX * 3 + ns) * N, *mz = c.means + (size_t)(M_LMBZ * 3 + ns) * N;
                ax = caddc(ax, mx[row]);
                az = caddc(az, mz[row]);
                mx[row] = mz[row] = c2(0, 0);
            }
            lx[o] = caddc(lx[o], ax);
            ly[o] = caddc(ly[o], cscale(sh[pp][row], -k2));
            lz[o] = caddc(lz[o], az);
            Hv[o] = Hg[o] = c2(0, 0);
        }
    }
}

// transpose of the friction evaluation: adjoints of U (msol) and W from the
// adjoints of the friction forces fx, fz (real) and of DY1 U, DY1 W (rows)
__device__ void friction_adj(const Chan &c, double2 *sdu, double2 *sdw, double afx, double afz, double2 *aU, double2 *aW) {
    const int N = c.NY;
    const double f = c.nu / c.Ly;
    __syncthreads();
    if (threadIdx.x == 0) {
        // fx = -nu/Ly (DY1U[N-1] - DY1U[0]) on the real parts
        sdu[N - 1].x -= f * afx;
        sdu[0].x += f * afx;
        sdw[N - 1].x -= f * afz;
        sdw[0].x += f * afz;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 u = c2(0, 0), w = c2(0, 0);
        const double *col = c.DY1 + (size_t)row * N;
        for (int j = 0; j < N; ++j) {
            const double a = col[j];
            u.x = fma(a, sdu[j].x, u.x);
            u.y = fma(a, sdu[j].y, u.y);
            w.x = fma(a, sdw[j].x, w.x);
            w.y = fma(a, sdw[j].y, w.y);
        }
        aU[row] = caddc(aU[row], u);
        aW[row] = caddc(aW[row], w);
    }
    __syncthreads();
}

// transpose of k_mean (one block). seed is added to the adjoint of this step's
// fx (the objective's weight on it). G, P: the transposed factor set.
template <int NQ>
__global__ __launch_bounds__(256) void k_mean_adj(Chan c, StepPtrs sp, const double *G, const int *P, int mf, double seed) {
    __shared__ double2 sdu[kMaxNY], sdw[kMaxNY], aU[kMaxNY], aW[kMaxNY];
    __shared__ double S[4][64 * NQ];
    __shared__ double red[2][256];
    const int N = c.NY, lane = threadIdx.x & 63, ns = sp.sn;
    const double dt = c.dt;
    const double afx = c.fric[ns] + seed, afz = c.fric[3 + ns];
    double2 *mvx = c.means + (size_t)(M_VELX * 3 + ns) * N, *mvz = c.means + (size_t)(M_VELZ * 3 + ns) * N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double2 x = c.cur[2 * c.W + j], z = c.cur[3 * c.W + j];
        sdu[j] = c2(-z.x, -z.y);  // omega_z(0,0) = -DY1 U
        sdw[j] = x;               // omega_x(0,0) = DY1 W
        aU[j] = caddc(c.cur[j], mvx[j]);
        aW[j] = caddc(c.cur[c.W + j], mvz[j]);
    }
    friction_adj(c, sdu, sdw, afx, afz, aU, aW);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        for (int f = 0; f < 4; ++f) c.cur[(size_t)f * c.W + j] = c2(0, 0);
        mvx[j] = mvz[j] = c2(0, 0);
    }
    if (threadIdx.x == 0) c.fric[ns] = c.fric[3 + ns] = 0.0;
    for (int j = threadIdx.x; j < 64 * NQ; j += blockDim.x) {
        const bool in = j < N;
        S[0][j] = in ? aU[j].x : 0.0;
        S[1][j] = in ? aU[j].y : 0.0;
        S[2][j] = in ? aW[j].x : 0.0;
        S[3][j] = in ? aW[j].y : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        // A^T y = b with A^T = L' D U' P: substitution in natural order, then y[perm[k]] = w[k]
        const int *Po = P + ((size_t)mf * 2 + 1) * N;
        double t[4][NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r) t[r][q] = S[r][lane + 64 * q];
        lu_substitute<NQ>(G + ((size_t)mf * 2 + 1) * N * N, N, lane, t);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const int k = lane + 64 * q;
            if (k < N) {
                int r = Po[k];
                if ((unsigned)r >= (unsigned)N) r = 0;
                const bool bc = r == 0 || r == N - 1;
                sdu[r] = bc ? c2(0, 0) : c2(t[0][q], t[1][q]);
                sdw[r] = bc ? c2(0, 0) : c2(t[2][q], t[3][q]);
            }
        }
    }
    __syncthreads();
    // mu_U = sdu, mu_W = sdw: into the adjoints of the Lamb, friction and velocity histories
    const double BE[3] = {BE0, BE1, BE2}, AL[3] = {AL0, AL1, AL2};
    double ru = 0.0, rw = 0.0;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double2 mu = sdu[j], mw = sdw[j];
        ru += mu.x;
        rw += mw.x;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            const int sl = sp.s[l];
            double2 *lmx = c.means + (size_t)(M_LMBX * 3 + sl) * N, *lmz = c.means + (size_t)(M_LMBZ * 3 + sl) * N;
            double2 *vx = c.means + (size_t)(M_VELX * 3 + sl) * N, *vz = c.means + (size_t)(M_VELZ * 3 + sl) * N;
            lmx[j] = caddc(lmx[j], cscale(mu, BE[l]));
            lmz[j] = caddc(lmz[j], cscale(mw, BE[l]));
            vx[j] = caddc(vx[j], cdiv(cscale(mu, AL[l]), dt));
            vz[j] = caddc(vz[j], cdiv(cscale(mw, AL[l]), dt));
        }
    }
    red[0][threadIdx.x] = ru;
    red[1][threadIdx.x] = rw;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int t2 = 0; t2 < (int)blockDim.x; ++t2) {
            a += red[0][t2];
            b += red[1][t2];
        }
        for (int l = 0; l < 3; ++l) {
            c.fric[sp.s[l]] += BE[l] * a;
            c.fric[3 + sp.s[l]] += BE[l] * b;
        }
    }
}

// transposes of the initial post-processing (ring 0). Order of use: init_mean,
// init_history, nonlinear (slot 1), Lamb, init_d2v, vor_from_vel.
__global__ __launch_bounds__(256) void k_init_mean_adj(Chan c) {
    __shared__ double2 sdu[kMaxNY], sdw[kMaxNY], aU[kMaxNY], aW[kMaxNY];
    const int N = c.NY;
    double2 *lx = reinterpret_cast<double2 *>(c.lmb), *lz = lx + 2 * c.W;
    double afx = 0.0, afz = 0.0;
    for (int l = 0; l < 3; ++l) {
        afx += c.fric[l];
        afz += c.fric[3 + l];
    }
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        double2 a = c2(0, 0), b = c2(0, 0), u = c2(0, 0), w = c2(0, 0);
        for (int s = 0; s < 3; ++s) {
            a = caddc(a, c.means[(size_t)(M_LMBX * 3 + s) * N + j]);
            b = caddc(b, c.means[(size_t)(M_LMBZ * 3 + s) * N + j]);
            u = caddc(u, c.means[(size_t)(M_VELX * 3 + s) * N + j]);
            w = caddc(w, c.means[(size_t)(M_VELZ * 3 + s) * N + j]);
        }
        lx[j] = caddc(lx[j], a);
        lz[j] = caddc(lz[j], b);
        aU[j] = u;
        aW[j] = w;
        sdu[j] = sdw[j] = c2(0, 0);
    }
    friction_adj(c, sdu, sdw, afx, afz, aU, aW);
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        c.cur[j] = caddc(c.cur[j], aU[j]);
        c.cur[c.W + j] = caddc(c.cur[c.W + j], aW[j]);
        for (int f = 0; f < 12; ++f) c.means[(size_t)f * N + j] = c2(0, 0);
    }
    __syncthreads();
    if (threadIdx.x < 6) c.fric[threadIdx.x] = 0.0;
}

__global__ void k_init_history_adj(Chan c) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < c.W; e += (size_t)gridDim.x * blockDim.x) {
        double2 v = c2(0, 0), g = c2(0, 0), hv = c2(0, 0), hg = c2(0, 0), dv = c2(0, 0);
        for (int s = 0; s < 3; ++s) {
            v = caddc(v, hist(c, H_V, s)[e]);
            g = caddc(g, hist(c, H_G, s)[e]);
            hv = caddc(hv, hist(c, H_HV, s)[e]);
            hg = caddc(hg, hist(c, H_HG, s)[e]);
            dv = caddc(dv, c.d2v[(size_t)s * c.W + e]);
            hist(c, H_V, s)[e] = hist(c, H_G, s)[e] = hist(c, H_HV, s)[e] = hist(c, H_HG, s)[e] = c2(0, 0);
            c.d2v[(size_t)s * c.W + e] = c2(0, 0);
        }
        hist(c, H_V, 0)[e] = v;
        hist(c, H_G, 0)[e] = g;
        hist(c, H_HV, 1)[e] = hv;
        hist(c, H_HG, 1)[e] = hg;
        c.d2v[e] = dv;
    }
}

// transposes of k_init_d2v and k_vor_from_vel: the gradient with respect to the
// initial u, v, w, written to gu, gv, gw (W each)
template <int TP>
__global__ __launch_bounds__(256) void k_init_vel_adj(Chan c, double2 *gu, double2 *gv, double2 *gw) {
    __shared__ double2 sz[TP][kMaxNY], sx[TP][kMaxNY], sd[TP][kMaxNY];
    const int N = c.NY;
    const double2 *au = c.cur, *aw = c.cur + c.W, *ax = c.cur + 2 * c.W, *az = c.cur + 3 * c.W;
    const double2 *av = hist(c, H_V, 0), *ay = hist(c, H_G, 0);
    const int p0 = blockIdx.x * TP;
    for (int e = threadIdx.x; e < TP * N; e += blockDim.x) {
        const int pp = e / N, j = e % N, p = p0 + pp;
        const bool in = p < c.npencil;
        const size_t o = (size_t)p * N + j;
        sz[pp][j] = in ? az[o] : c2(0, 0);
        sx[pp][j] = in ? ax[o] : c2(0, 0);
        sd[pp][j] = in ? c.d2v[o] : c2(0, 0);
    }
    __syncthreads();
    for (int row = threadIdx.x; row < N; row += blockDim.x) {
        double2 dz[TP], dx[TP], dd[TP];
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) dz[pp] = dx[pp] = dd[pp] = c2(0, 0);
        matvec_T<TP>(c.DY1, sz, N, row, dz);
        matvec_T<TP>(c.DY1, sx, N, row, dx);
        matvec_T<TP>(c.DY2, sd, N, row, dd);
    #pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            const int p = p0 + pp;
            if (p >= c.npencil) break;
            const int k = p / c.NXH, i = p % c.NXH;
            const double kx = c.kx[i], kz = c.kz[k];
            const size_t o = (size_t)p * N + row;
            const double2 Y = ay[o], X = sx[pp][row], Z = sz[pp][row];
            // omega_x = DY1 w - i kz v; omega_y = i kz u - i kx w; omega_z = i kx v - DY1 u
            gu[o] = csubc(caddc(au[o], imul(-kz, Y)), dz[pp]);
            gw[o] = caddc(caddc(aw[o], imul(kx, Y)), dx[pp]);
            gv[o] = caddc(caddc(caddc(av[o], dd[pp]), imul(kz, X)), imul(-kx, Z));
        }
    }
}

// dst[f] += src[f] for nf fields of n complex values; zeroes src when asked
struct AccTable {
    double2 *dst[6];
    double2 *src[6];
};
__global__ void k_accumulate(AccTable t, int nf, size_t n) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        for (int f = 0; f < nf; ++f) t.dst[f][e] = caddc(t.dst[f][e], t.src[f][e]);
}

// ---------------------------------------------------------------------------
// Plane statistics of the current level by Parseval over wave space (no
// reference counterpart; the reference prints the friction line only). For each
// y: the means U, V, W (the k = 0 pencil) and the plane averages of u'u', v'v',
// w'w', u'v', u'w', v'w' over every other pencil, weight 2 for kx > 0 (its
// conjugate partner is implicit) and 1 on the kx = 0 plane, the z-Nyquist
// pencils left out as the transforms to physical space leave them out.
// One block per y; out[q * NY + y] for q = 0..8.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_statistics(Chan c, const double2 *v, double *out) {
    __shared__ double red[6][256];
    const int N = c.NY, y = blockIdx.x;
    const double2 *u = c.cur, *w = c.cur + c.W;
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int p = 1 + threadIdx.x; p < c.npencil; p += blockDim.x) {
        const int k = p / c.NXH, i = p % c.NXH;
        if (k == c.NZH) continue;
        const double wt = i == 0 ? 1.0 : 2.0;
        const size_t o = (size_t)p * N + y;
        const double2 a = u[o], b = v[o], d = w[o];
        s[0] += wt * (a.x * a.x + a.y * a.y);
        s[1] += wt * (b.x * b.x + b.y * b.y);
        s[2] += wt * (d.x * d.x + d.y * d.y);
        s[3] += wt * (a.x * b.x + a.y * b.y);
        s[4] += wt * (a.x * d.x + a.y * d.y);
        s[5] += wt * (b.x * d.x + b.y * d.y);
    }
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0 * N + y] = u[y].x;
        out[1 * N + y] = v[y].x;
        out[2 * N + y] = w[y].x;
        for (int q = 0; q < 6; ++q) out[(3 + q) * N + y] = red[q][0];
    }
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kTP = 8;
// the horizontal kernel keeps three accumulators per pencil: fewer pencils per
// block keep its registers low enough for four waves per SIMD (measured faster)
constexpr int kTPH = 4;

static int blocks_tp(const Chan &c) { return (c.npencil + kTP - 1) / kTP; }

// the batched method's derivative stages run on the matrix-core kernels at every
// NY (NY <= kMaxNY); the plain method's in the reference's arithmetic
void launch_horizontal(const Chan &c, const StepPtrs &sp, bool exact, hipStream_t s, float2 *out32) {
    if (exact)
        hipLaunchKernelGGL(k_horizontal_exact<kTPH>, dim3((c.npencil + kTPH - 1) / kTPH), dim3(256), 0, s, c, sp);
    else if (use_mirror(c))
        hipLaunchKernelGGL(k_horizontal_mirror, dim3((c.npencil + kMP - 1) / kMP), dim3(512), 0, s, c, sp, out32);
    else if (c.NY <= 256)
        hipLaunchKernelGGL(k_horizontal_mfma<2>, dim3((c.npencil + kMP - 1) / kMP), dim3(512), 0, s, c, sp);
    else
        hipLaunchKernelGGL(k_horizontal_mfma<3>, dim3((c.npencil + kMP - 1) / kMP), dim3(512), 0, s, c, sp);
}

void launch_helmholtz_g(const Chan &c, const StepPtrs &sp, hipStream_t s) {
    if (use_mirror_g(c)) {
        hipLaunchKernelGGL(k_helmholtz_mirror, dim3((c.npencil + kMP - 1) / kMP), dim3(512), helmholtz_mirror_lds(c.NY), s, c,
                           sp);
        return;
    }
    const int M = c.NY - 2, ntile = (M + 15) / 16;
    const size_t lds = sizeof(double) * 2 * kMP * (16 * ntile + 1);
    const dim3 g((c.npencil + kMP - 1) / kMP);
    if (ntile <= 16) hipLaunchKernelGGL(k_helmholtz_g<2>, g, dim3(512), lds, s, c, sp);
    else hipLaunchKernelGGL(k_helmholtz_g<3>, g, dim3(512), lds, s, c, sp);
}

void launch_mean(const Chan &c, const StepPtrs &sp, const double *Fg, const int *Pg, hipStream_t s) {
    OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_mean<NQ>, dim3(1), dim3(256), 0, s, c, sp, Fg, Pg));
}

// ---- grouped launches of the step's stages over ensemble members ----
static int round8(int n) { return 8 * ((n + 7) / 8); }

int ens_key(const EnsEntry &e, int stage) {
    int blocks = 0;
    const Chan &c = e.c;
    switch (stage) {
        case ES_SOLVE: return ens_key_solve(e, &blocks);
        case ES_HELMHOLTZ: return use_mirror_g(c) ? 4 : (c.NY - 2 + 15) / 16 <= 16 ? 2 : 3;
        case ES_MEAN: return (c.NY + 63) / 64;
        case ES_HORIZONTAL: return use_mirror(c) ? 4 : c.NY <= 256 ? 2 : 3;
        case ES_NONLINEAR: return use_mirror(c) ? 8 : (c.NY <= 256 ? 2 : 3) * 2 + c.precision;
        default: return ens_key_fft(e, stage, &blocks);
    }
}

int ens_blocks(const EnsEntry &e, int stage) {
    int blocks = 0;
    switch (stage) {
        case ES_SOLVE: ens_key_solve(e, &blocks); return blocks;
        case ES_MEAN: return 8;  // one block does the work
        case ES_HELMHOLTZ:
        case ES_HORIZONTAL:
        case ES_NONLINEAR: return round8((e.c.npencil + kMP - 1) / kMP);
        default: ens_key_fft(e, stage, &blocks); return blocks;
    }
}

void launch_ens(int stage, const EnsEntry *ents, const int *first, int nm, int grid, const EnsEntry &rep, int ny_max,
                hipStream_t s) {
    const Chan &c = rep.c;
    const dim3 g(grid);
    switch (stage) {
        case ES_SOLVE: launch_ens_solve(ents, first, nm, grid, rep, s); break;
        case ES_HELMHOLTZ: {
            if (use_mirror_g(c)) {
                hipLaunchKernelGGL(k_helmholtz_mirror_ens, g, dim3(512), helmholtz_mirror_lds(ny_max), s, ents, first, nm);
                break;
            }
            const int ntile = (ny_max - 2 + 15) / 16;
            const size_t lds = sizeof(double) * 2 * kMP * (16 * ntile + 1);
            if ((c.NY - 2 + 15) / 16 <= 16) hipLaunchKernelGGL(k_helmholtz_g_ens<2>, g, dim3(512), lds, s, ents, first, nm);
            else hipLaunchKernelGGL(k_helmholtz_g_ens<3>, g, dim3(512), lds, s, ents, first, nm);
            break;
        }
        case ES_MEAN: OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_mean_ens<NQ>, g, dim3(256), 0, s, ents, first, nm)); break;
        case ES_HORIZONTAL:
            if (use_mirror(c)) hipLaunchKernelGGL(k_horizontal_mirror_ens, g, dim3(512), 0, s, ents, first, nm);
            else if (c.NY <= 256) hipLaunchKernelGGL(k_horizontal_mfma_ens<2>, g, dim3(512), 0, s, ents, first, nm);
            else hipLaunchKernelGGL(k_horizontal_mfma_ens<3>, g, dim3(512), 0, s, ents, first, nm);
            break;
        case ES_NONLINEAR:
            if (use_mirror(c)) {
                hipLaunchKernelGGL(k_nonlinear_mirror_ens<float2>, g, dim3(512), 0, s, ents, first, nm);
            } else if (c.NY <= 256) {
                if (c.precision == 0) hipLaunchKernelGGL((k_nonlinear_mfma_ens<2, double2>), g, dim3(512), 0, s, ents, first, nm);
                else hipLaunchKernelGGL((k_nonlinear_mfma_ens<2, float2>), g, dim3(512), 0, s, ents, first, nm);
            } else {
                if (c.precision == 0) hipLaunchKernelGGL((k_nonlinear_mfma_ens<3, double2>), g, dim3(512), 0, s, ents, first, nm);
                else hipLaunchKernelGGL((k_nonlinear_mfma_ens<3, float2>), g, dim3(512), 0, s, ents, first, nm);
            }
            break;
        default: launch_ens_fft(stage, ents, first, nm, grid, rep, s); break;
    }
}

void launch_nonlinear(const Chan &c, const StepPtrs &sp, bool exact, hipStream_t s) {
    if (exact) {
        if (c.precision == 0)
            hipLaunchKernelGGL((k_nonlinear_exact<kTP, double2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, sp);
        else
            hipLaunchKernelGGL((k_nonlinear_exact<kTP, float2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, sp);
    } else {
        const dim3 g((c.npencil + kMP - 1) / kMP);
        if (use_mirror(c)) {
            hipLaunchKernelGGL(k_nonlinear_mirror<float2>, g, dim3(512), 0, s, c, sp);
        } else if (c.NY <= 256) {
            if (c.precision == 0) hipLaunchKernelGGL((k_nonlinear_mfma<2, double2>), g, dim3(512), 0, s, c, sp);
            else hipLaunchKernelGGL((k_nonlinear_mfma<2, float2>), g, dim3(512), 0, s, c, sp);
        } else {
            if (c.precision == 0) hipLaunchKernelGGL((k_nonlinear_mfma<3, double2>), g, dim3(512), 0, s, c, sp);
            else hipLaunchKernelGGL((k_nonlinear_mfma<3, float2>), g, dim3(512), 0, s, c, sp);
        }
    }
}

void launch_pressure_rhs(const Chan &c, const double2 *dv, const double2 *v, double2 *R, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL((k_pressure_rhs<kTP, double2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, dv, v, R);
    else
        hipLaunchKernelGGL((k_pressure_rhs<kTP, float2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, dv, v, R);
}

// the stored wall-normal terms of the three levels between the two precisions'
// forms (d2v_store): mixed = 1 takes DY2 v to DY2 v - k^2 v, mixed = 0 back
__global__ void k_d2v_convert(Chan c, int mixed) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < 3 * c.W; e += (size_t)gridDim.x * blockDim.x) {
        const int s = (int)(e / c.W);
        const size_t r = e % c.W;
        const int p = (int)(r / c.NY);
        const double k2 = k2_of(c, p / c.NXH, p % c.NXH);
        const double2 kv = cscale(hist(c, H_V, s)[r], k2);
        c.d2v[e] = mixed ? csubc(c.d2v[e], kv) : caddc(c.d2v[e], kv);
    }
}
void launch_d2v_convert(const Chan &c, int mixed, hipStream_t s) {
    hipLaunchKernelGGL(k_d2v_convert, dim3(2048), dim3(256), 0, s, c, mixed);
}

void launch_horizontal_adj(const Chan &c, const StepPtrs &sp, hipStream_t s) {
    hipLaunchKernelGGL(k_horizontal_adj<kTPH>, dim3((c.npencil + kTPH - 1) / kTPH), dim3(256), 0, s, c, sp);
}

void launch_nonlinear_adj(const Chan &c, const StepPtrs &sp, hipStream_t s) {
    hipLaunchKernelGGL(k_nonlinear_adj<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c, sp);
}

void launch_mean_adj(const Chan &c, const StepPtrs &sp, const double *G, const int *P, int mf, double seed, hipStream_t s) {
    OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_mean_adj<NQ>, dim3(1), dim3(256), 0, s, c, sp, G, P, mf, seed));
}

void launch_init_adj_mean_history(const Chan &c, hipStream_t s) {
    hipLaunchKernelGGL(k_init_mean_adj, dim3(1), dim3(256), 0, s, c);
    hipLaunchKernelGGL(k_init_history_adj, dim3(1024), dim3(256), 0, s, c);
}

void launch_init_vel_adj(const Chan &c, double2 *gu, double2 *gv, double2 *gw, hipStream_t s) {
    hipLaunchKernelGGL(k_init_vel_adj<kTPH>, dim3((c.npencil + kTPH - 1) / kTPH), dim3(256), 0, s, c, gu, gv, gw);
}

void launch_accumulate(double2 *const *dst, double2 *const *src, int nf, size_t n, hipStream_t s) {
    AccTable t{};
    for (int f = 0; f < nf; ++f) {
        t.dst[f] = dst[f];
        t.src[f] = src[f];
    }
    hipLaunchKernelGGL(k_accumulate, dim3(1024), dim3(256), 0, s, t, nf, n);
}

void launch_statistics(const Chan &c, const double2 *v, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_statistics, dim3(c.NY), dim3(256), 0, s, c, v, out);
}

void launch_vor_from_vel(const Chan &c, hipStream_t s) {
    hipLaunchKernelGGL(k_vor_from_vel<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c);
}

void launch_init_d2v(const Chan &c, hipStream_t s) {
    hipLaunchKernelGGL(k_init_d2v<kTP>, dim3(blocks_tp(c)), dim3(256), 0, s, c);
}

void launch_init_history(const Chan &c, hipStream_t s) { hipLaunchKernelGGL(k_init_history, dim3(1024), dim3(256), 0, s, c); }

void launch_init_mean(const Chan &c, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL(k_init_mean<double2>, dim3(1), dim3(256), 0, s, c);
    else
        hipLaunchKernelGGL(k_init_mean<float2>, dim3(1), dim3(256), 0, s, c);
}

}  // namespace owt
