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


// ---------------------------------------------------------------------------
// horizontal velocity and vorticity (src/solver.cpp:516-539)
// ---------------------------------------------------------------------------
template <int TP>
__global__ __launch_bounds__(256) void k_horizontal(Chan c, StepPtrs sp) {
    __shared__ double2 sv[TP][kMaxNY];
    __shared__ double2 sg[TP][kMaxNY];
    const int N = c.NY;
    const double2 *v = sp.vn, *g = sp.gn;
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
    // operator columns JC at a time, all loads in flight before use
    for (int j0 = 0; j0 < N; j0 += JC) {
        double da[JC], db[JC];
#pragma unroll
        for (int cc = 0; cc < JC; ++cc) {
            const size_t e = (size_t)(j0 + cc < N ? j0 + cc : N - 1) * N + row;
            da[cc] = c.DY1[e];
            db[cc] = c.DY2[e];
        }
#pragma unroll
        for (int cc = 0; cc < JC; ++cc) {
            const int j = j0 + cc;
            if (j >= N) break;
            const double a = da[cc], b = db[cc];
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
    }
    double2 *u = c.cur, *w = c.cur + c.W, *ox = c.cur + 2 * c.W, *oz = c.cur + 3 * c.W;
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) {
        const int p = p0 + pp;
        if (p >= c.npencil) break;
        const size_t o = (size_t)p * N + row;
        // DY2 v of the new level, kept for the next steps' right-hand sides
        sp.dvn[o] = d2v[pp];
        const int k = p / c.NXH, i = p % c.NXH;
        const double kx = c.kx[i], kz = c.kz[k];
        const double k2 = k2_of(c, k, i);
        if (k2 < 1.e-8) continue;  // the mean pencil is set by the mean-flow kernels
        const double ix = __ddiv_rn(kx, k2), iz = __ddiv_rn(kz, k2), ir = __ddiv_rn(1.0, k2);
        const double2 vv = sv[pp][row], gg = sg[pp][row];
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
// Ag(0) (factor mf, the k = 0 pencil's) in wave 0, then the friction and mean
// vorticity.
template <int NQ>
__global__ __launch_bounds__(256) void k_mean(Chan c, StepPtrs sp, const double *F, const int *P, int mf) {
    __shared__ double S[4][64 * NQ];
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
        lu_gather<NQ>(S, P + ((size_t)mf * 2 + 1) * N, N, lane, t);
        lu_substitute<NQ>(F + ((size_t)mf * 2 + 1) * N * N, N, lane, t);
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

// ---------------------------------------------------------------------------
// nonlinear source terms (src/solver.cpp:240-260); the k = 0 pencil of the
// Lamb vector also enters the mean-flow history (src/solver.cpp:293-303)
// ---------------------------------------------------------------------------
template <int TP, typename LC>
__global__ __launch_bounds__(256) void k_nonlinear(Chan c, StepPtrs sp) {
    __shared__ double2 st[TP][kMaxNY];
    const int N = c.NY, ns = sp.sn;
    const LC *lx = reinterpret_cast<const LC *>(c.lmb);
    const LC *ly = lx + c.W, *lz = lx + 2 * c.W;
    double2 *Hv = sp.hvn, *Hg = sp.hgn;
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
    for (int j0 = 0; j0 < N; j0 += JC) {
        double dj[JC];
#pragma unroll
        for (int cc = 0; cc < JC; ++cc) dj[cc] = c.DY1[(size_t)(j0 + cc < N ? j0 + cc : N - 1) * N + row];
#pragma unroll
        for (int cc = 0; cc < JC; ++cc) {
            const int j = j0 + cc;
            if (j >= N) break;
#pragma unroll
            for (int pp = 0; pp < TP; ++pp) {
                acc[pp].x = fma(dj[cc], st[pp][j].x, acc[pp].x);
                acc[pp].y = fma(dj[cc], st[pp][j].y, acc[pp].y);
            }
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

// ---------------------------------------------------------------------------
// dynamic pressure P = p + |u|^2/2 of the current level (the reference has no
// pressure). The divergence of the rotational momentum equation
// du/dt = L - grad P + nu lap u (L the Lamb vector, u x omega) gives
//   rows 1..N-2:  (DY2 - k^2) P = i kx Lx + DY1 Ly + i kz Lz
// and its y component at the walls, where v = 0:
//   rows 0, N-1:  DY1 P = Ly + nu DY2 v
// The k = 0 pencil's row N-1 is its gauge row (right-hand side 0). This kernel
// writes the right-hand sides; lu_kernels.hip solves.
// ---------------------------------------------------------------------------
template <int TP, typename LC>
__global__ __launch_bounds__(256) void k_pressure_rhs(Chan c, const double2 *dv, double2 *R) {
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
    const int row = threadIdx.x;
    if (row >= N) return;
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
            for (int pp = 0; pp < TP; ++pp) {
                acc[pp].x = fma(dj[cc], st[pp][j].x, acc[pp].x);
                acc[pp].y = fma(dj[cc], st[pp][j].y, acc[pp].y);
            }
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
            r = caddc(st[pp][row], cscale(dv[o], c.nu));
            if (row == N - 1 && k2_of(c, k, i) < 1.e-8) r = c2(0, 0);
        } else {
            r = caddc(caddc(imul(c.kx[i], ld2(lx + o)), imul(c.kz[k], ld2(lz + o))), acc[pp]);
        }
        R[o] = r;
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

// DY2 v of the initial field (slot 0), with the mat-vec of the horizontal kernel
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
    const int row = threadIdx.x;
    if (row >= N) return;
    double2 acc[TP];
#pragma unroll
    for (int pp = 0; pp < TP; ++pp) acc[pp] = c2(0, 0);
    for (int j = 0; j < N; ++j) {
        const double b = c.DY2[(size_t)j * N + row];
#pragma unroll
        for (int pp = 0; pp < TP; ++pp) {
            acc[pp].x = fma(b, sv[pp][j].x, acc[pp].x);
            acc[pp].y = fma(b, sv[pp][j].y, acc[pp].y);
        }
    }
#pragma unroll
    for (int pp = 0; pp < TP; ++pp)
        if (p0 + pp < c.npencil) c.d2v[(size_t)(p0 + pp) * N + row] = acc[pp];
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
// launch helpers
// ---------------------------------------------------------------------------
constexpr int kTP = 8;
// the horizontal kernel keeps three accumulators per pencil: fewer pencils per
// block keep its registers low enough for four waves per SIMD (measured faster)
constexpr int kTPH = 4;

static int blocks_tp(const Chan &c) { return (c.npencil + kTP - 1) / kTP; }

void launch_horizontal(const Chan &c, const StepPtrs &sp, hipStream_t s) {
    hipLaunchKernelGGL(k_horizontal<kTPH>, dim3((c.npencil + kTPH - 1) / kTPH), dim3(256), 0, s, c, sp);
}

void launch_mean(const Chan &c, const StepPtrs &sp, const double *F, const int *P, int mf, hipStream_t s) {
    OWT_NQ_DISPATCH(c.NY, hipLaunchKernelGGL(k_mean<NQ>, dim3(1), dim3(256), 0, s, c, sp, F, P, mf));
}

void launch_nonlinear(const Chan &c, const StepPtrs &sp, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL((k_nonlinear<kTP, double2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, sp);
    else
        hipLaunchKernelGGL((k_nonlinear<kTP, float2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, sp);
}

void launch_pressure_rhs(const Chan &c, const double2 *dv, double2 *R, hipStream_t s) {
    if (c.precision == 0)
        hipLaunchKernelGGL((k_pressure_rhs<kTP, double2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, dv, R);
    else
        hipLaunchKernelGGL((k_pressure_rhs<kTP, float2>), dim3(blocks_tp(c)), dim3(256), 0, s, c, dv, R);
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
