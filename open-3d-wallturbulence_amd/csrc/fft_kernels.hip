// Dealiased horizontal transforms of the nonlinear term (reference
// src/FFT.cpp:49-194 and the Lamb product src/solver.cpp:284-286).
//
// Every kernel works on tiles of TJ consecutive wall-normal points y, which is
// the contiguous axis of all fields, so each global access is a run of TJ
// complex values. A tile of lines is transformed in LDS by a Stockham
// autosort FFT with radix-4/3/2 stages; each thread keeps its butterfly
// inputs in registers between the read and write barriers of a stage.
//
//   z_inverse : wave [NZ][NXH][NY] -> zero-padded, z-Nyquist dropped,
//               backward FFT of length NZ2 -> [NZ2][NXH][NY]
//   x_lamb    : six z-padded fields -> three Hermitian-extended c2r transforms
//               (two real fields per complex transform) -> u x omega in
//               registers -> two forward transforms (lx + i ly, lz) ->
//               kx < NXH kept, / NX2 -> three z-padded fields
//   z_forward : forward FFT of length NZ2, low and high bands kept, / NZ2
// The compute type T is double (plain method) or float (mixed precision).
#include <hip/hip_runtime.h>

#include <climits>

#include <type_traits>

#include "device.hpp"

namespace owt {

template <typename T>
struct CT;
template <>
struct CT<double> {
    using t = double2;
};
template <>
struct CT<float> {
    using t = float2;
};

template <typename C>
__device__ __forceinline__ C mk(decltype(C::x) a, decltype(C::x) b) {
    C r;
    r.x = a;
    r.y = b;
    return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { return mk<C>(a.x + b.x, a.y + b.y); }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { return mk<C>(a.x - b.x, a.y - b.y); }
// The transform arithmetic states every fused multiply-add itself and lets the
// compiler contract nothing, so that every kernel instance (tile shape, length
// plan, single or ensemble launch) rounds the same way and gives the same bits.
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
#pragma clang fp contract(off)
    return mk<C>(fmx(a.x, b.x, -(a.y * b.y)), fmx(a.x, b.y, a.y * b.x));
}

// a b - c d: in FP64 each product and the difference rounded separately, as
// the reference's pointwise Field operations do (src/linalg.hpp:756-816; no
// contraction into a fused multiply-add); in FP32 one fused multiply-add
__device__ __forceinline__ double cross_term(double a, double b, double c, double d) {
#pragma clang fp contract(off)
    return a * b - c * d;
}
__device__ __forceinline__ float cross_term(float a, float b, float c, float d) {
#pragma clang fp contract(off)
    return fmx(a, b, -(c * d));
}

template <typename C, int R>
__device__ __forceinline__ void dft(C (&x)[R], int sign) {
#pragma clang fp contract(off)
    using T = decltype(C::x);
    if constexpr (R == 2) {
        C a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const T sn = (T)sign * (T)0.86602540378443864676372317075294;
        C t = cadd(x[1], x[2]), d = csub(x[1], x[2]);
        C m = mk<C>(fmx((T)-0.5, t.x, x[0].x), fmx((T)-0.5, t.y, x[0].y));
        x[0] = cadd(x[0], t);
        x[1] = mk<C>(fmx(-sn, d.y, m.x), fmx(sn, d.x, m.y));
        x[2] = mk<C>(fmx(sn, d.y, m.x), fmx(-sn, d.x, m.y));
    } else if constexpr (R == 4) {
        C a = cadd(x[0], x[2]), b = csub(x[0], x[2]), c = cadd(x[1], x[3]), d = csub(x[1], x[3]);
        const T s = (T)sign;
        x[0] = cadd(a, c);
        x[2] = csub(a, c);
        x[1] = mk<C>(b.x - s * d.y, b.y + s * d.x);
        x[3] = mk<C>(b.x + s * d.y, b.y - s * d.x);
    } else {
        // radix 5 with w = e^{sign 2 pi i / 5} = c1 + i sign s1, w^2 = c2 + i sign s2:
        //   X1, X4 = x0 + c1 (x1 + x4) + c2 (x2 + x3) +- i sign (s1 (x1 - x4) + s2 (x2 - x3))
        //   X2, X3 = x0 + c2 (x1 + x4) + c1 (x2 + x3) +- i sign (s2 (x1 - x4) - s1 (x2 - x3))
        const T c1 = (T)0.30901699437494742410229341718281906, c2 = (T)-0.80901699437494742410229341718281906;
        const T s1 = (T)sign * (T)0.95105651629515357211643933337938214,
                s2 = (T)sign * (T)0.58778525229247312916870595463907277;
        const C t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]), t3 = csub(x[1], x[4]), t4 = csub(x[2], x[3]);
        const C a1 = mk<C>(fmx(c2, t2.x, fmx(c1, t1.x, x[0].x)), fmx(c2, t2.y, fmx(c1, t1.y, x[0].y)));
        const C a2 = mk<C>(fmx(c1, t2.x, fmx(c2, t1.x, x[0].x)), fmx(c1, t2.y, fmx(c2, t1.y, x[0].y)));
        const C b1 = mk<C>(fmx(s2, t4.x, s1 * t3.x), fmx(s2, t4.y, s1 * t3.y));
        const C b2 = mk<C>(fmx(-s1, t4.x, s2 * t3.x), fmx(-s1, t4.y, s2 * t3.y));
        x[0] = mk<C>(x[0].x + t1.x + t2.x, x[0].y + t1.y + t2.y);
        x[1] = mk<C>(a1.x - b1.y, a1.y + b1.x);
        x[4] = mk<C>(a1.x + b1.y, a1.y - b1.x);
        x[2] = mk<C>(a2.x - b2.y, a2.y + b2.x);
        x[3] = mk<C>(a2.x + b2.y, a2.y - b2.x);
    }
}

// One Stockham stage of radix R over TJ lines of length L held in buf[n*TJ + jj].
// Ns = product of the radices of the earlier stages, mNs its division magic,
// tstep = L / (Ns R). tw[m] = e^{-2 pi i m/L}.
template <typename C, int R, int TJ, int TL, int MAXB>
__device__ __forceinline__ void stage(C *buf, const C *tw, int L, int Ns, unsigned mNs, int tstep, int sign, int jj, int tl) {
    const int Lr = L / R;
    C out[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tl + b * TL;
        if (j < Lr) {
            const int k = j - (int)(mNs ? __umulhi((unsigned)j, mNs) : (unsigned)j) * Ns;
            C x[R];
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = buf[(j + q * Lr) * TJ + jj];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                C w = tw[q * k * tstep];
                if (sign > 0) w.y = -w.y;
                x[q] = cmul(x[q], w);
            }
            dft<C, R>(x, sign);
#pragma unroll
            for (int q = 0; q < R; ++q) out[b][q] = x[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tl + b * TL;
        if (j < Lr) {
            const int k = j - (int)(mNs ? __umulhi((unsigned)j, mNs) : (unsigned)j) * Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[(base + q * Ns) * TJ + jj] = out[b][q];
        }
    }
    __syncthreads();
}

// The same stage with the length, span and twiddle stride known at compile
// time: k = j mod Ns and every LDS address fold into shifts, adds and
// immediate offsets, and rounds that are full need no guard. The arithmetic on
// the data is the runtime stage's, so results are the same bit for bit.
template <typename C, int R, int TJ, int TL, int L, int Ns>
__device__ __forceinline__ void stage_fixed(C *buf, const C *tw, int sign, int jj, int tl) {
    constexpr int Lr = L / R, tstep = L / (Ns * R), MAXB = (Lr + TL - 1) / TL;
    C out[MAXB][R];
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tl + b * TL;
        if ((b + 1) * TL <= Lr || j < Lr) {
            const int k = j % Ns;
            C x[R];
#pragma unroll
            for (int q = 0; q < R; ++q) x[q] = buf[(j + q * Lr) * TJ + jj];
            // the first stage's twiddles are all tw[0] = 1: its products change at
            // most the sign of a zero, and are left out
            if constexpr (Ns > 1) {
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    C w = tw[q * k * tstep];
                    if (sign > 0) w.y = -w.y;
                    x[q] = cmul(x[q], w);
                }
            }
            dft<C, R>(x, sign);
#pragma unroll
            for (int q = 0; q < R; ++q) out[b][q] = x[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; ++b) {
        const int j = tl + b * TL;
        if ((b + 1) * TL <= Lr || j < Lr) {
            const int k = j % Ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) buf[(base + q * Ns) * TJ + jj] = out[b][q];
        }
    }
    __syncthreads();
}

// compile-time plan of a length: make_plan's radix order (4s, 3s, 2s, 5s)
struct FixedPlan {
    int nst = 0;
    int radix[kMaxStages] = {0};
    int ns[kMaxStages] = {0};
};
constexpr FixedPlan fixed_plan(int L) {
    FixedPlan p{};
    int n = L;
    while (n % 4 == 0) p.radix[p.nst++] = 4, n /= 4;
    while (n % 3 == 0) p.radix[p.nst++] = 3, n /= 3;
    while (n % 2 == 0) p.radix[p.nst++] = 2, n /= 2;
    while (n % 5 == 0) p.radix[p.nst++] = 5, n /= 5;
    int ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        p.ns[s] = ns;
        ns *= p.radix[s];
    }
    return p;
}

template <typename C, int TJ, int TL, int L, int S>
__device__ __forceinline__ void fft_fixed(C *buf, const C *tw, int sign, int jj, int tl) {
    constexpr FixedPlan P = fixed_plan(L);
    if constexpr (S < P.nst) {
        stage_fixed<C, P.radix[S], TJ, TL, L, P.ns[S]>(buf, tw, sign, jj, tl);
        fft_fixed<C, TJ, TL, L, S + 1>(buf, tw, sign, jj, tl);
    }
}

// Full transform of the TJ lines in buf (natural order in and out).
// Callers synchronise before the call; the call ends synchronised.
// The lengths of the benchmark and test grids (the largest of each length
// class, and 144, 480 and 960) run the compile-time plan; others the plan P.
template <typename C, int TJ, int TL, int LMAX>
__device__ void fft_lines(C *buf, const C *tw, const FFTPlan &P, int sign, int jj, int tl) {
    if (P.L == LMAX) {
        fft_fixed<C, TJ, TL, LMAX, 0>(buf, tw, sign, jj, tl);
        return;
    }
    if constexpr (LMAX == 192 || LMAX == 576 || LMAX == 1152) {
        constexpr int L2 = LMAX == 192 ? 144 : LMAX == 576 ? 480 : 960;
        if (P.L == L2) {
            fft_fixed<C, TJ, TL, L2, 0>(buf, tw, sign, jj, tl);
            return;
        }
    }
    for (int s = 0; s < P.nst; ++s) {
        const int R = P.radix[s];
        if (R == 4)
            stage<C, 4, TJ, TL, (LMAX / 4 + TL - 1) / TL>(buf, tw, P.L, P.ns[s], P.mns[s], P.tstep[s], sign, jj, tl);
        else if (R == 5)
            stage<C, 5, TJ, TL, (LMAX / 5 + TL - 1) / TL>(buf, tw, P.L, P.ns[s], P.mns[s], P.tstep[s], sign, jj, tl);
        else if (R == 3)
            stage<C, 3, TJ, TL, (LMAX / 3 + TL - 1) / TL>(buf, tw, P.L, P.ns[s], P.mns[s], P.tstep[s], sign, jj, tl);
        else
            stage<C, 2, TJ, TL, (LMAX / 2 + TL - 1) / TL>(buf, tw, P.L, P.ns[s], P.mns[s], P.tstep[s], sign, jj, tl);
    }
}

template <typename C>
__device__ __forceinline__ C load_c(const double2 *p) {
    double2 v = *p;
    return mk<C>((decltype(C::x))v.x, (decltype(C::x))v.y);
}
template <typename C>
__device__ __forceinline__ C load_c(const float2 *p) {
    float2 v = *p;
    return mk<C>((decltype(C::x))v.x, (decltype(C::x))v.y);
}

template <typename C>
__device__ __forceinline__ const C *tw_of(const Chan &c, bool xdir);
template <>
__device__ __forceinline__ const double2 *tw_of<double2>(const Chan &c, bool xdir) { return xdir ? c.twx : c.twz; }
template <>
__device__ __forceinline__ const float2 *tw_of<float2>(const Chan &c, bool xdir) { return xdir ? c.twxf : c.twzf; }

// Tile index of a 1-D grid launched with 8 * ceil(ntiles / 8) blocks. The
// hardware hands consecutive workgroups to the 8 XCDs in turn (block b runs on
// XCD b % 8, each XCD with its own L2), so tiles are numbered to give every XCD
// a contiguous range: the y-tiles of one line set, which share cache lines,
// run side by side in one L2. Padding blocks get -1.
__device__ __forceinline__ int xcd_tile_of(int ntiles, int b) {
    const int per = (ntiles + 7) / 8;
    const int t = (b % 8) * per + b / 8;
    return t < ntiles ? t : -1;
}
__device__ __forceinline__ int xcd_tile(int ntiles) { return xcd_tile_of(ntiles, blockIdx.x); }

static inline int xcd_grid(int ntiles) { return 8 * ((ntiles + 7) / 8); }

// ---------------------------------------------------------------------------
// z inverse (src/FFT.cpp:68-95): per (kx index i, y tile, field)
// ---------------------------------------------------------------------------
// Input fields of the z inverse
struct PtrTable {
    const double2 *p[6];
};

// div divides the input (1 in the step); nyq = 1 also feeds the mode at index
// NZH into the Nyquist slot NZ (0 in the step, which drops it): with div = NZ2
// and nyq = 1 the kernel is the transpose of k_z_forward.
// src_of(f): field f's input, double2 or float2 (the FP32 copies of the step's
// six fields hold the rounding the load would apply)
template <typename T, int TJ, int TL, int LMAX, typename SrcOf>
__device__ __forceinline__ void z_inverse_src(const Chan &c, SrcOf src_of, typename CT<T>::t *out, int nf, T div, int nyq,
                                              const int bid) {
    using C = typename CT<T>::t;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NZ2, NZ = c.NZ, NZH = c.NZH, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile_of(nyt * NXH * nf, bid);
    if (tile < 0) return;
    const int i = (tile / nyt) % NXH, f = tile / (nyt * NXH);
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    const C *twg = tw_of<C>(c, false);
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = twg[n];
    const auto *src = src_of(f);
    const bool scale = div != (T)1;  // x / 1 == x: the step's call skips the divisions
    for (int n = tl; n < L; n += TL) {
        // slot n <- mode n (n < NZH); slots NZH..NZ are padding and the dropped z-Nyquist
        const int mode = n < NZH ? n : (n < NZ || (n == NZ && !nyq) ? -1 : n - NZH);
        C v = mk<C>(0, 0);
        if (mode >= 0 && valid) v = load_c<C>(src + ((size_t)mode * NXH + i) * NY + j);
        buf[n * TJ + jj] = scale ? mk<C>(v.x / div, v.y / div) : v;
    }
    __syncthreads();
    fft_lines<C, TJ, TL, LMAX>(buf, tw, c.pz, +1, jj, tl);
    C *dst = out + (size_t)f * c.I;
    if (valid)
        for (int n = tl; n < L; n += TL) dst[((size_t)n * NXH + i) * NY + j] = buf[n * TJ + jj];
}
template <typename T, int TJ, int TL, int LMAX>
__device__ __forceinline__ void z_inverse_body(const Chan &c, const double2 *const *in, typename CT<T>::t *out, int nf, T div, int nyq,
                                               const int bid) {
    z_inverse_src<T, TJ, TL, LMAX>(c, [in](int f) { return in[f]; }, out, nf, div, nyq, bid);
}
template <typename T, int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_z_inverse(Chan c, PtrTable in, typename CT<T>::t *out, int nf, T div, int nyq) {
    z_inverse_body<T, TJ, TL, LMAX>(c, in.p, out, nf, div, nyq, blockIdx.x);
}
// the step's z inverse in mixed precision from the six fields' FP32 copies
// (6 W float2, field f at six + f W)
template <int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_z_inverse_f32in(Chan c, const float2 *six, float2 *out) {
    const size_t W = c.W;
    z_inverse_src<float, TJ, TL, LMAX>(c, [six, W](int f) { return six + (size_t)f * W; }, out, 6, 1.0f, 0, blockIdx.x);
}
template <typename T, int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_z_inverse_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    z_inverse_body<T, TJ, TL, LMAX>(e.c, e.six, (typename CT<T>::t *)e.c.zbuf, 6, (T)1, 0, bid);
}

// ---------------------------------------------------------------------------
// z forward (src/FFT.cpp:171-190): per (i, y tile, field)
// ---------------------------------------------------------------------------
// div divides the output (NZ2 in the step); nyq = 0 sets the mode at index NZH
// to zero (1 in the step): with div = 1 and nyq = 0 the kernel is the transpose
// of k_z_inverse.
template <typename T, int TJ, int TL, int LMAX, typename OutC>
__device__ __forceinline__ void z_forward_body(const Chan &c, const typename CT<T>::t *in, OutC *out, int nf, T div, int nyq,
                                               const int bid) {
    using C = typename CT<T>::t;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NZ2, NZ = c.NZ, NZH = c.NZH, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile_of(nyt * NXH * nf, bid);
    if (tile < 0) return;
    const int i = (tile / nyt) % NXH, f = tile / (nyt * NXH);
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    const C *twg = tw_of<C>(c, false);
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = twg[n];
    const C *src = in + (size_t)f * c.I;
    for (int n = tl; n < L; n += TL) buf[n * TJ + jj] = valid ? src[((size_t)n * NXH + i) * NY + j] : mk<C>(0, 0);
    __syncthreads();
    fft_lines<C, TJ, TL, LMAX>(buf, tw, c.pz, -1, jj, tl);
    OutC *dst = out + (size_t)f * c.W;
    const T rdiv = (T)1 / div;
    if (valid)
        for (int k = tl; k < NZ; k += TL) {
            // mode k <- bin k (k < NZH); mode k >= NZH <- bin k + NZH (negative frequencies)
            const int bin = k < NZH ? k : k + NZH;
            C v = buf[bin * TJ + jj];
            if (k == NZH && !nyq) v = mk<C>(0, 0);
            OutC o;
            if constexpr (std::is_same<T, float>::value) {
                // FP32 (the mixed transform phase): the product with the rounded
                // reciprocal of div, within one FP32 rounding of the quotient
                o.x = v.x * rdiv;
                o.y = v.y * rdiv;
            } else {
                o.x = v.x / div;
                o.y = v.y / div;
            }
            dst[((size_t)k * NXH + i) * NY + j] = o;
        }
}
template <typename T, int TJ, int TL, int LMAX, typename OutC>
__global__ __launch_bounds__(TJ *TL) void k_z_forward(Chan c, const typename CT<T>::t *in, OutC *out, int nf, T div, int nyq) {
    z_forward_body<T, TJ, TL, LMAX, OutC>(c, in, out, nf, div, nyq, blockIdx.x);
}
template <typename T, int TJ, int TL, int LMAX, typename OutC>
__global__ __launch_bounds__(TJ *TL) void k_z_forward_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    z_forward_body<T, TJ, TL, LMAX, OutC>(e.c, (const typename CT<T>::t *)e.c.xbuf, (OutC *)e.c.lmb, 3, (T)e.c.NZ2, 1, bid);
}

// ---------------------------------------------------------------------------
// Fused x phase (src/FFT.cpp:101-127, src/solver.cpp:284-286, src/FFT.cpp:147-164)
// per (z plane k2, y tile). Each thread owns the x positions n = tl + e*TL.
// ---------------------------------------------------------------------------
// The spectra of a pair of fields, fetched into registers (modes m = tl + e TL)
// so that the loads of the next pair are in flight during the current transform.
template <typename C, int MP>
struct PairRegs {
    C a[MP], b[MP];
};

template <typename C, int MP>
__device__ __forceinline__ void fetch_pair_spectrum(PairRegs<C, MP> &r, const C *A, const C *B, int NXH, size_t stride_i,
                                                    bool valid, int tl, int TL) {
#pragma unroll
    for (int e = 0; e < MP; ++e) {
        const int m = tl + e * TL;
        const bool in = valid && m < NXH;
        const size_t o = (size_t)(in ? m : 0) * stride_i;
        r.a[e] = in ? A[o] : mk<C>(0, 0);
        r.b[e] = in ? B[o] : mk<C>(0, 0);
    }
}

// Z = Aext + i Bext with Hermitian extensions; the kx = 0 entries enter
// through their real parts (the reference keeps Re of the backward transform)
template <typename C, int MP>
__device__ __forceinline__ void place_pair_spectrum(C *buf, const PairRegs<C, MP> &r, int L, int NXH, int jj, int tl, int TJ,
                                                    int TL) {
#pragma unroll
    for (int e = 0; e < MP; ++e) {
        const int m = tl + e * TL;
        if (m >= NXH) break;
        const C a = r.a[e], b = r.b[e];
        if (m == 0) {
            buf[jj] = mk<C>(a.x, b.x);
        } else {
            buf[m * TJ + jj] = mk<C>(a.x - b.y, a.y + b.x);
            buf[(L - m) * TJ + jj] = mk<C>(a.x + b.y, b.x - a.y);
        }
    }
    for (int n = NXH + tl; n <= L - NXH; n += TL) buf[n * TJ + jj] = mk<C>(0, 0);
}

template <typename T, int TJ, int TL, int LMAX>
__device__ __forceinline__ void x_lamb_body(const Chan &c, const typename CT<T>::t *in, typename CT<T>::t *out, const int bid) {
    using C = typename CT<T>::t;
    constexpr int EPT = (LMAX + TL - 1) / TL;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NX2, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile_of(nyt * c.NZ2, bid);
    if (tile < 0) return;
    const int k2 = tile / nyt;
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    const C *twg = tw_of<C>(c, true);
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = twg[n];
    const size_t plane = (size_t)k2 * NXH * NY + j;
    // modes per thread of one spectrum: NXH <= NX2 / 3 + 1 <= LMAX / 3 + 1
    constexpr int MP = (LMAX / 3 + 1 + TL - 1) / TL;
    PairRegs<C, MP> pr;
    fetch_pair_spectrum<C, MP>(pr, in + plane, in + c.I + plane, NXH, (size_t)NY, valid, tl, TL);
    C a[3][EPT];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        place_pair_spectrum<C, MP>(buf, pr, L, NXH, jj, tl, TJ, TL);
        __syncthreads();
        if (q < 2)
            fetch_pair_spectrum<C, MP>(pr, in + (size_t)(2 * q + 2) * c.I + plane, in + (size_t)(2 * q + 3) * c.I + plane, NXH,
                                       (size_t)NY, valid, tl, TL);
        fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, +1, jj, tl);
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int n = tl + e * TL;
            if (n < L) a[q][e] = buf[n * TJ + jj];
        }
        __syncthreads();
    }
    // Lamb vector u x omega (src/solver.cpp:284-286):
    // pair 0 = (u, v), pair 1 = (w, omx), pair 2 = (omy, omz)
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const T u = a[0][e].x, v = a[0][e].y, w = a[1][e].x, ox = a[1][e].y, oy = a[2][e].x, oz = a[2][e].y;
        const T lx = cross_term(v, oz, w, oy);
        const T ly = cross_term(w, ox, u, oz);
        const T lz = cross_term(u, oy, v, ox);
        a[0][e] = mk<C>(lx, ly);
        a[1][e] = mk<C>(lz, (T)0);
    }
    const T sc = (T)L;
    // the forward transform's 1/NX2 (src/FFT.cpp:161): divided in FP64; in FP32
    // (the mixed transform phase) a product with the rounded reciprocal, within
    // one FP32 rounding of the quotient and an eighth of the instructions
    constexpr bool DIV = !std::is_same<T, float>::value;
    const T rs = (T)1 / sc, hrs = (T)0.5 / sc;
    // FP32 with an even number of lines: the real z component of lines 2m and
    // 2m+1 goes through one complex transform (lz_2m + i lz_2m+1) of the TJ/2
    // line pairs, all threads taking part, and the two spectra are separated
    // like pass 0's; half the butterflies of a transform per line
    constexpr bool PAIR_LZ = !DIV && TJ % 2 == 0;
    constexpr int NPASS = PAIR_LZ ? 1 : 2;
#pragma unroll
    for (int pass = 0; pass < NPASS; ++pass) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int n = tl + e * TL;
            if (n < L) buf[n * TJ + jj] = a[pass][e];
        }
        __syncthreads();
        fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, -1, jj, tl);
        if (valid) {
            for (int n = tl; n < NXH; n += TL) {
                const C zn = buf[n * TJ + jj];
                if (pass == 0) {
                    const C zm = buf[((L - n) % L) * TJ + jj];
                    // spectra of the two real sequences packed as lx + i ly
                    const T ax = zn.x + zm.x, ay = zn.y - zm.y, bx = zn.y + zm.y, by = zm.x - zn.x;
                    C X, Y;
                    if constexpr (DIV) {
                        X = mk<C>(ax * (T)0.5 / sc, ay * (T)0.5 / sc);
                        Y = mk<C>(bx * (T)0.5 / sc, by * (T)0.5 / sc);
                    } else {
                        X = mk<C>(ax * hrs, ay * hrs);
                        Y = mk<C>(bx * hrs, by * hrs);
                    }
                    out[0 * c.I + plane + (size_t)n * NY] = X;
                    out[1 * c.I + plane + (size_t)n * NY] = Y;
                } else {
                    out[2 * c.I + plane + (size_t)n * NY] = DIV ? mk<C>(zn.x / sc, zn.y / sc) : mk<C>(zn.x * rs, zn.y * rs);
                }
            }
        }
        __syncthreads();
    }
    if constexpr (PAIR_LZ) {
        constexpr int TJ2 = TJ / 2, TL2 = 2 * TL;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int n = tl + e * TL;
            if (n < L) reinterpret_cast<T *>(&buf[n * TJ2 + jj / 2])[jj & 1] = a[1][e].x;
        }
        __syncthreads();
        const int jp = threadIdx.x % TJ2, tp = threadIdx.x / TJ2;
        fft_lines<C, TJ2, TL2, LMAX>(buf, tw, c.px, -1, jp, tp);
        const int j0 = (tile % nyt) * TJ + 2 * jp;
        const size_t p0 = (size_t)k2 * NXH * NY + j0;
        for (int n = tp; n < NXH; n += TL2) {
            const C zn = buf[n * TJ2 + jp], zm = buf[((L - n) % L) * TJ2 + jp];
            const T ax = zn.x + zm.x, ay = zn.y - zm.y, bx = zn.y + zm.y, by = zm.x - zn.x;
            if (j0 < NY) out[2 * c.I + p0 + (size_t)n * NY] = mk<C>(ax * hrs, ay * hrs);
            if (j0 + 1 < NY) out[2 * c.I + p0 + 1 + (size_t)n * NY] = mk<C>(bx * hrs, by * hrs);
        }
    }
}
template <typename T, int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_lamb(Chan c, const typename CT<T>::t *in, typename CT<T>::t *out) {
    x_lamb_body<T, TJ, TL, LMAX>(c, in, out, blockIdx.x);
}
template <typename T, int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_lamb_ens(const EnsEntry *__restrict__ ents, const int *__restrict__ first, int nm) {
    const int m = ens_member(first, nm, blockIdx.x);
    const EnsEntry &e = ents[m];
    const int bid = blockIdx.x - first[m];
    x_lamb_body<T, TJ, TL, LMAX>(e.c, (const typename CT<T>::t *)e.c.zbuf, (typename CT<T>::t *)e.c.xbuf, bid);
}

// ---------------------------------------------------------------------------
// Linearized and adjoint x phase (FP64), for the tangent and adjoint steps.
// base holds the six z-transformed fields of the state linearized about (u, v,
// w, omx, omy, omz), in the layout of the step's z buffer.
//
// k_x_tangent: tangent six fields (same layout) -> dL = du x omega + u x domega,
//   transformed as the Lamb vector (three fields, the step's x buffer layout).
// k_x_adjoint: the transpose. lam holds the adjoint of the three x-phase
//   outputs; with lam_phys = (its transpose of the forward x transform) the
//   adjoints of the six inputs are omega x lam_phys (for u) and lam_phys x u
//   (for omega), taken back through the transpose of the c2r transform.
//   Transposes, with theta = 2 pi / NX2 and modes m < NXH:
//     forward  X_m = (1/NX2) sum_j x_j e^{-i theta m j}
//       -> x*_j = (1/NX2) (Re a_0 + Re sum_{m>=1} a_m e^{i theta m j})
//          = the c2r transform of (a_0, a_m / 2) divided by NX2
//     c2r      x_j = Re c_0 + 2 Re sum_{m>=1} c_m e^{i theta m j}
//       -> c*_0 = sum_j x*_j,  c*_m = 2 sum_j x*_j e^{-i theta m j}
//          = the unnormalised forward transform, modes m >= 1 doubled
// ---------------------------------------------------------------------------
template <typename C>
__device__ __forceinline__ void load_pair_scaled(C *buf, const C *A, const C *B, int L, int NXH, size_t stride_i, bool valid,
                                                 int jj, int tl, int TL, int TJ, double s0, double sm) {
    for (int m = tl; m < NXH; m += TL) {
        C a = mk<C>(0, 0), b = mk<C>(0, 0);
        if (valid) {
            a = A[(size_t)m * stride_i];
            if (B) b = B[(size_t)m * stride_i];
        }
        const double sc = m == 0 ? s0 : sm;
        a = mk<C>(a.x * sc, a.y * sc);
        b = mk<C>(b.x * sc, b.y * sc);
        if (m == 0) {
            buf[jj] = mk<C>(a.x, b.x);
        } else {
            buf[m * TJ + jj] = mk<C>(a.x - b.y, a.y + b.x);
            buf[(L - m) * TJ + jj] = mk<C>(a.x + b.y, b.x - a.y);
        }
    }
    for (int n = NXH + tl; n <= L - NXH; n += TL) buf[n * TJ + jj] = mk<C>(0, 0);
}

// c2r of field pairs (f0, f1) of src into physical values held per thread
template <int TJ, int TL, int LMAX, int EPT>
__device__ __forceinline__ void x_c2r_pair(double2 *buf, const double2 *tw, const Chan &c, const double2 *A, const double2 *B,
                                           bool valid, int jj, int tl, double s0, double sm, double2 (&out)[EPT]) {
    load_pair_scaled<double2>(buf, A, B, c.NX2, c.NXH, (size_t)c.NY, valid, jj, tl, TL, TJ, s0, sm);
    __syncthreads();
    fft_lines<double2, TJ, TL, LMAX>(buf, tw, c.px, +1, jj, tl);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int n = tl + e * TL;
        if (n < c.NX2) out[e] = buf[n * TJ + jj];
    }
    __syncthreads();
}

template <int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_tangent(Chan c, const double2 *base, const double2 *in, double2 *out) {
    using C = double2;
    constexpr int EPT = (LMAX + TL - 1) / TL;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NX2, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile(nyt * c.NZ2);
    if (tile < 0) return;
    const int k2 = tile / nyt;
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = c.twx[n];
    const size_t plane = (size_t)k2 * NXH * NY + j;
    // base pairs (u, v), (w, omx), (omy, omz)
    C b[3][EPT];
#pragma unroll
    for (int q = 0; q < 3; ++q)
        x_c2r_pair<TJ, TL, LMAX, EPT>(buf, tw, c, base + (size_t)(2 * q) * c.I + plane, base + (size_t)(2 * q + 1) * c.I + plane,
                                      valid, jj, tl, 1.0, 1.0, b[q]);
    double lx[EPT], ly[EPT], lz[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) lx[e] = ly[e] = lz[e] = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        C t[EPT];
        x_c2r_pair<TJ, TL, LMAX, EPT>(buf, tw, c, in + (size_t)(2 * q) * c.I + plane, in + (size_t)(2 * q + 1) * c.I + plane, valid,
                                      jj, tl, 1.0, 1.0, t);
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const double u = b[0][e].x, v = b[0][e].y, w = b[1][e].x, ox = b[1][e].y, oy = b[2][e].x, oz = b[2][e].y;
            if (q == 0) {  // du, dv
                const double du = t[e].x, dv = t[e].y;
                lx[e] += dv * oz;
                ly[e] -= du * oz;
                lz[e] += du * oy - dv * ox;
            } else if (q == 1) {  // dw, domx
                const double dw = t[e].x, dox = t[e].y;
                lx[e] -= dw * oy;
                ly[e] += dw * ox + w * dox;
                lz[e] -= v * dox;
            } else {  // domy, domz
                const double doy = t[e].x, doz = t[e].y;
                lx[e] += v * doz - w * doy;
                ly[e] -= u * doz;
                lz[e] += u * doy;
            }
        }
    }
    const double sc = (double)L;
#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int n = tl + e * TL;
            if (n < L) buf[n * TJ + jj] = pass == 0 ? mk<C>(lx[e], ly[e]) : mk<C>(lz[e], 0.0);
        }
        __syncthreads();
        fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, -1, jj, tl);
        if (valid) {
            for (int n = tl; n < NXH; n += TL) {
                const C zn = buf[n * TJ + jj];
                if (pass == 0) {
                    const C zm = buf[((L - n) % L) * TJ + jj];
                    out[0 * c.I + plane + (size_t)n * NY] = mk<C>((zn.x + zm.x) * 0.5 / sc, (zn.y - zm.y) * 0.5 / sc);
                    out[1 * c.I + plane + (size_t)n * NY] = mk<C>((zn.y + zm.y) * 0.5 / sc, (zm.x - zn.x) * 0.5 / sc);
                } else {
                    out[2 * c.I + plane + (size_t)n * NY] = mk<C>(zn.x / sc, zn.y / sc);
                }
            }
        }
        __syncthreads();
    }
}

template <int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_adjoint(Chan c, const double2 *base, const double2 *lam, double2 *out) {
    using C = double2;
    constexpr int EPT = (LMAX + TL - 1) / TL;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NX2, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile(nyt * c.NZ2);
    if (tile < 0) return;
    const int k2 = tile / nyt;
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = c.twx[n];
    const size_t plane = (size_t)k2 * NXH * NY + j;
    C b[3][EPT], lp[2][EPT];
#pragma unroll
    for (int q = 0; q < 3; ++q)
        x_c2r_pair<TJ, TL, LMAX, EPT>(buf, tw, c, base + (size_t)(2 * q) * c.I + plane, base + (size_t)(2 * q + 1) * c.I + plane,
                                      valid, jj, tl, 1.0, 1.0, b[q]);
    const double il = 1.0 / (double)L;
    // (lam_x, lam_y) and (lam_z, 0): the transpose of the forward x transform
    x_c2r_pair<TJ, TL, LMAX, EPT>(buf, tw, c, lam + plane, lam + (size_t)c.I + plane, valid, jj, tl, il, 0.5 * il, lp[0]);
    x_c2r_pair<TJ, TL, LMAX, EPT>(buf, tw, c, lam + (size_t)2 * c.I + plane, nullptr, valid, jj, tl, il, 0.5 * il, lp[1]);
    // output pairs (u*, v*), (w*, omx*), (omy*, omz*)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            const int n = tl + e * TL;
            const double u = b[0][e].x, v = b[0][e].y, w = b[1][e].x, ox = b[1][e].y, oy = b[2][e].x, oz = b[2][e].y;
            const double ax = lp[0][e].x, ay = lp[0][e].y, az = lp[1][e].x;
            C r;
            if (q == 0) r = mk<C>(oy * az - oz * ay, oz * ax - ox * az);
            else if (q == 1) r = mk<C>(ox * ay - oy * ax, ay * w - az * v);
            else r = mk<C>(az * u - ax * w, ax * v - ay * u);
            if (n < L) buf[n * TJ + jj] = r;
        }
        __syncthreads();
        fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, -1, jj, tl);
        if (valid) {
            for (int n = tl; n < NXH; n += TL) {
                const C zn = buf[n * TJ + jj], zm = buf[((L - n) % L) * TJ + jj];
                // spectra of the two real sequences, modes m >= 1 doubled
                const double f = n == 0 ? 0.5 : 1.0;
                out[(size_t)(2 * q) * c.I + plane + (size_t)n * NY] = mk<C>((zn.x + zm.x) * f, (zn.y - zm.y) * f);
                out[(size_t)(2 * q + 1) * c.I + plane + (size_t)n * NY] = mk<C>((zn.y + zm.y) * f, (zm.x - zn.x) * f);
            }
        }
        __syncthreads();
    }
}

// physical-space utilities (FP64): c2r of nf z-padded fields -> [NZ2][NX2][NY]
template <int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_c2r(Chan c, const double2 *in, int nf, double *phys) {
    using C = double2;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NX2, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile(nyt * c.NZ2 * ((nf + 1) / 2));
    if (tile < 0) return;
    const int k2 = (tile / nyt) % c.NZ2, pr = tile / (nyt * c.NZ2);
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = c.twx[n];
    const size_t plane = (size_t)k2 * NXH * NY + j;
    const int f0 = 2 * pr, f1 = 2 * pr + 1;
    const bool has2 = f1 < nf;
    // the second field of an odd count is taken as zero: reuse the first pointer and zero it
    const C *A = in + (size_t)f0 * c.I + plane;
    const C *B = in + (size_t)(has2 ? f1 : f0) * c.I + plane;
    for (int m = tl; m < NXH; m += TL) {
        C a = mk<C>(0, 0), b = mk<C>(0, 0);
        if (valid) {
            a = A[(size_t)m * NY];
            if (has2) b = B[(size_t)m * NY];
        }
        if (m == 0) buf[jj] = mk<C>(a.x, b.x);
        else {
            buf[m * TJ + jj] = mk<C>(a.x - b.y, a.y + b.x);
            buf[(L - m) * TJ + jj] = mk<C>(a.x + b.y, b.x - a.y);
        }
    }
    for (int n = NXH + tl; n <= L - NXH; n += TL) buf[n * TJ + jj] = mk<C>(0, 0);
    __syncthreads();
    fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, +1, jj, tl);
    const size_t P = (size_t)c.NZ2 * L * NY;
    if (valid)
        for (int n = tl; n < L; n += TL) {
            const C z = buf[n * TJ + jj];
            const size_t o = ((size_t)k2 * L + n) * NY + j;
            phys[(size_t)f0 * P + o] = z.x;
            if (has2) phys[(size_t)f1 * P + o] = z.y;
        }
}

// r2c of nf physical fields -> nf z-padded fields (kx < NXH, / NX2)
template <int TJ, int TL, int LMAX>
__global__ __launch_bounds__(TJ *TL) void k_x_r2c(Chan c, const double *phys, int nf, double2 *out) {
    using C = double2;
    extern __shared__ unsigned char smem[];
    C *tw = reinterpret_cast<C *>(smem);
    C *buf = tw + LMAX;
    const int L = c.NX2, NXH = c.NXH, NY = c.NY;
    const int nyt = (NY + TJ - 1) / TJ;
    const int tile = xcd_tile(nyt * c.NZ2 * ((nf + 1) / 2));
    if (tile < 0) return;
    const int k2 = (tile / nyt) % c.NZ2, pr = tile / (nyt * c.NZ2);
    const int jj = threadIdx.x % TJ, tl = threadIdx.x / TJ;
    const int j = (tile % nyt) * TJ + jj;
    const bool valid = j < NY;
    for (int n = threadIdx.x; n < L; n += TJ * TL) tw[n] = c.twx[n];
    const int f0 = 2 * pr, f1 = 2 * pr + 1;
    const bool has2 = f1 < nf;
    const size_t P = (size_t)c.NZ2 * L * NY;
    for (int n = tl; n < L; n += TL) {
        double x = 0, y = 0;
        if (valid) {
            const size_t o = ((size_t)k2 * L + n) * NY + j;
            x = phys[(size_t)f0 * P + o];
            if (has2) y = phys[(size_t)f1 * P + o];
        }
        buf[n * TJ + jj] = mk<C>(x, y);
    }
    __syncthreads();
    fft_lines<C, TJ, TL, LMAX>(buf, tw, c.px, -1, jj, tl);
    const size_t plane = (size_t)k2 * NXH * NY + j;
    const double sc = (double)L;
    if (valid)
        for (int n = tl; n < NXH; n += TL) {
            const C zn = buf[n * TJ + jj], zm = buf[((L - n) % L) * TJ + jj];
            if (has2) {
                out[(size_t)f0 * c.I + plane + (size_t)n * NY] = mk<C>((zn.x + zm.x) * 0.5 / sc, (zn.y - zm.y) * 0.5 / sc);
                out[(size_t)f1 * c.I + plane + (size_t)n * NY] = mk<C>((zn.y + zm.y) * 0.5 / sc, (zm.x - zn.x) * 0.5 / sc);
            } else {
                out[(size_t)f0 * c.I + plane + (size_t)n * NY] = mk<C>(zn.x / sc, zn.y / sc);
            }
        }
}

__global__ void k_peak3(const double *p, size_t n, unsigned long long *out) {
    double m = 0.0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const double u = p[e], v = p[n + e], w = p[2 * n + e];
        const double s = u * u + v * v + w * w;
        m = s > m ? s : m;
    }
    // non-negative doubles order like their bit patterns
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// ---------------------------------------------------------------------------
// launch helpers: tile shape chosen by the padded length
// ---------------------------------------------------------------------------
namespace {

template <typename T, int TJ, int TL, int LMAX>
size_t smem_bytes() {
    return sizeof(typename CT<T>::t) * (size_t)LMAX * (TJ + 1);
}

// Tile shapes: TJ consecutive y-points per block (each global access is a run
// of TJ complex values), TL threads per line. Wider runs (8 or 16 points, for
// 64- or 128-byte runs) cost LDS per block, and with it blocks per CU; at the
// lengths of the Re_tau 550 grid they measured slower than these.
#define OWT_Z_DISPATCH(L, T, BODY)                                                   \
    do {                                                                             \
        if ((L) <= 96) { constexpr int TJ = 16, TL = 16, LMAX = 96; BODY; }          \
        else if ((L) <= 192) { constexpr int TJ = 8, TL = 32, LMAX = 192; BODY; }    \
        else if ((L) <= 576) { constexpr int TJ = 4, TL = 64, LMAX = 576; BODY; }    \
        else { constexpr int TJ = 2, TL = 128, LMAX = 1152; BODY; }                  \
    } while (0)
// Tuned variants for the step's three transform kernels: candidate SEL of the
// length class (0 the shape above; every candidate has 256 threads and fits 64 KB
// of LDS in FP64)
#define OWT_TUNED_DISPATCH(L, SEL, BODY)                                                            \
    do {                                                                                            \
        const int sel_ = (SEL);                                                                     \
        if ((L) <= 96) {                                                                            \
            if (sel_ == 1) { constexpr int TJ = 8, TL = 32, LMAX = 96; BODY; }                      \
            else if (sel_ == 2) { constexpr int TJ = 32, TL = 8, LMAX = 96; BODY; }                 \
            else { constexpr int TJ = 16, TL = 16, LMAX = 96; BODY; }                               \
        } else if ((L) <= 192) {                                                                    \
            if (sel_ == 1) { constexpr int TJ = 16, TL = 16, LMAX = 192; BODY; }                    \
            else if (sel_ == 2) { constexpr int TJ = 4, TL = 64, LMAX = 192; BODY; }                \
            else { constexpr int TJ = 8, TL = 32, LMAX = 192; BODY; }                               \
        } else if ((L) <= 576) {                                                                    \
            if (sel_ == 1) { constexpr int TJ = 2, TL = 128, LMAX = 576; BODY; }                    \
            else if (sel_ == 2) { constexpr int TJ = 1, TL = 256, LMAX = 576; BODY; }               \
            else { constexpr int TJ = 4, TL = 64, LMAX = 576; BODY; }                               \
        } else {                                                                                    \
            if (sel_ == 1) { constexpr int TJ = 1, TL = 256, LMAX = 1152; BODY; }                   \
            else { constexpr int TJ = 2, TL = 128, LMAX = 1152; BODY; }                             \
        }                                                                                           \
    } while (0)
// FP32 (mixed precision) variant: candidate 3 doubles the y-run of the default
// shape of the two long length classes, which only fits 64 KB of LDS in FP32
#define OWT_TUNED_DISPATCH_F32(L, SEL, BODY)                                                        \
    do {                                                                                            \
        const int sel_ = (SEL);                                                                     \
        if ((L) <= 96) {                                                                            \
            if (sel_ == 1) { constexpr int TJ = 8, TL = 32, LMAX = 96; BODY; }                      \
            else if (sel_ == 2) { constexpr int TJ = 32, TL = 8, LMAX = 96; BODY; }                 \
            else { constexpr int TJ = 16, TL = 16, LMAX = 96; BODY; }                               \
        } else if ((L) <= 192) {                                                                    \
            if (sel_ == 1) { constexpr int TJ = 16, TL = 16, LMAX = 192; BODY; }                    \
            else if (sel_ == 2) { constexpr int TJ = 4, TL = 64, LMAX = 192; BODY; }                \
            else { constexpr int TJ = 8, TL = 32, LMAX = 192; BODY; }                               \
        } else if ((L) <= 576) {                                                                    \
            if (sel_ == 1) { constexpr int TJ = 2, TL = 128, LMAX = 576; BODY; }                    \
            else if (sel_ == 2) { constexpr int TJ = 1, TL = 256, LMAX = 576; BODY; }               \
            else if (sel_ == 3) { constexpr int TJ = 8, TL = 32, LMAX = 576; BODY; }                \
            else { constexpr int TJ = 4, TL = 64, LMAX = 576; BODY; }                               \
        } else {                                                                                    \
            if (sel_ == 1) { constexpr int TJ = 1, TL = 256, LMAX = 1152; BODY; }                   \
            else if (sel_ == 3) { constexpr int TJ = 4, TL = 64, LMAX = 1152; BODY; }               \
            else { constexpr int TJ = 2, TL = 128, LMAX = 1152; BODY; }                             \
        }                                                                                           \
    } while (0)
// the same shapes for the x phase (spelled out: a forwarding macro would expand
// the launch in BODY before substitution)
#define OWT_X_DISPATCH(L, T, BODY)                                                   \
    do {                                                                             \
        if ((L) <= 96) { constexpr int TJ = 16, TL = 16, LMAX = 96; BODY; }          \
        else if ((L) <= 192) { constexpr int TJ = 8, TL = 32, LMAX = 192; BODY; }    \
        else if ((L) <= 576) { constexpr int TJ = 4, TL = 64, LMAX = 576; BODY; }    \
        else { constexpr int TJ = 2, TL = 128, LMAX = 1152; BODY; }                  \
    } while (0)

// The FP32 x phase pairs its lines (x_lamb_body): its one-line candidates run
// the two-line shape of their length class instead
int x_sel_f32(int L, int sel) {
    if (L > 576) return sel == 1 ? 0 : sel;
    if (L > 192) return sel == 2 ? 1 : sel;
    return sel;
}

}  // namespace

void launch_z_inverse(const Chan &c, const double2 *const *in, int nf, void *out, hipStream_t s) {
    PtrTable t{};
    for (int f = 0; f < nf; ++f) t.p[f] = in[f];
    if (c.precision == 0)
        OWT_TUNED_DISPATCH(c.NZ2, c.tune[T_Z_INVERSE_TILE], {
            dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * nf));
            hipLaunchKernelGGL((k_z_inverse<double, TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, t,
                               (double2 *)out, nf, 1.0, 0);
        });
    else
        OWT_TUNED_DISPATCH_F32(c.NZ2, c.tune[T_Z_INVERSE_TILE], {
            dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * nf));
            hipLaunchKernelGGL((k_z_inverse<float, TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<float, TJ, TL, LMAX>()), s, c, t,
                               (float2 *)out, nf, 1.0f, 0);
        });
}

void launch_z_inverse_f32(const Chan &c, const float2 *six, hipStream_t s) {
    OWT_TUNED_DISPATCH_F32(c.NZ2, c.tune[T_Z_INVERSE_TILE], {
        dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * 6));
        hipLaunchKernelGGL((k_z_inverse_f32in<TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<float, TJ, TL, LMAX>()), s, c, six,
                           (float2 *)c.zbuf);
    });
}

void launch_x_lamb(const Chan &c, hipStream_t s) {
    if (c.precision == 0)
        OWT_TUNED_DISPATCH(c.NX2, c.tune[T_X_TILE], {
            dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ)));
            hipLaunchKernelGGL((k_x_lamb<double, TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c,
                               (const double2 *)c.zbuf, (double2 *)c.xbuf);
        });
    else
        OWT_TUNED_DISPATCH_F32(c.NX2, x_sel_f32(c.NX2, c.tune[T_X_TILE]), {
            dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ)));
            hipLaunchKernelGGL((k_x_lamb<float, TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<float, TJ, TL, LMAX>()), s, c,
                               (const float2 *)c.zbuf, (float2 *)c.xbuf);
        });
}

void launch_z_forward(const Chan &c, hipStream_t s) {
    if (c.precision == 0)
        OWT_TUNED_DISPATCH(c.NZ2, c.tune[T_Z_FORWARD_TILE], {
            dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * 3));
            hipLaunchKernelGGL((k_z_forward<double, TJ, TL, LMAX, double2>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s,
                               c, (const double2 *)c.xbuf, (double2 *)c.lmb, 3, (double)c.NZ2, 1);
        });
    else
        OWT_TUNED_DISPATCH_F32(c.NZ2, c.tune[T_Z_FORWARD_TILE], {
            dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * 3));
            hipLaunchKernelGGL((k_z_forward<float, TJ, TL, LMAX, float2>), g, dim3(TJ * TL), (smem_bytes<float, TJ, TL, LMAX>()), s, c,
                               (const float2 *)c.xbuf, (float2 *)c.lmb, 3, (float)c.NZ2, 1);
        });
}

void launch_z_forward_to(const Chan &c, const void *in, int nf, double2 *out, hipStream_t s) {
    Chan d = c;
    d.precision = 0;
    OWT_Z_DISPATCH(c.NZ2, double, {
        dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * nf));
        hipLaunchKernelGGL((k_z_forward<double, TJ, TL, LMAX, double2>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, d,
                           (const double2 *)in, out, nf, (double)c.NZ2, 1);
    });
}

void launch_x_tangent(const Chan &c, const double2 *base, hipStream_t s) {
    OWT_X_DISPATCH(c.NX2, double, {
        dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ)));
        hipLaunchKernelGGL((k_x_tangent<TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, base,
                           (const double2 *)c.zbuf, (double2 *)c.xbuf);
    });
}

void launch_x_adjoint(const Chan &c, const double2 *base, const double2 *lam, double2 *out, hipStream_t s) {
    OWT_X_DISPATCH(c.NX2, double, {
        dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ)));
        hipLaunchKernelGGL((k_x_adjoint<TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, base, lam, out);
    });
}

// transposes of the z transforms of the step (FP64): of k_z_forward (divide by
// NZ2, the Nyquist mode fed) and of k_z_inverse (no division, Nyquist mode zero)
void launch_z_forward_T(const Chan &c, const double2 *const *in, int nf, double2 *out, hipStream_t s) {
    PtrTable t{};
    for (int f = 0; f < nf; ++f) t.p[f] = in[f];
    OWT_Z_DISPATCH(c.NZ2, double, {
        dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * nf));
        hipLaunchKernelGGL((k_z_inverse<double, TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, t, out, nf,
                           (double)c.NZ2, 1);
    });
}

void launch_z_inverse_T(const Chan &c, const double2 *in, int nf, double2 *out, hipStream_t s) {
    OWT_Z_DISPATCH(c.NZ2, double, {
        dim3 g(xcd_grid(c.NXH * ((c.NY + TJ - 1) / TJ) * nf));
        hipLaunchKernelGGL((k_z_forward<double, TJ, TL, LMAX, double2>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, in,
                           out, nf, 1.0, 0);
    });
}

void launch_x_c2r(const Chan &c, const void *in, int nf, double *phys, hipStream_t s) {
    OWT_X_DISPATCH(c.NX2, double, {
        dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ) * ((nf + 1) / 2)));
        hipLaunchKernelGGL((k_x_c2r<TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c,
                           (const double2 *)in, nf, phys);
    });
}

void launch_x_r2c(const Chan &c, const double *phys, int nf, void *out, hipStream_t s) {
    OWT_X_DISPATCH(c.NX2, double, {
        dim3 g(xcd_grid(c.NZ2 * ((c.NY + TJ - 1) / TJ) * ((nf + 1) / 2)));
        hipLaunchKernelGGL((k_x_r2c<TJ, TL, LMAX>), g, dim3(TJ * TL), (smem_bytes<double, TJ, TL, LMAX>()), s, c, phys, nf,
                           (double2 *)out);
    });
}

void launch_peak3(const double *phys, size_t n, unsigned long long *out, hipStream_t s) {
    hipLaunchKernelGGL(k_peak3, dim3(1024), dim3(256), 0, s, phys, n, out);
}

// ---- grouped launches of the transform stages over ensemble members ----
// The instance is fixed by the precision, the length class (LMAX) and the tuned
// y-run TJ (every candidate has 256 threads); the blocks are the single launch's.
int ens_key_fft(const EnsEntry &e, int stage, int *blocks) {
    const Chan &c = e.c;
    const int L = stage == ES_X_LAMB ? c.NX2 : c.NZ2;
    int sel = c.tune[stage == ES_Z_INVERSE ? T_Z_INVERSE_TILE : stage == ES_X_LAMB ? T_X_TILE : T_Z_FORWARD_TILE];
    if (stage == ES_X_LAMB && c.precision == 1) sel = x_sel_f32(L, sel);
    int key = 0, tj = 1;
    if (c.precision == 0) OWT_TUNED_DISPATCH(L, sel, { key = LMAX * 64 + TJ; tj = TJ; });
    else OWT_TUNED_DISPATCH_F32(L, sel, { key = (1 << 20) + LMAX * 64 + TJ; tj = TJ; });
    const int nyt = (c.NY + tj - 1) / tj;
    *blocks = xcd_grid(stage == ES_X_LAMB ? c.NZ2 * nyt : c.NXH * nyt * (stage == ES_Z_INVERSE ? 6 : 3));
    return key;
}

void launch_ens_fft(int stage, const EnsEntry *ents, const int *first, int nm, int grid, const EnsEntry &rep, hipStream_t s) {
    const Chan &c = rep.c;
    const int L = stage == ES_X_LAMB ? c.NX2 : c.NZ2;
    int sel = c.tune[stage == ES_Z_INVERSE ? T_Z_INVERSE_TILE : stage == ES_X_LAMB ? T_X_TILE : T_Z_FORWARD_TILE];
    if (stage == ES_X_LAMB && c.precision == 1) sel = x_sel_f32(L, sel);
    const dim3 g(grid);
    if (c.precision == 0)
        OWT_TUNED_DISPATCH(L, sel, {
            const size_t lds = (smem_bytes<double, TJ, TL, LMAX>());
            if (stage == ES_Z_INVERSE)
                hipLaunchKernelGGL((k_z_inverse_ens<double, TJ, TL, LMAX>), g, dim3(TJ * TL), lds, s, ents, first, nm);
            else if (stage == ES_X_LAMB)
                hipLaunchKernelGGL((k_x_lamb_ens<double, TJ, TL, LMAX>), g, dim3(TJ * TL), lds, s, ents, first, nm);
            else
                hipLaunchKernelGGL((k_z_forward_ens<double, TJ, TL, LMAX, double2>), g, dim3(TJ * TL), lds, s, ents, first, nm);
        });
    else
        OWT_TUNED_DISPATCH_F32(L, sel, {
            const size_t lds = (smem_bytes<float, TJ, TL, LMAX>());
            if (stage == ES_Z_INVERSE)
                hipLaunchKernelGGL((k_z_inverse_ens<float, TJ, TL, LMAX>), g, dim3(TJ * TL), lds, s, ents, first, nm);
            else if (stage == ES_X_LAMB)
                hipLaunchKernelGGL((k_x_lamb_ens<float, TJ, TL, LMAX>), g, dim3(TJ * TL), lds, s, ents, first, nm);
            else
                hipLaunchKernelGGL((k_z_forward_ens<float, TJ, TL, LMAX, float2>), g, dim3(TJ * TL), lds, s, ents, first, nm);
        });
}

}  // namespace owt
