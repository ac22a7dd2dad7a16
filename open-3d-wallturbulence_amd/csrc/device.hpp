// Device-side data structures and launch entry points of the HIP kernels.
// One Chan describes one channel resident in HBM; it is passed to every kernel
// by value. Wave-space fields use the reference layout [NZ][NXH][NY] (complex,
// y fastest, src/setup.hpp:58-69); the z-padded intermediate is [NZ2][NXH][NY].
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace owt {

constexpr int kMaxStages = 8;
constexpr int kMaxNY = 288;

// Stored LU factors use N*N doubles in a packed layout: first the strictly lower
// part of every column (rows k+1..N-1 of column k, columns in order), then the
// upper part with the diagonal (rows 0..k of column k, columns in order). The
// forward substitution then streams the first region and the back substitution
// the second, each cache line fetched once; with plain column-major storage the
// line holding a column's diagonal and the line shared by two columns are
// fetched by both passes.
__host__ __device__ __forceinline__ size_t lu_lower_at(int N, int k) { return (size_t)k * (2 * N - k - 1) / 2; }
__host__ __device__ __forceinline__ size_t lu_upper_at(int N, int k) {
    return (size_t)N * (N - 1) / 2 + (size_t)k * (k + 1) / 2;
}
// element (row, col) of a packed factor
__host__ __device__ __forceinline__ size_t lu_at(int N, int row, int col) {
    return row > col ? lu_lower_at(N, col) + (row - col - 1) : lu_upper_at(N, col) + row;
}

// 1-D transform plan: length L = prod(radix[s]), radices in {2, 3, 4, 5}.
// Stage s has span ns[s] (product of the earlier radices), twiddle stride
// tstep[s] = L / (ns[s] * radix[s]) and mns[s] = ceil(2^32 / ns[s]), with which
// j / ns[s] = umulhi(j, mns[s]) exactly for j < 2^16 (mns = 0 marks ns = 1).
struct FFTPlan {
    int L = 0;
    int nst = 0;
    int radix[kMaxStages] = {0};
    int ns[kMaxStages] = {0};
    int tstep[kMaxStages] = {0};
    unsigned mns[kMaxStages] = {0};
};

// history field indices inside Chan::hist ([field][slot] of W elements each)
enum HistField { H_V = 0, H_G = 1, H_HV = 2, H_HG = 3 };
// mean-flow history indices inside Chan::means ([quantity][slot][NY])
enum MeanField { M_LMBX = 0, M_LMBZ = 1, M_VELX = 2, M_VELZ = 3 };
// control words inside Chan::ctrl
enum Ctrl { C_RING = 0, C_STEP = 1, C_ERR = 2 };  // C_RING is written only into checkpoints
// error bits
enum ErrBits { E_SINGULAR = 1 };

// Launch shapes (candidate indices, 0 = the default shape) chosen by the
// autotuner; every candidate computes the same values in the same order.
enum TuneKnob { T_SOLVE_JOBS = 0, T_HORIZONTAL_TP, T_NONLINEAR_TP, T_Z_INVERSE_TILE, T_X_TILE, T_Z_FORWARD_TILE, T_COUNT };
constexpr int kTuneCandidates = 4;

struct Chan {
    int NX, NY, NZ, NXH, NZH, NX2, NZ2;
    int npencil;  // NZ * NXH
    int nsys;     // NXH * (NZH + 1): distinct k^2 shared by +-kz
    size_t W;     // NZ * NXH * NY complex elements
    size_t I;     // NZ2 * NXH * NY complex elements
    double nu, dt, Ly;
    int precision;  // 0 fp64 transforms, 1 fp32 transforms
    double2 *hist;  // 12 W: v, omega_y, Hv, Hg at 3 levels
    double2 *cur;   // 4 W: u, w, omega_x, omega_z
    void *lmb;      // 3 W of double2 or float2
    void *zbuf;     // 6 I of double2 or float2 (z-inverse output)
    void *xbuf;     // 3 I of double2 or float2 (x-phase output)
    const double *DY1, *DY2, *DY4;  // N x N column-major
    const double *kx, *kz;
    double2 *means;  // 4 x 3 x NY
    double *fric;    // fx[3], fz[3]
    double2 *msol;   // 2 x NY mean-flow solutions
    double2 *d2v;    // 3 W: DY2 v at the 3 history levels (slots as in hist); DY2 v - k^2 v in mixed precision
    const double *qw;  // NY trapezoid weights of y (pressure gauge)
    int *ctrl;
    double *diag;    // diag_cap x 3: fx_fric, bottom, upper per friction evaluation
    int diag_cap;
    const double2 *twx, *twz;  // e^{-2 pi i m / L}, fp64
    const float2 *twxf, *twzf; // same, fp32
    FFTPlan px, pz;
    int tune[T_COUNT];
    // omega_y solves of the batched method: D = DY2[1..N-2, 1..N-2] = Q B Q^-1
    // (M = N - 2, column-major M x M), per row the eigenvalue's real part hlre and
    // imaginary part hlim (0 for a real one; rows 2m, 2m+1 of a pair hold a, b of
    // a +- i b and form the block [[a, b], [-b, a]])
    const double *hQ, *hQi, *hlre, *hlim;
    // mirror halves of DY1 and DY2 (A then B, h x h column-major each,
    // h = (NY+1)/2): the mixed-precision derivative products
    const double *DY1m, *DY2m;
    // the omega_y solves' mirror-parity blocks (helmholtz_mirror_basis); null
    // where their decomposition is unavailable
    const double *hmir;
    // the v-solve inverses in mirror form (k_mirror_inverse), per distinct k^2
    // mirror_inv_stride(NY) doubles; null where unavailable
    const double *Xm;
};
// per k^2: A and B (h x h column-major each, h = (NY+1)/2), P (NY x 2
// column-major), then Qe_0, Qo_0, Qe_1, Qo_1 (h each)
__host__ __device__ inline size_t mirror_inv_stride(int N) {
    const size_t h = (N + 1) / 2;
    return 2 * h * h + 2 * (size_t)N + 4 * h;
}
// the mixed-precision step takes the DY1 and DY2 products as mirror halves
// (NY <= 256: one row tile of the half range per wave)
__host__ __device__ inline bool use_mirror(const Chan &c) { return c.precision == 1 && c.NY <= 256; }
__host__ __device__ inline bool use_mirror_g(const Chan &c) { return use_mirror(c) && c.hmir != nullptr; }
__host__ __device__ inline bool use_mirror_v(const Chan &c) { return use_mirror(c) && c.Xm != nullptr; }

// History slots of one step: level 0 is the current field, levels 1 and 2 the
// two before it; the new level overwrites level 2 (slot sn = s[2]).
struct StepPtrs {
    const double2 *v[3], *g[3], *hv[3], *hg[3], *dv[3];
    double2 *vn, *gn, *hvn, *hgn, *dvn;
    int s[3], sn;
};

// ---- ensembles: one launch of a step stage over several channels ----
// A member of a grouped launch: its channel (the operator pointers may be those
// of an earlier member with the same parameters, whose operators are the same
// bit for bit), its history slots at the ensemble's ring phase, and the
// operands the single-channel launches take as arguments.
struct EnsEntry {
    Chan c;
    StepPtrs sp;
    const double2 *six[6];  // z-inverse inputs: u, v, w, omega_x, omega_y, omega_z of the new level
    const double *X;        // explicit inverses of Av (v solves)
    const int4 *jobs;
    int njobs;
    const double *Fg;  // LU factor of Ag(0) and its pivot rows (mean flow)
    const int *Pg;
};
// member of block b of a grouped launch: the largest m < nm with first[m] <= b
__device__ __forceinline__ int ens_member(const int *__restrict__ first, int nm, int b) {
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// the step's stages in order
enum EnsStage { ES_SOLVE = 0, ES_HELMHOLTZ, ES_MEAN, ES_HORIZONTAL, ES_Z_INVERSE, ES_X_LAMB, ES_Z_FORWARD, ES_NONLINEAR, ES_COUNT };
// The kernel instance a member runs in a stage (members with equal keys share a
// launch) and the blocks it takes there: a multiple of 8, so that each member's
// blocks start on the same XCD as the grid and the per-XCD tile numbering of the
// transforms and solves holds inside every member.
int ens_key(const EnsEntry &e, int stage);
int ens_blocks(const EnsEntry &e, int stage);
// one launch of `stage` over nm members of one key: ents and first are device
// arrays (first[m] the first block of member m), grid the sum of their blocks,
// rep a host copy of one of the entries, ny_max the largest NY among them
void launch_ens(int stage, const EnsEntry *ents, const int *first, int nm, int grid, const EnsEntry &rep, int ny_max,
                hipStream_t s);
// the parts of the above in fft_kernels.hip (z inverse, x phase, z forward) and
// lu_kernels.hip (v solves)
int ens_key_fft(const EnsEntry &e, int stage, int *blocks);
void launch_ens_fft(int stage, const EnsEntry *ents, const int *first, int nm, int grid, const EnsEntry &rep, hipStream_t s);
int ens_key_solve(const EnsEntry &e, int *blocks);
void launch_ens_solve(const EnsEntry *ents, const int *first, int nm, int grid, const EnsEntry &rep, hipStream_t s);

// ---- scratch of the allocation-free entry points (noalloc.hip) ----
// device address of the file-scope scratch array on `device`, looked up on the
// first call for that device (which must be current); and its size in bytes
double2 *noalloc_scratch(int device);
size_t noalloc_scratch_bytes();

// ---- FFT phase launches (fft_kernels.hip) ----
// z inverse: nf wave fields -> nf z-padded fields; fields given as pointer table
void launch_z_inverse(const Chan &c, const double2 *const *in, int nf, void *out, hipStream_t s);
// the step's z inverse (FP32 transforms) from the six fields in FP32 (6 W float2,
// written by launch_horizontal's out32) into Chan::zbuf
void launch_z_inverse_f32(const Chan &c, const float2 *six, hipStream_t s);
// fused x phase: 6 z-padded fields -> Lamb vector (3 z-padded fields)
void launch_x_lamb(const Chan &c, hipStream_t s);
// z forward: 3 z-padded fields -> 3 wave fields (c.lmb)
void launch_z_forward(const Chan &c, hipStream_t s);
// stand-alone x transforms used by the physical-space utilities:
// c2r: nf z-padded fields -> physical [NZ2][NX2][NY] doubles; r2c: inverse
void launch_x_c2r(const Chan &c, const void *in, int nf, double *phys, hipStream_t s);
void launch_x_r2c(const Chan &c, const double *phys, int nf, void *out, hipStream_t s);
// z forward of nf fields into given double2 wave outputs
void launch_z_forward_to(const Chan &c, const void *in, int nf, double2 *out, hipStream_t s);
// max over the physical grid of u^2+v^2+w^2 (three consecutive fields), stored as bits
void launch_peak3(const double *phys, size_t n, unsigned long long *out, hipStream_t s);

// ---- wall-normal phase launches (lu_kernels.hip, y_kernels.hip) ----
// LU factors for nfac factor indices (see lu_kernels.hip): operators op0 ..
// op0 + nops - 1 (0 Av, 1 Ag, 2 pressure), stored nops per factor index;
// frep[f]: representative system or pencil of factor f (null: f itself)
// Above 256 rows (or with force_panel) the factorization runs by panels and
// needs scratch for chunk x nops matrices of NY x NY doubles; launches are
// issued per chunk of factor indices. factor_scratch_rows: NY when scratch is
// needed, else 0.
int factor_scratch_rows(const Chan &c, bool force_panel);
void launch_factor(const Chan &c, double *F, int *P, int nfac, const int *frep, int op0, int nops, double *scratch,
                   int chunk, bool force_panel, hipStream_t s);
// dynamic pressure of the current level: right-hand sides into R (W) from the
// Lamb vector (in the transform precision) and dv = DY2 v of the level; then the
// solves with the pressure factors (op 2 of the batched factor tables) into out
void launch_pressure_rhs(const Chan &c, const double2 *dv, const double2 *v, double2 *R, hipStream_t s);
// Chan::d2v's three levels into the form of the given precision (1 mixed, 0 FP64)
void launch_d2v_convert(const Chan &c, int mixed, hipStream_t s);
void launch_pressure_solve(const Chan &c, const double *F, const int *P, const int4 *jobs, int njobs, const double2 *R,
                           double2 *out, hipStream_t s);
// right-hand sides + solves; plain: njobs = pencils, each with its own factor;
// batched: job j = {factor, system, system or -1}. nops: operators per factor
// index in F (2: Av and Ag, both solved; 1: Av only, omega_y by launch_helmholtz_g)
void launch_solve(const Chan &c, const StepPtrs &sp, const double *F, const int *P, const int4 *jobs, int njobs, bool plain,
                  int nops, hipStream_t s);
// explicit inverses X_f = Av_f^-1 (N x N column-major each) of nfac Av factors
// (bops 1), refined once; scratch holds chunk matrices
void launch_invert(const Chan &c, const double *F, const int *P, int nfac, double *X, double *scratch,
                   int chunk, hipStream_t s);
// mirror form of nfac inverses X (mirror_inv_stride doubles each into Xm) from
// the rank-2 factors UV of the odd part of Av (mirror_lowrank); NY <= 256.
// *bad (zeroed by the caller) becomes nonzero if a k^2's Woodbury matrix is
// far from the identity
void launch_mirror_inverse(const Chan &c, const double *X, int nfac, const double *UV, double *Xm, int *bad, hipStream_t s);
// the batched method's v solves with the inverses X of Av (one per factor index);
// in mirror form from c.Xm where use_mirror_v(c)
void launch_solve_inv(const Chan &c, const StepPtrs &sp, const double *X, const int4 *jobs, int njobs, hipStream_t s);
// omega_y of the new level by the eigendecomposition of DY2 (batched method)
void launch_helmholtz_g(const Chan &c, const StepPtrs &sp, hipStream_t s);
// exact: the plain method's kernels in the reference's arithmetic (column-order
// mat-vecs, separate roundings); otherwise the matrix-core kernels
// out32 (the mirror kernel only, else ignored): the z inverse's six input fields
// in FP32 go there instead of u, w, omega_x, omega_z into Chan::cur
void launch_horizontal(const Chan &c, const StepPtrs &sp, bool exact, hipStream_t s, float2 *out32 = nullptr);
// mean flow, friction force and the k = 0 pencil (Fg, Pg: the LU factor of Ag(0))
void launch_mean(const Chan &c, const StepPtrs &sp, const double *Fg, const int *Pg, hipStream_t s);
void launch_nonlinear(const Chan &c, const StepPtrs &sp, bool exact, hipStream_t s);
// initialisation helpers (src/solver.cpp:379-425)
void launch_vor_from_vel(const Chan &c, hipStream_t s);
void launch_init_history(const Chan &c, hipStream_t s);
void launch_init_mean(const Chan &c, hipStream_t s);
// DY2 v of the initial field into d2v slot 0
void launch_init_d2v(const Chan &c, hipStream_t s);

// ---- tangent and adjoint (FP64; fft_kernels.hip, lu_kernels.hip, y_kernels.hip) ----
// x phase of the linearized Lamb vector: c.zbuf (tangent six) with base (base six,
// z-transformed) -> c.xbuf
void launch_x_tangent(const Chan &c, const double2 *base, hipStream_t s);
// transpose of the x phase about base: lam (3 z-padded) -> out (6 z-padded)
void launch_x_adjoint(const Chan &c, const double2 *base, const double2 *lam, double2 *out, hipStream_t s);
// transposes of the z forward (nf wave fields -> nf z-padded) and of the z inverse
// (nf z-padded -> nf wave fields)
void launch_z_forward_T(const Chan &c, const double2 *const *in, int nf, double2 *out, hipStream_t s);
void launch_z_inverse_T(const Chan &c, const double2 *in, int nf, double2 *out, hipStream_t s);
// factor set -> the form solving with A^T (nmat matrices of N x N, in place)
void launch_factor_transpose(double *F, int nmat, int N, hipStream_t s);
// transposes of the step's stages on an adjoint state c at the ring phase of sp
// Xv: the forward step's explicit inverses of Av (null: the v solve transposes the LU factors too)
void launch_solve_adj(const Chan &c, const StepPtrs &sp, const double *G, const int *P, const double *Xv, const int4 *jobs,
                      int njobs, hipStream_t s);
void launch_horizontal_adj(const Chan &c, const StepPtrs &sp, hipStream_t s);
void launch_mean_adj(const Chan &c, const StepPtrs &sp, const double *G, const int *P, int mf, double seed, hipStream_t s);
void launch_nonlinear_adj(const Chan &c, const StepPtrs &sp, hipStream_t s);
// transposes of the initial post-processing: mean and history, then the
// gradient with respect to the initial u, v, w
void launch_init_adj_mean_history(const Chan &c, hipStream_t s);
void launch_init_vel_adj(const Chan &c, double2 *gu, double2 *gv, double2 *gw, hipStream_t s);
// plane statistics of the current level (v: its v slot) into out[9][NY]:
// U, V, W, <u'u'>, <v'v'>, <w'w'>, <u'v'>, <u'w'>, <v'w'>
void launch_statistics(const Chan &c, const double2 *v, double *out, hipStream_t s);
// dst[f] += src[f], f < nf (<= 6), n complex values each
void launch_accumulate(double2 *const *dst, double2 *const *src, int nf, size_t n, hipStream_t s);

}  // namespace owt
