// Internal declarations shared by the host side and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace owt {

// Raised by host code; the C ABI turns it into a status code + message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define OWT_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::owt::Error(3, std::string("HIP error: ") + hipGetErrorString(e_) + " at " + \
                                      __FILE__ + ":" + std::to_string(__LINE__));              \
    } while (0)

// Reference toml keys (src/IO.cpp:14-25) and derived sizes (src/IO.cpp:27-34).
struct Params {
    double nu = 0, dt = 0, dkx = 0, dkz = 0, alpha_mesh = 0;
    long NX = 0, NY = 0, NZ = 0, NT = 0, PX = 1, PZ = 1, save_interval = 1;
    long NXH = 0, NZH = 0, NX2 = 0, NZ2 = 0;
    // extension keys (namespaced in the toml as owt.*)
    int precision = 0;  // 0 = FP64 everywhere, 1 = FP32 transform phase
    int method = 1;     // 0 = plain (per-step LU, eager), 1 = batched (factor once, graph)
    uint64_t seed = 0;
    double amp = 0.15;
    // initial field of a run from a case file: 1 = the reference's init_fields
    // (laminar + (2, +-2) wave), 2 = seeded recipe, 3 = laminar only, 4 = restart
    // from the full-state checkpoint `restart`
    int init_mode = 1;
    std::string restart = "DATA/restart.ckpt";
};

Params load_params(const std::string &path);
void derive_sizes(Params &p);

// Host-side wall-normal discretisation (src/solver.cpp:15-216, src/compdiff.hpp).
struct Operators {
    int N = 0;
    double ds = 0;
    std::vector<double> s, y, dyds1, dyds2, dyds3, dyds4;
    std::vector<double> c11, c21, c22, c31, c32, c33, c41, c42, c43, c44;
    // column-major N x N
    std::vector<double> A[4], B[4], D[4], DY1, DY2, DY3, DY4;
};
Operators build_operators(int NY, double alpha_mesh);

// Mirror halves of a wall-normal operator M (column-major N x N) that is even
// (s = +1: DY2) or odd (s = -1: DY1) under the reflection j -> N-1-j, up to the
// rounding of its construction (a few hundred ulps). With h = (N+1)/2 and, for
// j < h, e_j = x_j + x_{N-1-j} and o_j = x_j - x_{N-1-j}, the symmetrised
// operator Ms = (M + s J M J) / 2 gives on rows i < h
//   (Ms x)_i = sum_j A_ij e_j + B_ij o_j,
//   A_ij = (Ms_ij + Ms_i,N-1-j) / 2,   B_ij = (Ms_ij - Ms_i,N-1-j) / 2
// (for odd N the middle column c: A_ic = Ms_ic / 2, B_ic = 0), and on the rows
// below (Ms x)_{N-1-i} = s (sum_j A_ij e_j - B_ij o_j). Returns A then B,
// h x h column-major each.
std::vector<double> mirror_halves(const std::vector<double> &M, int N, int s);

// The part of the v operator Av (row N-2 negated) that is odd under the
// reflection: E = -nu (DY4 - J DY4 J) / 2 on rows 2..N-3, the same for every
// k^2 (DY1 and DY2 are mirror symmetric to rounding, the boundary rows map onto
// each other). The reference's DY4 makes E rank 2: E = U V^T. UV receives U
// then V (N x 2 column-major each). False if E is not rank 2 to 1e-12 of its
// largest entry.
bool mirror_lowrank(const std::vector<double> &DY4, int N, double nu, std::vector<double> &UV);

// D = DY2's interior block (rows and columns 1..N-2) as Q B Q^-1, B block
// diagonal: rows (2m, 2m+1), m < npair, a 2 x 2 block [[a, b], [-b, a]] of the
// pair a +- i b (lre = a, lim = b on both rows); the other rows real eigenvalues
// (lim = 0). Q, Qinv: M x M column-major, M = N - 2. residual = max|D Q - Q B| /
// (|D| |Q|), identity = max|Q Qinv - I| (eigen.cpp)
struct HelmholtzBasis {
    int M = 0, npair = 0;
    std::vector<double> Q, Qinv, lre, lim;
    double residual = 0, identity = 0;
};
HelmholtzBasis helmholtz_basis(const std::vector<double> &DY2, int N);

// The same decomposition split by mirror parity (the mixed-precision step):
// the symmetrised interior block acts on even and odd vectors through two blocks
// of the half sizes hm = (M+1)/2 and ho = M/2, decomposed each (eigen.cpp).
// Returns, hm x hm column-major each (the odd block padded with a zero last
// row and column when M is odd): Q_E^-1 / 2, Q_E, Q_O^-1 / 2, Q_O, then the
// eigenvalue rows lre_E, lim_E, lre_O, lim_O (hm each, the pad 0).
std::vector<double> helmholtz_mirror_basis(const std::vector<double> &DY2, int N);

// host dense LU with partial pivoting (dgetrf semantics), column-major
int host_lu_factor(int n, double *A, int *ipiv);
void host_lu_solve(int n, const double *LU, const int *ipiv, double *b);

}  // namespace owt
