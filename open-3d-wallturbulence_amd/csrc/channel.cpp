// Channel runtime: device memory, initial fields, the step sequence and its
// hipGraph capture, field access and full-state checkpoints.
#include "channel.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>

#include "snapshot.hpp"

namespace owt {

FFTPlan make_plan(int L) {
    FFTPlan p;
    p.L = L;
    int n = L;
    while (n % 4 == 0 && p.nst < kMaxStages) p.radix[p.nst++] = 4, n /= 4;
    while (n % 3 == 0 && p.nst < kMaxStages) p.radix[p.nst++] = 3, n /= 3;
    while (n % 2 == 0 && p.nst < kMaxStages) p.radix[p.nst++] = 2, n /= 2;
    while (n % 5 == 0 && p.nst < kMaxStages) p.radix[p.nst++] = 5, n /= 5;
    if (n != 1) throw Error(2, "transform length " + std::to_string(L) + " is not 2^a 3^b 5^c with at most 8 stages");
    int ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        p.ns[s] = ns;
        p.tstep[s] = L / (ns * p.radix[s]);
        p.mns[s] = ns == 1 ? 0u : (unsigned)((((unsigned long long)1 << 32) + (unsigned)ns - 1) / (unsigned)ns);
        ns *= p.radix[s];
    }
    return p;
}

namespace {

std::vector<double2> twiddles(int L) {
    std::vector<double2> t(L);
    for (int m = 0; m < L; ++m) {
        long double th = 2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)L;
        t[m] = make_double2((double)cosl(th), -(double)sinl(th));
    }
    return t;
}

}  // namespace

void *Channel::dmalloc(size_t bytes) {
    void *p = nullptr;
    if (bytes == 0) bytes = 16;
    OWT_HIP(hipMalloc(&p, bytes));
    allocs_.push_back(p);
    return p;
}

Channel::Channel(const Params &p, int device, bool verbose, hipStream_t stream) : P_(p), device_(device), verbose_(verbose) {
    const char *fp = std::getenv("OWT_FACTOR_PANEL");
    force_panel_ = fp && fp[0] == '1';
    const char *fm = std::getenv("OWT_MIRROR_SOLVE");
    mirror_solve_ = !(fm && fm[0] == '0');
    int ndev = 0;
    OWT_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(3, "no HIP device " + std::to_string(device));
    OWT_HIP(hipSetDevice(device));
    if (stream) {
        s_ = stream;
        own_stream_ = false;
    } else {
        OWT_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    }
    OWT_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    OWT_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    OWT_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    noalloc_ = noalloc_scratch(device);
    op_ = build_operators((int)P_.NY, P_.alpha_mesh);
    try {
        alloc();
        upload_operators();
    } catch (...) {
        free_all();
        throw;
    }
}

Channel::~Channel() {
    (void)hipSetDevice(device_);
    tangent_.reset();
    adjoint_.reset();
    drop_graph();
    free_all();
}

void Channel::free_all() {
    if (s_) (void)hipStreamSynchronize(s_);
    if (side_) (void)hipStreamSynchronize(side_);
    for (void *p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (side_) (void)hipStreamDestroy(side_);
    ev_fork_ = ev_join_ = nullptr;
    side_ = nullptr;
    if (s_ && own_stream_) (void)hipStreamDestroy(s_);
    s_ = nullptr;
}

void Channel::alloc() {
    Chan &c = c_;
    c.NX = (int)P_.NX;
    c.NY = (int)P_.NY;
    c.NZ = (int)P_.NZ;
    c.NXH = (int)P_.NXH;
    c.NZH = (int)P_.NZH;
    c.NX2 = (int)P_.NX2;
    c.NZ2 = (int)P_.NZ2;
    c.npencil = c.NZ * c.NXH;
    c.nsys = c.NXH * (c.NZH + 1);
    c.W = (size_t)c.NZ * c.NXH * c.NY;
    c.I = (size_t)c.NZ2 * c.NXH * c.NY;
    c.nu = P_.nu;
    c.dt = P_.dt;
    c.Ly = op_.y[c.NY - 1] - op_.y[0];
    c.precision = P_.precision;
    const size_t es = P_.precision == 0 ? sizeof(double2) : sizeof(float2);
    c.hist = (double2 *)dmalloc(12 * c.W * sizeof(double2));
    c.cur = (double2 *)dmalloc(4 * c.W * sizeof(double2));
    six32_ = (float2 *)dmalloc(6 * c.W * sizeof(float2));
    // transform buffers sized for FP64 so the precision can be switched in place
    c.lmb = dmalloc(3 * c.W * sizeof(double2));
    c.zbuf = dmalloc(6 * c.I * sizeof(double2));
    c.xbuf = dmalloc(3 * c.I * sizeof(double2));
    (void)es;
    const size_t N = c.NY;
    c.DY1 = (const double *)dmalloc(N * N * sizeof(double));
    c.DY2 = (const double *)dmalloc(N * N * sizeof(double));
    c.DY4 = (const double *)dmalloc(N * N * sizeof(double));
    c.qw = (const double *)dmalloc(N * sizeof(double));
    c.hQ = (const double *)dmalloc((N - 2) * (N - 2) * sizeof(double));
    c.hQi = (const double *)dmalloc((N - 2) * (N - 2) * sizeof(double));
    c.hlre = (const double *)dmalloc((N - 2) * sizeof(double));
    c.hlim = (const double *)dmalloc((N - 2) * sizeof(double));
    const size_t h = (N + 1) / 2;
    c.DY1m = (const double *)dmalloc(2 * h * h * sizeof(double));
    c.DY2m = (const double *)dmalloc(2 * h * h * sizeof(double));
    c.kx = (const double *)dmalloc(c.NXH * sizeof(double));
    c.kz = (const double *)dmalloc(c.NZ * sizeof(double));
    c.means = (double2 *)dmalloc(12 * N * sizeof(double2));
    c.fric = (double *)dmalloc(8 * sizeof(double));
    c.msol = (double2 *)dmalloc(2 * N * sizeof(double2));
    c.d2v = (double2 *)dmalloc(3 * c.W * sizeof(double2));
    c.ctrl = (int *)dmalloc(16 * sizeof(int));
    c.diag_cap = 1 << 16;
    c.diag = (double *)dmalloc((size_t)c.diag_cap * 3 * sizeof(double));
    c.twx = (const double2 *)dmalloc(c.NX2 * sizeof(double2));
    c.twz = (const double2 *)dmalloc(c.NZ2 * sizeof(double2));
    c.twxf = (const float2 *)dmalloc(c.NX2 * sizeof(float2));
    c.twzf = (const float2 *)dmalloc(c.NZ2 * sizeof(float2));
    c.px = make_plan(c.NX2);
    c.pz = make_plan(c.NZ2);
    OWT_HIP(hipMemsetAsync(c.hist, 0, 12 * c.W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c.cur, 0, 4 * c.W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c.lmb, 0, 3 * c.W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c.ctrl, 0, 16 * sizeof(int), s_));
    OWT_HIP(hipMemsetAsync(c.diag, 0, (size_t)c.diag_cap * 3 * sizeof(double), s_));
}

void Channel::upload_operators() {
    const size_t N = c_.NY;
    OWT_HIP(hipMemcpyAsync((void *)c_.DY1, op_.DY1.data(), N * N * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.DY2, op_.DY2.data(), N * N * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.DY4, op_.DY4.data(), N * N * sizeof(double), hipMemcpyHostToDevice, s_));
    const std::vector<double> m1 = mirror_halves(op_.DY1, (int)N, -1), m2 = mirror_halves(op_.DY2, (int)N, +1);
    OWT_HIP(hipMemcpyAsync((void *)c_.DY1m, m1.data(), m1.size() * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.DY2m, m2.data(), m2.size() * sizeof(double), hipMemcpyHostToDevice, s_));
    // trapezoid weights of the y grid (pressure gauge): (y[j+1] - y[j-1]) / 2,
    // one-sided at the walls
    std::vector<double> qw(N);
    for (size_t j = 0; j < N; ++j) {
        const double lo = op_.y[j > 0 ? j - 1 : 0], hi = op_.y[j + 1 < N ? j + 1 : N - 1];
        qw[j] = (hi - lo) / 2.0;
    }
    OWT_HIP(hipMemcpyAsync((void *)c_.qw, qw.data(), N * sizeof(double), hipMemcpyHostToDevice, s_));
    // wavenumbers (src/solver.cpp:22-37) with PX = PZ = 1
    std::vector<double> kx(c_.NXH), kz(c_.NZ);
    for (int i = 0; i < c_.NXH; ++i) kx[i] = (double)i * P_.dkx;
    for (int k = 0; k < c_.NZ; ++k) {
        double t = (double)k;
        if (t >= c_.NZH) t -= c_.NZ;
        kz[k] = t * P_.dkz;
    }
    OWT_HIP(hipMemcpyAsync((void *)c_.kx, kx.data(), kx.size() * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.kz, kz.data(), kz.size() * sizeof(double), hipMemcpyHostToDevice, s_));
    // eigendecomposition of DY2's interior block for the omega_y solves
    const HelmholtzBasis hb = helmholtz_basis(op_.DY2, (int)N);
    const size_t M = hb.M;
    OWT_HIP(hipMemcpyAsync((void *)c_.hQ, hb.Q.data(), M * M * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.hQi, hb.Qinv.data(), M * M * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.hlre, hb.lre.data(), M * sizeof(double), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.hlim, hb.lim.data(), M * sizeof(double), hipMemcpyHostToDevice, s_));
    // its mirror-parity blocks (the mixed step at NY <= 256); the full
    // decomposition serves where they are unavailable
    c_.hmir = nullptr;
    if (N <= 256) {
        std::vector<double> mb;
        try {
            mb = helmholtz_mirror_basis(op_.DY2, (int)N);
        } catch (const Error &) {
            mb.clear();
        }
        if (!mb.empty()) {
            c_.hmir = (const double *)dmalloc(mb.size() * sizeof(double));
            OWT_HIP(hipMemcpyAsync((void *)c_.hmir, mb.data(), mb.size() * sizeof(double), hipMemcpyHostToDevice, s_));
        }
    }
    auto tx = twiddles(c_.NX2), tz = twiddles(c_.NZ2);
    std::vector<float2> txf(tx.size()), tzf(tz.size());
    for (size_t m = 0; m < tx.size(); ++m) txf[m] = make_float2((float)tx[m].x, (float)tx[m].y);
    for (size_t m = 0; m < tz.size(); ++m) tzf[m] = make_float2((float)tz[m].x, (float)tz[m].y);
    OWT_HIP(hipMemcpyAsync((void *)c_.twx, tx.data(), tx.size() * sizeof(double2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.twz, tz.data(), tz.size() * sizeof(double2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.twxf, txf.data(), txf.size() * sizeof(float2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync((void *)c_.twzf, tzf.data(), tzf.size() * sizeof(float2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipStreamSynchronize(s_));
}

// The operators Av(k^2), Ag(k^2) do not change in time and depend on k^2
// only: factor each distinct value once (the reference refactors every pencil
// every step). Systems are grouped by the bit pattern of k^2 as the device
// computes it, kx*kx + kz*kz with each product rounded. The step solves the
// omega_y systems through the eigendecomposition of DY2 (k_helmholtz_g) and the
// v systems with the explicit inverse of each Av (k_solve_inv), so Av is
// factored per k^2 (bops_ = 1) and inverted, plus Ag(0) is factored for the mean
// flow; an adjoint channel, whose transposed solves use LU factors of both,
// keeps both (bops_ = 2).
void Channel::factor_operators() {
    if (factored_) return;
    const size_t N = c_.NY;
    const int ns = c_.nsys, NXH = c_.NXH;
    if (!Fb_) {
        std::map<double, std::vector<int>> groups;
        for (int sys = 0; sys < ns; ++sys) {
            const int mz = sys / NXH, i = sys % NXH;
            double kzv = (double)mz;
            if (kzv >= c_.NZH) kzv -= c_.NZ;
            volatile double a = ((double)i * P_.dkx) * ((double)i * P_.dkx);
            volatile double b = (kzv * P_.dkz) * (kzv * P_.dkz);
            groups[a + b].push_back(sys);
        }
        std::vector<int4> jobs;
        std::vector<int> frep;
        fmap_host_.assign(ns, 0);
        for (auto &g : groups) {
            const int f = (int)frep.size();
            const std::vector<int> &sys = g.second;
            frep.push_back(sys.front());
            for (int s : sys) fmap_host_[s] = f;
            for (size_t a = 0; a < sys.size(); a += 2)
                jobs.push_back(make_int4(f, sys[a], a + 1 < sys.size() ? sys[a + 1] : -1, 0));
        }
        nfac_ = (int)frep.size();
        njobs_ = (int)jobs.size();
        Fb_ = (double *)dmalloc((size_t)bops_ * nfac_ * N * N * sizeof(double));
        Pb_ = (int *)dmalloc((size_t)bops_ * nfac_ * N * sizeof(int));
        jobs_ = (int4 *)dmalloc(njobs_ * sizeof(int4));
        frep_ = (int *)dmalloc(nfac_ * sizeof(int));
        OWT_HIP(hipMemcpy(jobs_, jobs.data(), njobs_ * sizeof(int4), hipMemcpyHostToDevice));
        OWT_HIP(hipMemcpy(frep_, frep.data(), nfac_ * sizeof(int), hipMemcpyHostToDevice));
        if (bops_ == 1) {
            // the mean flow's Ag(0) (system 0: kx = kz = 0), the one omega_y operator
            // kept as an LU factor
            Fg0_ = (double *)dmalloc(N * N * sizeof(double));
            Pg0_ = (int *)dmalloc(N * sizeof(int));
            sys0_ = (int *)dmalloc(sizeof(int));
            OWT_HIP(hipMemset(sys0_, 0, sizeof(int)));
        }
    }
    factor_set(Fb_, Pb_, nfac_, frep_, 0, bops_);
    if (bops_ == 1) {
        factor_set(Fg0_, Pg0_, 1, sys0_, 1, 1);
        Fmean_ = Fg0_;
        Pmean_ = Pg0_;
        invert_operators();
    } else {
        Fmean_ = Fb_ + ((size_t)fmap_host_[0] * 2 + 1) * N * N;
        Pmean_ = Pb_ + ((size_t)fmap_host_[0] * 2 + 1) * N;
    }
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    check_errors();
    factored_ = true;
}

// Dynamic pressure of the current level into out (W complex). The pressure
// operator depends on DY1, DY2 and k^2 only: its factors (one per distinct k^2,
// the batched tables' grouping) are computed on the first request and kept.
// The step's transform buffer zbuf is free between steps and holds the
// right-hand sides and the solution.
void Channel::pressure(double *out) {
    factor_operators();
    const size_t N = c_.NY, W = c_.W;
    if (!Fq_) {
        Fq_ = (double *)dmalloc((size_t)nfac_ * N * N * sizeof(double));
        Pq_ = (int *)dmalloc((size_t)nfac_ * N * sizeof(int));
        factor_set(Fq_, Pq_, nfac_, frep_, 2, 1);
        OWT_HIP(hipGetLastError());
        OWT_HIP(hipStreamSynchronize(s_));
        check_errors();
    }
    double2 *R = (double2 *)c_.zbuf, *X = R + W;
    launch_pressure_rhs(c_, c_.d2v + (size_t)ring_ * W, c_.hist + (size_t)(H_V * 3 + ring_) * W, R, s_);
    launch_pressure_solve(c_, Fq_, Pq_, jobs_, njobs_, R, X, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipMemcpyAsync(out, X, W * sizeof(double2), hipMemcpyDeviceToHost, s_));
    OWT_HIP(hipStreamSynchronize(s_));
}

void Channel::get_factor(int sys, int op, double *lu, int *perm) {
    if (sys < 0 || sys >= c_.nsys || op < 0 || op > 1) throw Error(2, "factor index out of range");
    factor_operators();
    const size_t N = c_.NY;
    const double *F = Fb_;
    const int *P = Pb_;
    size_t f = (size_t)fmap_host_[sys] * bops_ + op;
    if (op == 1 && bops_ == 1) {
        // the step keeps no Ag factors: factor them on request
        if (!Fag_) {
            Fag_ = (double *)dmalloc((size_t)nfac_ * N * N * sizeof(double));
            Pag_ = (int *)dmalloc((size_t)nfac_ * N * sizeof(int));
            factor_set(Fag_, Pag_, nfac_, frep_, 1, 1);
            OWT_HIP(hipGetLastError());
            OWT_HIP(hipStreamSynchronize(s_));
            check_errors();
        }
        F = Fag_;
        P = Pag_;
        f = (size_t)fmap_host_[sys];
    }
    // the device keeps the packed layout of lu_at; hand out dgetrf's column-major one
    std::vector<double> packed(N * N);
    OWT_HIP(hipMemcpy(packed.data(), F + f * N * N, N * N * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t col = 0; col < N; ++col)
        for (size_t row = 0; row < N; ++row) lu[col * N + row] = packed[lu_at((int)N, (int)row, (int)col)];
    OWT_HIP(hipMemcpy(perm, P + f * N, N * sizeof(int), hipMemcpyDeviceToHost));
}

// LU factors of nfac operators (launch_factor). Above 256 rows the factorization
// runs by panels with a scratch copy of each matrix: factor indices go in
// chunks that keep the scratch at about 1 GiB, allocated on first use. The
// environment variable OWT_FACTOR_PANEL=1, read when the channel is created,
// forces the panel kernel at every NY (its factors equal the register kernel's).
void Channel::factor_set(double *F, int *P, int nfac, const int *frep, int op0, int nops) {
    const int rows = factor_scratch_rows(c_, force_panel_);
    if (rows > 0 && !fscratch_) {
        const size_t per = (size_t)2 * rows * rows * sizeof(double);
        fchunk_ = (int)std::max<size_t>(1, ((size_t)1 << 30) / per);
        fscratch_ = (double *)dmalloc((size_t)fchunk_ * per);
    }
    launch_factor(c_, F, P, nfac, frep, op0, nops, fscratch_, fchunk_ > 0 ? fchunk_ : 1, force_panel_, s_);
}

// Explicit inverses of the Av factors for the batched v solves (launch_invert),
// the unrefined inverses of up to about 1 GiB of matrices at a time in scratch.
void Channel::invert_operators() {
    const size_t per = (size_t)c_.NY * c_.NY * sizeof(double);
    if (!Xb_) {
        Xb_ = (double *)dmalloc((size_t)nfac_ * per);
        xchunk_ = (int)std::min<size_t>((size_t)nfac_, std::max<size_t>(1, ((size_t)1 << 30) / per));
        xscratch_ = (double *)dmalloc((size_t)xchunk_ * per);
    }
    launch_invert(c_, Fb_, Pb_, nfac_, Xb_, xscratch_, xchunk_, s_);
    // mirror form (k_mirror_inverse): half the bytes per step for the mixed
    // step's v solves; the full inverses stay for FP64. The environment variable
    // OWT_MIRROR_SOLVE=0, read when the channel is created, keeps the mixed step
    // on the full inverses.
    if (mirror_solve_ && c_.NY <= 256 && !UV_) {
        std::vector<double> uv;
        if (mirror_lowrank(op_.DY4, c_.NY, c_.nu, uv)) {
            UV_ = (double *)dmalloc(uv.size() * sizeof(double) + sizeof(int));
            OWT_HIP(hipMemcpy(UV_, uv.data(), uv.size() * sizeof(double), hipMemcpyHostToDevice));
            Xm_ = (double *)dmalloc((size_t)nfac_ * mirror_inv_stride(c_.NY) * sizeof(double));
        }
    }
    if (Xm_) {
        int *bad = (int *)(UV_ + 4 * (size_t)c_.NY), flag = 0;
        OWT_HIP(hipMemsetAsync(bad, 0, sizeof(int), s_));
        launch_mirror_inverse(c_, Xb_, nfac_, UV_, Xm_, bad, s_);
        OWT_HIP(hipMemcpyAsync(&flag, bad, sizeof(int), hipMemcpyDeviceToHost, s_));
        OWT_HIP(hipStreamSynchronize(s_));
        c_.Xm = flag ? nullptr : Xm_;
    }
}

// the plain method's per-pencil factor set (refilled every step)
void Channel::ensure_plain_factors() {
    if (Fp_) return;
    const size_t N = c_.NY, np = c_.npencil;
    Fp_ = (double *)dmalloc(2 * np * N * N * sizeof(double));
    Pp_ = (int *)dmalloc(2 * np * N * sizeof(int));
}

void Channel::check_errors() {
    int err = 0;
    OWT_HIP(hipMemcpy(&err, c_.ctrl + C_ERR, sizeof(int), hipMemcpyDeviceToHost));
    if (err & E_SINGULAR) {
        int zero = 0;
        OWT_HIP(hipMemcpy(c_.ctrl + C_ERR, &zero, sizeof(int), hipMemcpyHostToDevice));
        throw Error(4, "LU factorization failed. Matrix A may be singular.");
    }
}

void Channel::drop_graph() {
    ++generation_;
    for (int r = 0; r < 3; ++r) {
        if (gexec_[r]) (void)hipGraphExecDestroy(gexec_[r]);
        if (graph_[r]) (void)hipGraphDestroy(graph_[r]);
        gexec_[r] = nullptr;
        graph_[r] = nullptr;
    }
}

void Channel::set_method(int method) {
    if (method != 0 && method != 1) throw Error(2, "method must be 0 (plain) or 1 (batched)");
    P_.method = method;
    drop_graph();
}

void Channel::set_precision(int precision) {
    if (precision != 0 && precision != 1) throw Error(2, "precision must be 0 (fp64) or 1 (mixed)");
    if (precision == P_.precision) return;
    sync();
    refresh_cur();  // while the mirror kernels still apply
    // the Lamb vector is kept in the transform precision: convert it in place
    std::vector<double> l(3 * c_.W * 2);
    if (P_.precision == 0) {
        OWT_HIP(hipMemcpy(l.data(), c_.lmb, l.size() * sizeof(double), hipMemcpyDeviceToHost));
        std::vector<float> f(l.begin(), l.end());
        OWT_HIP(hipMemcpy(c_.lmb, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(l.size());
        OWT_HIP(hipMemcpy(f.data(), c_.lmb, f.size() * sizeof(float), hipMemcpyDeviceToHost));
        std::vector<double> d(f.begin(), f.end());
        OWT_HIP(hipMemcpy(c_.lmb, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    P_.precision = precision;
    c_.precision = precision;
    // the stored wall-normal terms of the history take the new precision's form
    launch_d2v_convert(c_, precision, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    drop_graph();
}

// ---------------------------------------------------------------------------
// initial fields
// ---------------------------------------------------------------------------
namespace {

uint64_t splitmix(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// complex N(0,1) keyed by (seed, field, i, m, p): SplitMix64 + Box-Muller
void seeded_gauss(uint64_t seed, int field, int i, int m, int p, double &re, double &im) {
    uint64_t h = splitmix(seed ^ 0x5851F42D4C957F2DULL);
    h = splitmix(h ^ (uint64_t)(int64_t)field);
    h = splitmix(h ^ (uint64_t)(int64_t)i);
    h = splitmix(h ^ (uint64_t)(int64_t)m);
    h = splitmix(h ^ (uint64_t)(int64_t)p);
    const uint64_t h2 = splitmix(h);
    const double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 6.283185307179586476925286766559 * u2;
    re = r * std::cos(th);
    im = r * std::sin(th);
}

}  // namespace

double Channel::peak_velocity_dev(const std::vector<double> &u, const std::vector<double> &v, const std::vector<double> &w) {
    const size_t W = c_.W, I = c_.I, P = (size_t)c_.NZ2 * c_.NX2 * c_.NY;
    std::vector<void *> tmp;
    auto tm = [&](size_t b) {
        void *p = nullptr;
        OWT_HIP(hipMalloc(&p, b));
        tmp.push_back(p);
        return p;
    };
    double peak = 0.0;
    try {
        double2 *wave = (double2 *)tm(3 * W * sizeof(double2));
        double2 *zp = (double2 *)tm(3 * I * sizeof(double2));
        double *phys = (double *)tm(3 * P * sizeof(double));
        unsigned long long *mx = (unsigned long long *)tm(sizeof(unsigned long long));
        OWT_HIP(hipMemcpyAsync(wave, u.data(), W * sizeof(double2), hipMemcpyHostToDevice, s_));
        OWT_HIP(hipMemcpyAsync(wave + W, v.data(), W * sizeof(double2), hipMemcpyHostToDevice, s_));
        OWT_HIP(hipMemcpyAsync(wave + 2 * W, w.data(), W * sizeof(double2), hipMemcpyHostToDevice, s_));
        OWT_HIP(hipMemsetAsync(mx, 0, sizeof(unsigned long long), s_));
        Chan d = c_;
        d.precision = 0;
        const double2 *in[3] = {wave, wave + W, wave + 2 * W};
        launch_z_inverse(d, in, 3, zp, s_);
        launch_x_c2r(d, zp, 3, phys, s_);
        launch_peak3(phys, P, mx, s_);
        OWT_HIP(hipGetLastError());
        unsigned long long bits = 0;
        OWT_HIP(hipMemcpyAsync(&bits, mx, sizeof(bits), hipMemcpyDeviceToHost, s_));
        OWT_HIP(hipStreamSynchronize(s_));
        double m2;
        std::memcpy(&m2, &bits, sizeof(m2));
        peak = std::sqrt(m2);
    } catch (...) {
        for (void *p : tmp) (void)hipFree(p);
        throw;
    }
    for (void *p : tmp) (void)hipFree(p);
    return peak;
}

void Channel::init_fields(int mode, uint64_t seed, double amp) {
    const int NZ = c_.NZ, NXH = c_.NXH, NY = c_.NY, NZH = c_.NZH;
    const size_t W = c_.W;
    std::vector<double> u(2 * W, 0.0), v(2 * W, 0.0), w(2 * W, 0.0);
    auto at = [&](int k, int i) { return ((size_t)k * NXH + i) * NY; };
    std::vector<double> kx(NXH), kz(NZ);
    for (int i = 0; i < NXH; ++i) kx[i] = (double)i * P_.dkx;
    for (int k = 0; k < NZ; ++k) kz[k] = ((double)k >= NZH ? (double)k - NZ : (double)k) * P_.dkz;
    const std::vector<double> &y = op_.y;
    auto laminar = [&](std::vector<double> &uu, bool add) {
        for (int j = 0; j < NY; ++j) {
            const double U = 1.5 * (1 - std::pow(y[j], 2));
            uu[2 * j] = add ? uu[2 * j] + U : U;
            if (!add) uu[2 * j + 1] = 0.0;
        }
    };
    if (mode == 1) {
        // laminar profile plus the fixed wave at (kx, kz) = (2, +-2) (src/solver.cpp:336-377)
        laminar(u, false);
        const double k1 = 2.0, k2 = 2.0, al = 1.0;
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NXH; ++i) {
                const bool plus = kz[k] == k2 && kx[i] == k1, minus = kz[k] == -k2 && kx[i] == k1;
                if (!plus && !minus) continue;
                for (int j = 0; j < NY; ++j) {
                    const double yy = y[j];
                    const size_t o = 2 * (at(k, i) + j);
                    const double fu = al / k1 * (yy - std::pow(yy, 3));
                    const double fv = al * (0.5 * std::pow(yy, 2) - 0.25 * std::pow(yy, 4) - 0.25) * (0.25);
                    const double fw = -2 * al / k2 * (yy - std::pow(yy, 3));
                    // f * (-0.25 i) = (f * -0, f * -0.25); the real part adds a signed zero
                    u[o] += fu * -0.0;
                    u[o + 1] += fu * -0.25;
                    v[o] += fv;
                    w[o] += fw * (plus ? -0.0 : 0.0);
                    w[o + 1] += fw * (plus ? -0.25 : 0.25);
                }
            }
    } else if (mode == 2) {
        // seeded divergence-free perturbation (recipe defined by this build)
        std::vector<double> g(2 * W, 0.0);
        for (int k = 0; k < NZ; ++k) {
            if (k == NZH) continue;
            const int m = k < NZH ? k : k - NZ;
            for (int i = 0; i < NXH; ++i) {
                const int r2 = i * i + m * m;
                if (r2 < 1 || r2 > 64) continue;
                if (i == 0 && m < 0) continue;
                double cr[4], ci[4], dr[4], di[4];
                for (int p = 0; p < 4; ++p) {
                    seeded_gauss(seed, 0, i, m, p, cr[p], ci[p]);
                    seeded_gauss(seed, 1, i, m, p, dr[p], di[p]);
                }
                const size_t o = at(k, i), oc = (i == 0) ? at(NZ - m, 0) : 0;
                for (int j = 0; j < NY; ++j) {
                    const double yy = y[j];
                    const double T[4] = {1.0, yy, 2.0 * yy * yy - 1.0, (4.0 * yy * yy - 3.0) * yy};
                    const double w1 = 1.0 - yy * yy;
                    // complex * real products summed left to right as the oracle's std::complex does
                    double svr = ((cr[0] * T[0] + cr[1] * T[1]) + cr[2] * T[2]) + cr[3] * T[3];
                    double svi = ((ci[0] * T[0] + ci[1] * T[1]) + ci[2] * T[2]) + ci[3] * T[3];
                    double sgr = ((dr[0] * T[0] + dr[1] * T[1]) + dr[2] * T[2]) + dr[3] * T[3];
                    double sgi = ((di[0] * T[0] + di[1] * T[1]) + di[2] * T[2]) + di[3] * T[3];
                    const double ww = w1 * w1;
                    v[2 * (o + j)] = ww * svr;
                    v[2 * (o + j) + 1] = ww * svi;
                    g[2 * (o + j)] = w1 * sgr;
                    g[2 * (o + j) + 1] = w1 * sgi;
                    if (i == 0) {
                        v[2 * (oc + j)] = ww * svr;
                        v[2 * (oc + j) + 1] = -(ww * svi);
                        g[2 * (oc + j)] = w1 * sgr;
                        g[2 * (oc + j) + 1] = -(w1 * sgi);
                    }
                }
            }
        }
        // u, w from v, omega_y with the discrete DY1 (src/solver.cpp:532-533)
        std::vector<double> d1r(NY), d1i(NY);
        for (int k = 0; k < NZ; ++k)
            for (int i = 0; i < NXH; ++i) {
                const double k2 = std::pow(kx[i], 2) + std::pow(kz[k], 2);
                if (k2 < 1.e-8) continue;
                const size_t o = at(k, i);
                // only the pencils that received modes are non-zero
                const int m = k < NZH ? k : k - NZ;
                if (i * i + m * m > 64) continue;
                std::fill(d1r.begin(), d1r.end(), 0.0);
                std::fill(d1i.begin(), d1i.end(), 0.0);
                for (int j = 0; j < NY; ++j) {
                    const double xr = v[2 * (o + j)], xi = v[2 * (o + j) + 1];
                    const double *col = op_.DY1.data() + (size_t)j * NY;
                    for (int r = 0; r < NY; ++r) {
                        d1r[r] = d1r[r] + col[r] * xr;
                        d1i[r] = d1i[r] + col[r] * xi;
                    }
                }
                const double ix = kx[i] / k2, iz = kz[k] / k2;
                for (int j = 0; j < NY; ++j) {
                    const double gr = g[2 * (o + j)], gi = g[2 * (o + j) + 1];
                    // (0, ix) * d1 - (0, iz) * g ;  (0, ix) * g + (0, iz) * d1
                    u[2 * (o + j)] = (0.0 * d1r[j] - ix * d1i[j]) - (0.0 * gr - iz * gi);
                    u[2 * (o + j) + 1] = (0.0 * d1i[j] + ix * d1r[j]) - (0.0 * gi + iz * gr);
                    w[2 * (o + j)] = (0.0 * gr - ix * gi) + (0.0 * d1r[j] - iz * d1i[j]);
                    w[2 * (o + j) + 1] = (0.0 * gi + ix * gr) + (0.0 * d1i[j] + iz * d1r[j]);
                }
            }
        const double peak = peak_velocity_dev(u, v, w);
        if (peak > 0.0) {
            const double f = amp / peak;
            for (size_t e = 0; e < 2 * W; ++e) {
                u[e] *= f;
                v[e] *= f;
                w[e] *= f;
            }
        }
        laminar(u, true);
    } else if (mode == 3) {
        // the laminar profile alone (the README's initial_condition 1)
        laminar(u, false);
    } else {
        throw Error(2, "init mode must be 1 (reference), 2 (seeded) or 3 (laminar)");
    }
    set_velocity(u.data(), v.data(), w.data());
}

void Channel::set_velocity(const double *u, const double *v, const double *w) {
    const size_t W = c_.W;
    int ctrl[16] = {0};
    OWT_HIP(hipMemcpyAsync(c_.ctrl, ctrl, sizeof(ctrl), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync(c_.cur, u, W * sizeof(double2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync(c_.cur + W, w, W * sizeof(double2), hipMemcpyHostToDevice, s_));
    OWT_HIP(hipMemcpyAsync(c_.hist + (size_t)(H_V * 3) * W, v, W * sizeof(double2), hipMemcpyHostToDevice, s_));
    host_steps_ = 0;
    ring_ = 0;
    tangent_.reset();
    init_post();
}

// src/solver.cpp:379-425
void Channel::init_post() {
    cur_stale_ = false;  // the initial u, w, omega_x, omega_z are computed into cur
    const size_t W = c_.W;
    launch_vor_from_vel(c_, s_);
    launch_init_d2v(c_, s_);
    const double2 *six[6] = {c_.cur, c_.hist + (size_t)(H_V * 3) * W, c_.cur + W, c_.cur + 2 * W, c_.hist + (size_t)(H_G * 3) * W,
                             c_.cur + 3 * W};
    launch_z_inverse(c_, six, 6, c_.zbuf, s_);
    if (tangent_base_) launch_x_tangent(c_, (const double2 *)tangent_base_->c_.zbuf, s_);
    else launch_x_lamb(c_, s_);
    launch_z_forward(c_, s_);
    // Hv, Hg of the initial field land in slot 1 (the new slot of ring phase 0);
    // in the reference's arithmetic for either method
    launch_nonlinear(c_, step_ptrs(0), true, s_);
    launch_init_history(c_, s_);
    launch_init_mean(c_, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
}

// ---------------------------------------------------------------------------
// time step
// ---------------------------------------------------------------------------
StepPtrs Channel::step_ptrs(int ring) const {
    StepPtrs sp{};
    const size_t W = c_.W;
    auto slot = [&](int f, int sl) { return c_.hist + (size_t)(f * 3 + sl) * W; };
    for (int l = 0; l < 3; ++l) {
        const int sl = (ring - l + 3) % 3;
        sp.s[l] = sl;
        sp.v[l] = slot(H_V, sl);
        sp.g[l] = slot(H_G, sl);
        sp.hv[l] = slot(H_HV, sl);
        sp.hg[l] = slot(H_HG, sl);
        sp.dv[l] = c_.d2v + (size_t)sl * W;
    }
    sp.sn = (ring + 1) % 3;
    sp.vn = slot(H_V, sp.sn);
    sp.gn = slot(H_G, sp.sn);
    sp.hvn = slot(H_HV, sp.sn);
    sp.hgn = slot(H_HG, sp.sn);
    sp.dvn = c_.d2v + (size_t)sp.sn * W;
    return sp;
}

// time_Integration (src/solver.cpp:541-627) for the ring phase `ring`
void Channel::enqueue_step_body(int ring) {
    const bool plain = P_.method == 0;
    const StepPtrs sp = step_ptrs(ring);
    const double *F = plain ? Fp_ : Fb_;
    const int *P = plain ? Pp_ : Pb_;
    if (plain) factor_set(Fp_, Pp_, c_.npencil, nullptr, 0, 2);
    // The mean flow (one block: pencil 0's u, w, omega_x, omega_z, the mean
    // histories and the friction) and the v solves read and write nothing of the
    // omega_y solves: they run on a side stream forked from the step's stream (a
    // parallel branch of the captured graph), the mean first, and join before the
    // horizontal stage. The omega_y solves go first on the step's stream: their
    // 512-thread blocks take the CUs as the step starts and the v solves' one-wave
    // blocks fill the slots they free (the other way round, the omega_y blocks
    // were starved by the v solves: 0.57 ms alone, 2.4 ms beside them).
    if (!plain && bops_ == 1) {
        OWT_HIP(hipEventRecord(ev_fork_, s_));
        OWT_HIP(hipStreamWaitEvent(side_, ev_fork_, 0));
        launch_mean(c_, sp, Fmean_, Pmean_, side_);
        launch_solve_inv(c_, sp, Xb_, jobs_, njobs_, side_);
        OWT_HIP(hipEventRecord(ev_join_, side_));
        launch_helmholtz_g(c_, sp, s_);
        OWT_HIP(hipStreamWaitEvent(s_, ev_join_, 0));
        launch_horizontal(c_, sp, plain, s_, f32_six() ? six32_ : nullptr);
    } else {
        launch_solve(c_, sp, F, P, jobs_, plain ? c_.npencil : njobs_, plain, plain ? 2 : bops_, s_);
        launch_horizontal(c_, sp, plain, s_);
        // plain: pencil 0's own factors hold Ag(0) as its second operator
        launch_mean(c_, sp, plain ? Fp_ + (size_t)c_.NY * c_.NY : Fmean_, plain ? Pp_ + c_.NY : Pmean_, s_);
    }
    // the six transformed fields: u, v, w, omega_x, omega_y, omega_z of the new level
    const size_t W = c_.W;
    const double2 *six[6] = {c_.cur, sp.vn, c_.cur + W, c_.cur + 2 * W, sp.gn, c_.cur + 3 * W};
    if (f32_six()) launch_z_inverse_f32(c_, six32_, s_);
    else launch_z_inverse(c_, six, 6, c_.zbuf, s_);
    if (tangent_base_) launch_x_tangent(c_, (const double2 *)tangent_base_->c_.zbuf, s_);
    else launch_x_lamb(c_, s_);
    launch_z_forward(c_, s_);
    launch_nonlinear(c_, sp, plain, s_);
}

EnsEntry Channel::ens_entry(int ring) const {
    EnsEntry e{};
    e.c = c_;
    e.sp = step_ptrs(ring);
    const size_t W = c_.W;
    const double2 *six[6] = {c_.cur, e.sp.vn, c_.cur + W, c_.cur + 2 * W, e.sp.gn, c_.cur + 3 * W};
    for (int f = 0; f < 6; ++f) e.six[f] = six[f];
    e.X = Xb_;
    e.jobs = jobs_;
    e.njobs = njobs_;
    e.Fg = Fmean_;
    e.Pg = Pmean_;
    return e;
}

void Channel::build_graph(int ring) {
    OWT_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_step_body(ring);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OWT_HIP(hipStreamEndCapture(s_, &graph_[ring]));
    OWT_HIP(hipGraphInstantiate(&gexec_[ring], graph_[ring], nullptr, nullptr, 0));
}

// The mixed batched step (mirror kernels, NY <= 256, no tangent attached): the
// horizontal stage writes the z inverse's inputs in FP32 instead of cur.
bool Channel::f32_six() const { return P_.method == 1 && bops_ == 1 && use_mirror(c_) && !tangent_base_; }

// u, w, omega_x, omega_z of the current level in FP64 (cur), after steps that
// left them to the FP32 copies: the same mirror-half kernel, writing cur; the
// stored wall-normal term it also writes goes to the free transform buffer
void Channel::refresh_cur() {
    if (!cur_stale_) return;
    StepPtrs sp = step_ptrs(ring_);
    sp.vn = const_cast<double2 *>(sp.v[0]);
    sp.gn = const_cast<double2 *>(sp.g[0]);
    sp.dvn = (double2 *)c_.zbuf;
    launch_horizontal(c_, sp, false, s_, nullptr);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    cur_stale_ = false;
}

void Channel::enqueue_step() {
    cur_stale_ = f32_six();  // a step that writes cur leaves it current
    if (P_.method == 1) {
        factor_operators();
        if (!gexec_[ring_]) build_graph(ring_);
        OWT_HIP(hipGraphLaunch(gexec_[ring_], s_));
    } else {
        ensure_plain_factors();
        enqueue_step_body(ring_);
        OWT_HIP(hipGetLastError());
    }
    ring_ = (ring_ + 1) % 3;
    ++host_steps_;
}

void Channel::enqueue_kernel(int id) {
    const bool plain = P_.method == 0;
    if (plain) ensure_plain_factors();
    else factor_operators();
    const StepPtrs sp = step_ptrs(ring_);
    const double *F = plain ? Fp_ : Fb_;
    const int *P = plain ? Pp_ : Pb_;
    const size_t W = c_.W;
    const double2 *six[6] = {c_.cur, sp.vn, c_.cur + W, c_.cur + 2 * W, sp.gn, c_.cur + 3 * W};
    switch (id) {
        case 0:
            if (!plain && bops_ == 1) launch_solve_inv(c_, sp, Xb_, jobs_, njobs_, s_);
            else launch_solve(c_, sp, F, P, jobs_, plain ? c_.npencil : njobs_, plain, plain ? 2 : bops_, s_);
            break;
        case 1: launch_horizontal(c_, sp, plain, s_, f32_six() ? six32_ : nullptr); break;
        case 2:
            launch_mean(c_, sp, plain ? Fp_ + (size_t)c_.NY * c_.NY : Fmean_, plain ? Pp_ + c_.NY : Pmean_, s_);
            break;
        case 3:
            if (f32_six()) launch_z_inverse_f32(c_, six32_, s_);
            else launch_z_inverse(c_, six, 6, c_.zbuf, s_);
            break;
        case 4: launch_x_lamb(c_, s_); break;
        case 5: launch_z_forward(c_, s_); break;
        case 6: launch_nonlinear(c_, sp, plain, s_); break;
        case 7:
            if (plain) {
                factor_set(Fp_, Pp_, c_.npencil, nullptr, 0, 2);
            } else {
                factor_set(Fb_, Pb_, nfac_, frep_, 0, bops_);
                if (bops_ == 1) invert_operators();
            }
            break;
        case 8:
            if (!plain && bops_ == 1) launch_helmholtz_g(c_, sp, s_);
            break;
        default: throw Error(2, "kernel id must be 0..8");
    }
    OWT_HIP(hipGetLastError());
}

void Channel::sync() { OWT_HIP(hipStreamSynchronize(s_)); }

static double elapsed_ms(hipEvent_t a, hipEvent_t b) {
    OWT_HIP(hipEventSynchronize(b));
    float t = 0;
    OWT_HIP(hipEventElapsedTime(&t, a, b));
    return t;
}

double Channel::time_steps(int n) {
    if (n < 1) throw Error(2, "time_steps: step count must be positive");
    if (tangent_) throw Error(2, "time_steps: not with a tangent attached (owt_step steps both)");
    for (int k = 0; k < 3; ++k) enqueue_step();
    sync();
    hipEvent_t a = nullptr, b = nullptr;
    OWT_HIP(hipEventCreate(&a));
    OWT_HIP(hipEventCreate(&b));
    double ms = 0;
    try {
        OWT_HIP(hipEventRecord(a, s_));
        for (int k = 0; k < n; ++k) enqueue_step();
        OWT_HIP(hipEventRecord(b, s_));
        ms = elapsed_ms(a, b);
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    finish();
    return ms;
}

double Channel::time_kernel(int id, int reps) {
    if (reps < 1) throw Error(2, "time_kernel: reps must be positive");
    if (tangent_) throw Error(2, "time_kernel: not with a tangent attached");
    prepare_step();
    save_state(timing_);
    enqueue_kernel(id);  // warm launch (and, batched, the one-time factorization)
    sync();
    hipEvent_t a = nullptr, b = nullptr;
    OWT_HIP(hipEventCreate(&a));
    OWT_HIP(hipEventCreate(&b));
    double ms = 0;
    try {
        OWT_HIP(hipEventRecord(a, s_));
        for (int r = 0; r < reps; ++r) enqueue_kernel(id);
        OWT_HIP(hipEventRecord(b, s_));
        ms = elapsed_ms(a, b) / reps;
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        load_state(timing_);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    load_state(timing_);
    finish();
    return ms;
}

void Channel::prepare_step() {
    if (P_.method == 1) factor_operators();
    else ensure_plain_factors();
}

void Channel::finish() {
    sync();
    check_errors();
}

void Channel::step(int n) {
    if (n < 0) throw Error(2, "step count must be non-negative");
    for (int k = 0; k < n; ++k) {
        enqueue_step();
        if (tangent_) tangent_->enqueue_step();
        if (verbose_) {
            sync();
            double fx, b, u;
            wall_shear(&fx, &b, &u);
            std::printf("fx_fric, %g, bottom, %g, upper, %g\n", fx, b, u);
        }
    }
    sync();
    check_errors();
}

long Channel::steps_done() {
    sync();
    int ctrl[2];
    OWT_HIP(hipMemcpy(ctrl, c_.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    return ctrl[C_STEP];
}

// ---------------------------------------------------------------------------
// access
// ---------------------------------------------------------------------------
void Channel::get_field(int f, double *out) {
    if (f < 0 || f >= F_COUNT) throw Error(2, "unknown field id " + std::to_string(f));
    sync();
    refresh_cur();
    if (f == F_P) {
        pressure(out);
        return;
    }
    const int r = ring_;
    const size_t W = c_.W;
    const double2 *src = nullptr;
    switch (f) {
        case F_U: src = c_.cur; break;
        case F_V: src = c_.hist + (size_t)(H_V * 3 + r) * W; break;
        case F_W: src = c_.cur + W; break;
        case F_OMX: src = c_.cur + 2 * W; break;
        case F_OMY: src = c_.hist + (size_t)(H_G * 3 + r) * W; break;
        case F_OMZ: src = c_.cur + 3 * W; break;
        case F_HV: src = c_.hist + (size_t)(H_HV * 3 + r) * W; break;
        case F_HG: src = c_.hist + (size_t)(H_HG * 3 + r) * W; break;
        default: break;
    }
    if (src) {
        OWT_HIP(hipMemcpy(out, src, W * sizeof(double2), hipMemcpyDeviceToHost));
        return;
    }
    const int q = f - F_LMBX;
    if (c_.precision == 0) {
        OWT_HIP(hipMemcpy(out, (const double2 *)c_.lmb + q * W, W * sizeof(double2), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(2 * W);
        OWT_HIP(hipMemcpy(tmp.data(), (const float2 *)c_.lmb + q * W, W * sizeof(float2), hipMemcpyDeviceToHost));
        for (size_t e = 0; e < 2 * W; ++e) out[e] = tmp[e];
    }
}

void Channel::wall_shear(double *fx, double *bottom, double *upper) {
    sync();
    int ctrl[2];
    OWT_HIP(hipMemcpy(ctrl, c_.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    double d[3];
    OWT_HIP(hipMemcpy(d, c_.diag + (size_t)(ctrl[C_STEP] % c_.diag_cap) * 3, sizeof(d), hipMemcpyDeviceToHost));
    *fx = d[0];
    *bottom = d[1];
    *upper = d[2];
}

int Channel::diag(double *out, int max_rows) {
    sync();
    const long n = steps_done() + 1;
    long rows = n < max_rows ? n : max_rows;
    if (rows > c_.diag_cap) rows = c_.diag_cap;
    const long first = n - rows;
    for (long r = 0; r < rows; ++r) {
        const long idx = (first + r) % c_.diag_cap;
        OWT_HIP(hipMemcpy(out + 3 * r, c_.diag + idx * 3, 3 * sizeof(double), hipMemcpyDeviceToHost));
    }
    return (int)rows;
}

// ---------------------------------------------------------------------------
// stand-alone transforms (FP64)
// ---------------------------------------------------------------------------
void Channel::f2p(const double *wave, int nf, double *phys) {
    if (nf < 1 || nf > 6) throw Error(2, "f2p takes 1..6 fields");
    const size_t W = c_.W, I = c_.I, P = (size_t)c_.NZ2 * c_.NX2 * c_.NY;
    double2 *dw = nullptr, *dz = nullptr;
    double *dp = nullptr;
    OWT_HIP(hipMalloc(&dw, nf * W * sizeof(double2)));
    OWT_HIP(hipMalloc(&dz, nf * I * sizeof(double2)));
    OWT_HIP(hipMalloc(&dp, nf * P * sizeof(double)));
    OWT_HIP(hipMemcpy(dw, wave, nf * W * sizeof(double2), hipMemcpyHostToDevice));
    Chan d = c_;
    d.precision = 0;
    const double2 *in[6];
    for (int f = 0; f < nf; ++f) in[f] = dw + f * W;
    launch_z_inverse(d, in, nf, dz, s_);
    launch_x_c2r(d, dz, nf, dp, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    OWT_HIP(hipMemcpy(phys, dp, nf * P * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(dw);
    (void)hipFree(dz);
    (void)hipFree(dp);
}

// f2p on device buffers with the intermediate in the scratch array: no
// allocation. Every element of the intermediate that the c2r transform reads is
// written by the z inverse first, so the scratch needs no clearing; calls that
// overlap on one GPU share the array and race.
void Channel::f2p_noalloc(const double *wave_dev, int nf, double *phys_dev) {
    if (nf < 1 || nf > 6) throw Error(2, "f2p takes 1..6 fields");
    if (!wave_dev || !phys_dev) throw Error(2, "f2p_noalloc: null device buffer");
    const size_t W = c_.W, need = (size_t)nf * c_.I * sizeof(double2);
    if (need > noalloc_scratch_bytes())
        throw Error(2, "f2p_noalloc: " + std::to_string(nf) + " fields need " + std::to_string(need) +
                           " bytes of scratch; the limit is " + std::to_string(noalloc_scratch_bytes()));
    OWT_HIP(hipSetDevice(device_));
    Chan d = c_;
    d.precision = 0;
    const double2 *in[6];
    for (int f = 0; f < nf; ++f) in[f] = reinterpret_cast<const double2 *>(wave_dev) + f * W;
    launch_z_inverse(d, in, nf, noalloc_, s_);
    launch_x_c2r(d, noalloc_, nf, phys_dev, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
}

void Channel::p2f(const double *phys, int nf, double *wave) {
    if (nf < 1 || nf > 6) throw Error(2, "p2f takes 1..6 fields");
    const size_t W = c_.W, I = c_.I, P = (size_t)c_.NZ2 * c_.NX2 * c_.NY;
    double2 *dw = nullptr, *dz = nullptr;
    double *dp = nullptr;
    OWT_HIP(hipMalloc(&dw, nf * W * sizeof(double2)));
    OWT_HIP(hipMalloc(&dz, nf * I * sizeof(double2)));
    OWT_HIP(hipMalloc(&dp, nf * P * sizeof(double)));
    OWT_HIP(hipMemcpy(dp, phys, nf * P * sizeof(double), hipMemcpyHostToDevice));
    Chan d = c_;
    d.precision = 0;
    launch_x_r2c(d, dp, nf, dz, s_);
    launch_z_forward_to(d, dz, nf, dw, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    OWT_HIP(hipMemcpy(wave, dw, nf * W * sizeof(double2), hipMemcpyDeviceToHost));
    (void)hipFree(dw);
    (void)hipFree(dz);
    (void)hipFree(dp);
}

// Lamb vector of six given wave fields through the step's own transform kernels
void Channel::lamb(const double *six, double *three) {
    const size_t W = c_.W;
    double2 *dw = nullptr;
    OWT_HIP(hipMalloc(&dw, 6 * W * sizeof(double2)));
    OWT_HIP(hipMemcpy(dw, six, 6 * W * sizeof(double2), hipMemcpyHostToDevice));
    const double2 *in[6];
    for (int f = 0; f < 6; ++f) in[f] = dw + f * W;
    launch_z_inverse(c_, in, 6, c_.zbuf, s_);
    launch_x_lamb(c_, s_);
    launch_z_forward(c_, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipStreamSynchronize(s_));
    (void)hipFree(dw);
    for (int q = 0; q < 3; ++q) get_field(F_LMBX + q, three + 2 * W * q);
}

// ---------------------------------------------------------------------------
// full-state checkpoint: every history level, the mean-flow and friction
// history and the step counter (the reference's snapshots hold u, v, w only)
// ---------------------------------------------------------------------------
namespace {
struct CkptHeader {
    char magic[8];
    int32_t version, NX, NY, NZ, precision, pad;
    double nu, dt, dkx, dkz, alpha;
};
}  // namespace

// save_instant_field (src/IO.cpp:198-283): <dir>/F-<id_step, 8 digits>.H5. The
// dataset families are the reference's, enabled (velocity) and commented out
// (the rest) there: wave fields as real and imaginary [NZ][NXH][NY] doubles
// (WRITE_3D, src/IO.cpp:118-157), physical fields as [NZ2][NX2][NY] doubles
// (WRITE_3D_dealiased, src/IO.cpp:158-196), named <field>_real / _imag / _p.
void Channel::save_fields(const std::string &dir, int what) {
    if (what <= 0 || what >= 64) throw Error(2, "save: dataset mask must be in 1..63");
    const long id = steps_done();
    char name[40];
    std::snprintf(name, sizeof name, "F-%08ld.H5", id);
    const std::string path = dir.empty() ? std::string(name) : dir + "/" + name;
    const size_t W = c_.W, P = (size_t)c_.NZ2 * c_.NX2 * c_.NY;
    const long dims[3] = {c_.NZ, c_.NXH, c_.NY}, pdims[3] = {c_.NZ2, c_.NX2, c_.NY};
    std::vector<double> z(2 * W), re(W), im(W), ph;
    H5Writer f(path);
    struct Set { int id; const char *name; };
    const Set fam[3][3] = {{{F_U, "velx"}, {F_V, "vely"}, {F_W, "velz"}},
                           {{F_OMX, "vorx"}, {F_OMY, "vory"}, {F_OMZ, "vorz"}},
                           {{F_LMBX, "lmbx"}, {F_LMBY, "lmby"}, {F_LMBZ, "lmbz"}}};
    for (int g = 0; g < 3; ++g) {
        const bool wave = (what >> g) & 1, phys = (what >> (3 + g)) & 1;
        if (!wave && !phys) continue;
        for (const Set &s : fam[g]) {
            get_field(s.id, z.data());
            if (wave) {
                for (size_t e = 0; e < W; ++e) {
                    re[e] = z[2 * e];
                    im[e] = z[2 * e + 1];
                }
                f.write(std::string(s.name) + "_real", dims, re.data());
                f.write(std::string(s.name) + "_imag", dims, im.data());
            }
            if (phys) {
                ph.resize(P);
                f2p(z.data(), 1, ph.data());
                f.write(std::string(s.name) + "_p", pdims, ph.data());
            }
        }
    }
    f.close();
}

void Channel::save(const std::string &dir) { save_fields(dir, 1); }

// Plane statistics of the current level (k_statistics) and the friction Reynolds
// number Re_tau = u_tau h / nu, u_tau^2 = fx_fric h, h = Ly / 2 (the friction
// force balances the shear of both walls)
void Channel::statistics(double *prof, double *re_tau) {
    sync();
    refresh_cur();
    sync();
    double *d = (double *)c_.zbuf;
    launch_statistics(c_, c_.hist + (size_t)(H_V * 3 + ring_) * c_.W, d, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipMemcpyAsync(prof, d, 9 * (size_t)c_.NY * sizeof(double), hipMemcpyDeviceToHost, s_));
    OWT_HIP(hipStreamSynchronize(s_));
    double fx, b, u;
    wall_shear(&fx, &b, &u);
    const double h = c_.Ly / 2.0;
    *re_tau = std::sqrt(std::fabs(fx) * h) * h / c_.nu;
}

void Channel::checkpoint(const std::string &path) {
    sync();
    refresh_cur();
    sync();
    std::ofstream f(path, std::ios::binary);
    if (!f) throw Error(5, "cannot open " + path + " for writing");
    CkptHeader h{};
    std::memcpy(h.magic, "OWTCKPT2", 8);
    h.version = 2;  // 2: the stored wall-normal terms in their precision's form (d2v_store)
    h.NX = c_.NX;
    h.NY = c_.NY;
    h.NZ = c_.NZ;
    h.precision = c_.precision;
    h.nu = P_.nu;
    h.dt = P_.dt;
    h.dkx = P_.dkx;
    h.dkz = P_.dkz;
    h.alpha = P_.alpha_mesh;
    f.write((const char *)&h, sizeof(h));
    auto dump = [&](const void *d, size_t bytes) {
        std::vector<char> buf(bytes);
        OWT_HIP(hipMemcpy(buf.data(), d, bytes, hipMemcpyDeviceToHost));
        f.write(buf.data(), (std::streamsize)bytes);
    };
    const size_t W = c_.W, N = c_.NY;
    dump(c_.hist, 12 * W * sizeof(double2));
    dump(c_.d2v, 3 * W * sizeof(double2));
    dump(c_.cur, 4 * W * sizeof(double2));
    dump(c_.lmb, 3 * W * (c_.precision == 0 ? sizeof(double2) : sizeof(float2)));
    dump(c_.means, 12 * N * sizeof(double2));
    dump(c_.fric, 8 * sizeof(double));
    int ctrl[16];
    OWT_HIP(hipMemcpy(ctrl, c_.ctrl, sizeof(ctrl), hipMemcpyDeviceToHost));
    ctrl[C_RING] = ring_;
    f.write((const char *)ctrl, sizeof(ctrl));
    dump(c_.diag, (size_t)c_.diag_cap * 3 * sizeof(double));
    if (!f) throw Error(5, "write failed: " + path);
}

void Channel::restore(const std::string &path) {
    sync();
    tangent_.reset();
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(5, "cannot open " + path);
    CkptHeader h{};
    f.read((char *)&h, sizeof(h));
    if (!f || std::memcmp(h.magic, "OWTCKPT2", 8) != 0) throw Error(5, path + " is not a checkpoint");
    if (h.NX != c_.NX || h.NY != c_.NY || h.NZ != c_.NZ || h.nu != P_.nu || h.dt != P_.dt || h.dkx != P_.dkx || h.dkz != P_.dkz ||
        h.alpha != P_.alpha_mesh)
        throw Error(5, "checkpoint " + path + " was written for a different case");
    if (h.precision != c_.precision) set_precision(h.precision);
    auto load = [&](void *d, size_t bytes) {
        std::vector<char> buf(bytes);
        f.read(buf.data(), (std::streamsize)bytes);
        if (!f) throw Error(5, "truncated checkpoint " + path);
        OWT_HIP(hipMemcpy(d, buf.data(), bytes, hipMemcpyHostToDevice));
    };
    const size_t W = c_.W, N = c_.NY;
    load(c_.hist, 12 * W * sizeof(double2));
    load(c_.d2v, 3 * W * sizeof(double2));
    if (h.version < 2 && h.precision == 1) {
        // written when every precision stored DY2 v
        launch_d2v_convert(c_, 1, s_);
        OWT_HIP(hipGetLastError());
        OWT_HIP(hipStreamSynchronize(s_));
    }
    load(c_.cur, 4 * W * sizeof(double2));
    cur_stale_ = false;
    load(c_.lmb, 3 * W * (c_.precision == 0 ? sizeof(double2) : sizeof(float2)));
    load(c_.means, 12 * N * sizeof(double2));
    load(c_.fric, 8 * sizeof(double));
    int ctrl[16];
    f.read((char *)ctrl, sizeof(ctrl));
    if (!f || ctrl[C_RING] < 0 || ctrl[C_RING] > 2) throw Error(5, "corrupt checkpoint " + path);
    ring_ = ctrl[C_RING];
    ctrl[C_RING] = 0;
    OWT_HIP(hipMemcpy(c_.ctrl, ctrl, sizeof(ctrl), hipMemcpyHostToDevice));
    load(c_.diag, (size_t)c_.diag_cap * 3 * sizeof(double));
}

// ---------------------------------------------------------------------------
// tangent and adjoint
// ---------------------------------------------------------------------------
void Channel::save_state(State &st) {
    const size_t W = c_.W, N = c_.NY;
    if (!st.hist) {
        st.hist = (double2 *)dmalloc(12 * W * sizeof(double2));
        st.d2v = (double2 *)dmalloc(3 * W * sizeof(double2));
        st.cur = (double2 *)dmalloc(4 * W * sizeof(double2));
        st.lmb = (double2 *)dmalloc(3 * W * sizeof(double2));
        st.means = (double2 *)dmalloc(12 * N * sizeof(double2));
        st.fric = (double *)dmalloc(8 * sizeof(double));
        st.ctrl = (int *)dmalloc(16 * sizeof(int));
    }
    const auto d2d = hipMemcpyDeviceToDevice;
    OWT_HIP(hipMemcpyAsync(st.hist, c_.hist, 12 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.d2v, c_.d2v, 3 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.cur, c_.cur, 4 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.lmb, c_.lmb, 3 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.means, c_.means, 12 * N * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.fric, c_.fric, 8 * sizeof(double), d2d, s_));
    OWT_HIP(hipMemcpyAsync(st.ctrl, c_.ctrl, 16 * sizeof(int), d2d, s_));
    st.ring = ring_;
    st.host_steps = host_steps_;
    st.cur_stale = cur_stale_;
}

void Channel::load_state(const State &st) {
    const size_t W = c_.W, N = c_.NY;
    const auto d2d = hipMemcpyDeviceToDevice;
    OWT_HIP(hipMemcpyAsync(c_.hist, st.hist, 12 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.d2v, st.d2v, 3 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.cur, st.cur, 4 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.lmb, st.lmb, 3 * W * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.means, st.means, 12 * N * sizeof(double2), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.fric, st.fric, 8 * sizeof(double), d2d, s_));
    OWT_HIP(hipMemcpyAsync(c_.ctrl, st.ctrl, 16 * sizeof(int), d2d, s_));
    ring_ = st.ring;
    host_steps_ = st.host_steps;
    cur_stale_ = st.cur_stale;
}

void Channel::zero_state() {
    const size_t W = c_.W, N = c_.NY;
    cur_stale_ = false;
    OWT_HIP(hipMemsetAsync(c_.hist, 0, 12 * W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c_.d2v, 0, 3 * W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c_.cur, 0, 4 * W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c_.lmb, 0, 3 * W * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c_.means, 0, 12 * N * sizeof(double2), s_));
    OWT_HIP(hipMemsetAsync(c_.fric, 0, 8 * sizeof(double), s_));
}

// the six fields of the level in `slot` (cur holds that level's u, w, omega_x,
// omega_z) transformed in z into zbuf, as the step leaves them
void Channel::six_to_zbuf(int slot) {
    const size_t W = c_.W;
    const double2 *six[6] = {c_.cur, c_.hist + (size_t)(H_V * 3 + slot) * W, c_.cur + W, c_.cur + 2 * W,
                             c_.hist + (size_t)(H_G * 3 + slot) * W, c_.cur + 3 * W};
    launch_z_inverse(c_, six, 6, c_.zbuf, s_);
}

void Channel::tangent_init(const double *du, const double *dv, const double *dw) {
    if (c_.precision != 0) throw Error(2, "the tangent needs an FP64 channel");
    if (steps_done() != 0) throw Error(2, "the tangent starts from an initial field (no steps taken)");
    if (!tangent_) {
        Params q = P_;
        q.precision = 0;
        q.method = 1;
        tangent_ = std::make_unique<Channel>(q, device_, false, s_);
        tangent_->tangent_base_ = this;
    }
    six_to_zbuf(ring_);
    tangent_->set_velocity(du, dv, dw);
    check_errors();
}

// the adjoint state shares this channel's stream; its factor set is turned into
// the transposed form once
void Channel::adjoint_prepare() {
    if (transposed_) return;
    factor_operators();
    launch_factor_transpose(Fb_, 2 * nfac_, c_.NY, s_);
    OWT_HIP(hipGetLastError());
    transposed_ = true;
}

// transpose of the Lamb vector evaluation: adjoint lmb -> adjoints of the six
// fields (u, w, omega_x, omega_z in cur; v, omega_y in the history slot), about
// the base fields in base_zbuf; the adjoint lmb is consumed
void Channel::adjoint_lamb(const double2 *base_zbuf, int slot) {
    const size_t W = c_.W, I = c_.I;
    double2 *lm = (double2 *)c_.lmb, *zb = (double2 *)c_.zbuf, *xb = (double2 *)c_.xbuf;
    const double2 *in3[3] = {lm, lm + W, lm + 2 * W};
    launch_z_forward_T(c_, in3, 3, xb, s_);
    launch_x_adjoint(c_, base_zbuf, xb, zb, s_);
    double2 *hv = c_.hist + (size_t)(H_V * 3 + slot) * W, *hg = c_.hist + (size_t)(H_G * 3 + slot) * W;
    double2 *dst[2][3] = {{c_.cur, hv, c_.cur + W}, {c_.cur + 2 * W, hg, c_.cur + 3 * W}};
    for (int half = 0; half < 2; ++half) {
        launch_z_inverse_T(c_, zb + (size_t)half * 3 * I, 3, xb, s_);
        double2 *src[3] = {xb, xb + W, xb + 2 * W};
        launch_accumulate(dst[half], src, 3, W, s_);
    }
    OWT_HIP(hipMemsetAsync(c_.lmb, 0, 3 * W * sizeof(double2), s_));
}

// transpose of one step at ring phase `ring` (reverse stage order)
void Channel::adjoint_step(int ring, const double2 *base_zbuf, double seed, const double *Xv) {
    const StepPtrs sp = step_ptrs(ring);
    launch_nonlinear_adj(c_, sp, s_);
    adjoint_lamb(base_zbuf, sp.sn);
    launch_mean_adj(c_, sp, Fb_, Pb_, fmap_host_[0], seed, s_);
    launch_horizontal_adj(c_, sp, s_);
    launch_solve_adj(c_, sp, Fb_, Pb_, Xv, jobs_, njobs_, s_);
    OWT_HIP(hipGetLastError());
}

// transpose of the initial post-processing, down to the initial u, v, w
void Channel::adjoint_init(const double2 *base_zbuf, double *gu, double *gv, double *gw) {
    const size_t W = c_.W;
    launch_init_adj_mean_history(c_, s_);
    launch_nonlinear_adj(c_, step_ptrs(0), s_);
    adjoint_lamb(base_zbuf, 0);
    double2 *g = (double2 *)c_.zbuf;
    launch_init_vel_adj(c_, g, g + W, g + 2 * W, s_);
    OWT_HIP(hipGetLastError());
    OWT_HIP(hipMemcpyAsync(gu, g, W * sizeof(double2), hipMemcpyDeviceToHost, s_));
    OWT_HIP(hipMemcpyAsync(gv, g + W, W * sizeof(double2), hipMemcpyDeviceToHost, s_));
    OWT_HIP(hipMemcpyAsync(gw, g + 2 * W, W * sizeof(double2), hipMemcpyDeviceToHost, s_));
    OWT_HIP(hipStreamSynchronize(s_));
}

double Channel::adjoint_run(int K, int ckpt, int first, int last, double *gu, double *gv, double *gw, const double *tu,
                            const double *tv, const double *tw) {
    const bool terminal = tu && tv && tw;
    if (terminal) first = last = K;
    if (c_.precision != 0) throw Error(2, "the adjoint needs an FP64 channel");
    if (K < 1 || ckpt < 1 || first < 1 || first > last || last > K)
        throw Error(2, "adjoint: need K >= 1, ckpt >= 1 and 1 <= first <= last <= K");
    if (steps_done() != 0) throw Error(2, "the adjoint starts from an initial field (no steps taken)");
    if (K >= c_.diag_cap) throw Error(2, "adjoint: K exceeds the friction history");
    tangent_.reset();
    // forward sweep with checkpoints at steps 0, ckpt, 2 ckpt, ...
    const int nck = (K - 1) / ckpt + 1;
    if ((int)ckpts_.size() < nck) ckpts_.resize(nck);
    for (int n = 0; n < K; ++n) {
        if (n % ckpt == 0) save_state(ckpts_[n / ckpt]);
        enqueue_step();
    }
    sync();
    check_errors();
    std::vector<double> d((size_t)K * 3 + 3);
    OWT_HIP(hipMemcpy(d.data(), c_.diag, d.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double wgt = 1.0 / (double)(last - first + 1);
    double J = 0.0;
    if (!terminal) {
        for (int m = first; m <= last; ++m) J += d[(size_t)m * 3];
        J *= wgt;
    }
    if (!adjoint_) {
        Params q = P_;
        q.precision = 0;
        q.method = 1;
        adjoint_ = std::make_unique<Channel>(q, device_, false, s_);
        adjoint_->bops_ = 2;  // the transposed solves use LU factors of Av and Ag
    }
    Channel &A = *adjoint_;
    A.adjoint_prepare();
    A.zero_state();
    if (terminal) {
        // the velocity after step K: u and w in cur, v in the history slot of level K
        const size_t W = c_.W;
        const StepPtrs spK = A.step_ptrs((K - 1) % 3);
        OWT_HIP(hipMemcpyAsync(A.c_.cur, tu, W * sizeof(double2), hipMemcpyHostToDevice, s_));
        OWT_HIP(hipMemcpyAsync(A.c_.cur + W, tw, W * sizeof(double2), hipMemcpyHostToDevice, s_));
        OWT_HIP(hipMemcpyAsync(spK.vn, tv, W * sizeof(double2), hipMemcpyHostToDevice, s_));
    }
    // the transposed v solves use the forward step's explicit inverses (batched method)
    const double *Xv = P_.method == 1 ? Xb_ : nullptr;
    // reverse sweep: step n -> n+1 needs the base's six fields of level n+1
    for (int n = K - 1; n >= 0; --n) {
        const int c0 = (n / ckpt) * ckpt;
        load_state(ckpts_[n / ckpt]);
        for (int m = c0; m <= n; ++m) enqueue_step();
        const double seed = (!terminal && n + 1 >= first && n + 1 <= last) ? wgt : 0.0;
        A.adjoint_step(n % 3, (const double2 *)c_.zbuf, seed, Xv);
    }
    load_state(ckpts_[0]);
    six_to_zbuf(0);
    A.adjoint_init((const double2 *)c_.zbuf, gu, gv, gw);
    // leave the channel after step K
    load_state(ckpts_[(K - 1) / ckpt]);
    for (int m = ((K - 1) / ckpt) * ckpt; m < K; ++m) enqueue_step();
    sync();
    check_errors();
    A.check_errors();
    return J;
}

// ---------------------------------------------------------------------------
// launch-shape autotuner
// ---------------------------------------------------------------------------
void Channel::get_tune(int *knobs) const {
    for (int k = 0; k < T_COUNT; ++k) knobs[k] = c_.tune[k];
}

void Channel::set_tune(const int *knobs) {
    for (int k = 0; k < T_COUNT; ++k)
        if (knobs[k] < 0 || knobs[k] >= kTuneCandidates) throw Error(2, "tune candidate out of range");
    sync();
    for (int k = 0; k < T_COUNT; ++k) c_.tune[k] = knobs[k];
    drop_graph();
}

std::string Channel::autotune(int reps) {
    if (reps < 1) throw Error(2, "autotune: reps must be positive");
    static const char *names[T_COUNT] = {"solve_jobs_per_block", "horizontal_pencils_per_block", "nonlinear_pencils_per_block",
                                         "z_inverse_tile", "x_tile", "z_forward_tile"};
    static const char *shapes[T_COUNT][kTuneCandidates] = {
        {"2", "1", "4", ""}, {"mfma16", "", "", ""}, {"mfma16", "", "", ""}, {"default", "alt1", "alt2", "alt3"},
        {"default", "alt1", "alt2", "alt3"}, {"default", "alt1", "alt2", "alt3"}};
    // stage kernel of each knob (enqueue_kernel ids)
    static const int kernel[T_COUNT] = {0, 1, 6, 3, 4, 5};
    if (tangent_) throw Error(2, "autotune: not with a tangent attached");
    prepare_step();
    // the candidates run on the state, which they change (a stage writes the
    // new-level slots, which hold history level 2 between steps): saved here and
    // restored at the end
    save_state(timing_);
    sync();
    hipEvent_t a, b;
    OWT_HIP(hipEventCreate(&a));
    OWT_HIP(hipEventCreate(&b));
    std::string rep = "{";
    for (int k = 0; k < T_COUNT; ++k) {
        // transform tiles: three candidates per length class (two above 576), and in
        // FP32 a fourth for lengths above 192; the other knobs have three
        int ncand = 3;
        if (k == T_Z_INVERSE_TILE || k == T_Z_FORWARD_TILE || k == T_X_TILE) {
            const int L = k == T_X_TILE ? c_.NX2 : c_.NZ2;
            if (c_.precision == 1 && L > 192) ncand = 4;
            else if (L > 576) ncand = 2;
        }
        if (k == T_SOLVE_JOBS && P_.method == 0) ncand = 1;  // the plain solve has one shape
        // the derivative stages run on the matrix-core kernels, which have one shape
        if (k == T_HORIZONTAL_TP || k == T_NONLINEAR_TP) ncand = 1;
        // the batched v solves share the step's start with the omega_y solves
        // (parallel branches): their candidates are timed by whole steps
        const bool whole_step = k == T_SOLVE_JOBS && P_.method == 1 && bops_ == 1;
        auto run = [&]() {
            if (whole_step) enqueue_step_body(ring_);
            else enqueue_kernel(kernel[k]);
        };
        double best = 1e300;
        int arg = 0;
        std::string times;
        for (int cand = 0; cand < ncand; ++cand) {
            c_.tune[k] = cand;
            run();
            OWT_HIP(hipEventRecord(a, s_));
            for (int r = 0; r < reps; ++r) run();
            OWT_HIP(hipEventRecord(b, s_));
            OWT_HIP(hipEventSynchronize(b));
            float t = 0;
            OWT_HIP(hipEventElapsedTime(&t, a, b));
            const double ms = t / reps;
            char buf[64];
            std::snprintf(buf, sizeof buf, "%s\"%s\": %.5f", cand ? ", " : "", shapes[k][cand], ms);
            times += buf;
            if (ms < best) {
                best = ms;
                arg = cand;
            }
        }
        c_.tune[k] = arg;
        char buf[160];
        std::snprintf(buf, sizeof buf, "%s\"%s\": {\"choice\": %d, \"ms\": {", k ? ", " : "", names[k], arg);
        rep += buf + times + "}}";
    }
    rep += "}";
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    load_state(timing_);
    sync();
    check_errors();
    drop_graph();
    return rep;
}

}  // namespace owt
