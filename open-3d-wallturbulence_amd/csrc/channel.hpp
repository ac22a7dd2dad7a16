// One channel resident on one GPU: the drop-in for the reference's global
// solver state (src/setup.hpp:21-83) and its setup/step call sequence
// (src/main.cpp:63-87).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "device.hpp"
#include "owt_internal.hpp"

namespace owt {

// field ids of the C ABI (owt.h)
enum FieldId { F_U = 0, F_V, F_W, F_OMX, F_OMY, F_OMZ, F_LMBX, F_LMBY, F_LMBZ, F_HV, F_HG, F_P, F_COUNT };

class Channel {
  public:
    // stream: a stream to share (the channel does not own it), or null for its own
    Channel(const Params &p, int device, bool verbose, hipStream_t stream = nullptr);
    ~Channel();
    Channel(const Channel &) = delete;
    Channel &operator=(const Channel &) = delete;

    const Params &params() const { return P_; }
    const Operators &ops() const { return op_; }
    const Chan &dev() const { return c_; }
    hipStream_t stream() const { return s_; }

    // init_fields (src/solver.cpp:329-426); mode 1 reference wave, 2 seeded recipe
    void init_fields(int mode, uint64_t seed, double amp);
    // set u, v, w ([NZ][NXH][NY] complex interleaved) and run the init post-processing
    void set_velocity(const double *u, const double *v, const double *w);
    // n calls of time_Integration (src/solver.cpp:541-627)
    void step(int n);
    void get_field(int f, double *out);
    void wall_shear(double *fx, double *bottom, double *upper);
    int diag(double *out, int max_rows);
    long steps_done();
    void set_method(int method);
    void set_precision(int precision);
    // save_instant_field: HDF5 snapshot <dir>/F-<id_step>.H5 of u, v, w; save_fields
    // with a dataset mask (1 velocity, 2 vorticity, 4 Lamb vector in wave space;
    // 8, 16, 32 the same in physical space)
    void save(const std::string &dir);
    void save_fields(const std::string &dir, int what);
    // plane statistics: prof[9][NY] = U, V, W, <u'u'>, <v'v'>, <w'w'>, <u'v'>, <u'w'>, <v'w'>
    void statistics(double *prof, double *re_tau);
    void checkpoint(const std::string &path);
    void restore(const std::string &path);

    // stand-alone transforms, FP64 (stage-level parity)
    void f2p(const double *wave, int nf, double *phys);
    // the same on caller-owned device buffers, without allocating: the z-padded
    // intermediate lives in the device's scratch array (noalloc.hip)
    void f2p_noalloc(const double *wave_dev, int nf, double *phys_dev);
    void p2f(const double *phys, int nf, double *wave);
    void lamb(const double *six, double *three);

    // Timing (HIP events on the channel's stream). time_steps runs 3 steps (one
    // per ring phase, which builds the graphs) and then n timed steps: the state
    // advances by n + 3 steps. time_kernel launches one stage of the step reps
    // times at the current ring phase (after one untimed launch) on the state,
    // which it saves before and restores after. Both refuse a channel with an
    // attached tangent (whose step reads this channel's transform buffer) and
    // report errors like step(). Milliseconds: total for time_steps, per launch
    // for time_kernel.
    double time_steps(int n);
    double time_kernel(int id, int reps);
    // device-side timing helpers (bench): run one step on the stream without syncing
    void enqueue_step();
    // one launch of a single stage of the step at the current ring phase:
    // 0 solve, 1 horizontal, 2 mean, 3 z inverse, 4 x phase, 5 z forward,
    // 6 nonlinear, 7 factorization, 8 omega_y solves by the eigendecomposition
    void enqueue_kernel(int id);
    // Launch-shape autotuner: times every candidate of every knob (TuneKnob) on
    // this channel's data (saved before and restored after: fields, histories
    // and counters are unchanged), keeps the fastest, and recaptures the step
    // graphs. Returns a JSON report; get/set_tune read or force the choices.
    // Refused with a tangent attached.
    std::string autotune(int reps);
    void get_tune(int *knobs) const;
    void set_tune(const int *knobs);
    // stored factor of one +-kz system (batched factor set)
    void get_factor(int sys, int op, double *lu, int *perm);
    void sync();

    // used by Ensemble, which captures the step bodies of several channels into
    // one graph: the one-time preparation of a step (factors), the body of one
    // step at the current ring phase on this channel's stream, the ring advance
    // after such a body ran, and a counter that changes whenever this channel's
    // captured kernels would change (precision, method)
    void prepare_step();
    int ring() const { return ring_; }
    int method() const { return P_.method; }
    void enqueue_body() { enqueue_step_body(ring_); }
    void advance_ring() {
        ring_ = (ring_ + 1) % 3;
        ++host_steps_;
        cur_stale_ = false;  // the grouped launches write u, w, omega_x, omega_z into cur
    }
    unsigned generation() const { return generation_; }
    void finish();
    // grouped launches (Ensemble): whether this channel's step is the batched
    // one with the omega_y eigendecomposition and nothing attached (tangent),
    // and its member entry at ring phase `ring` (after prepare_step)
    bool ens_ready() const { return P_.method == 1 && bops_ == 1 && factored_ && !tangent_ && !tangent_base_; }
    EnsEntry ens_entry(int ring) const;

    // Tangent (forward-mode linearization, FP64). tangent_init attaches a tangent
    // state started from the direction (du, dv, dw) through the linearized
    // initial post-processing; the channel must hold an initial field. step()
    // then advances both; tangent() is the tangent state as a channel.
    void tangent_init(const double *du, const double *dv, const double *dw);
    Channel *tangent() { return tangent_.get(); }
    // Discrete adjoint (FP64) of K steps from the initial field held now, for
    // J = mean of the friction force fx over steps first..last: forward with a
    // checkpoint every ckpt steps, then the transposed steps in reverse, each
    // base state recomputed from its checkpoint. Returns J; grad_* get dJ/d(u,
    // v, w) of the initial field ([NZ][NXH][NY] complex, d/dRe + i d/dIm). The
    // channel is left after step K.
    // objective J = mean fx_fric over steps first..last; or, with a terminal seed
    // (tu, tv, tw non-null), the adjoint of the map initial velocity -> velocity
    // after K steps applied to the seed (J = 0)
    double adjoint_run(int K, int ckpt, int first, int last, double *gu, double *gv, double *gw, const double *tu = nullptr,
                       const double *tv = nullptr, const double *tw = nullptr);

  private:
    // device copy of the whole evolving state (in-memory checkpoint)
    struct State {
        double2 *hist = nullptr, *d2v = nullptr, *cur = nullptr, *lmb = nullptr, *means = nullptr;
        double *fric = nullptr;
        int *ctrl = nullptr;
        int ring = 0;
        long host_steps = 0;
        bool cur_stale = false;
    };
    void save_state(State &st);
    void load_state(const State &st);
    void zero_state();
    void six_to_zbuf(int slot);
    // the mixed batched step at NY <= 256 hands the z inverse its six fields in
    // FP32 (six32_) and leaves u, w, omega_x, omega_z of cur as they were
    // (cur_stale_); refresh_cur recomputes them in FP64 for the calls that read them
    bool f32_six() const;
    void refresh_cur();
    void adjoint_prepare();
    void adjoint_lamb(const double2 *base_zbuf, int slot);
    void adjoint_step(int ring, const double2 *base_zbuf, double seed, const double *Xv);
    void adjoint_init(const double2 *base_zbuf, double *gu, double *gv, double *gw);
    void alloc();
    void free_all();
    void upload_operators();
    void factor_operators();
    void ensure_plain_factors();
    void factor_set(double *F, int *P, int nfac, const int *frep, int op0, int nops);
    void invert_operators();
    void pressure(double *out);
    void init_post();
    StepPtrs step_ptrs(int ring) const;
    void enqueue_step_body(int ring);
    void build_graph(int ring);
    void drop_graph();
    void check_errors();
    void *dmalloc(size_t bytes);
    double peak_velocity_dev(const std::vector<double> &u, const std::vector<double> &v, const std::vector<double> &w);

    Params P_;
    Operators op_;
    Chan c_{};
    int device_ = 0;
    double2 *noalloc_ = nullptr;  // this device's scratch array (noalloc.hip)
    bool verbose_ = false;
    hipStream_t s_ = nullptr;
    // side stream and events of the step's parallel branch (omega_y solves and
    // mean flow beside the v solves)
    hipStream_t side_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    std::vector<void *> allocs_;
    // factor sets (lu_kernels.hip): batched, one per +-kz system, computed once;
    // plain, one per pencil, recomputed every step
    double *Fb_ = nullptr, *Fp_ = nullptr;
    int *Pb_ = nullptr, *Pp_ = nullptr;
    // pressure operator factors (batched grouping), computed on first request
    double *Fq_ = nullptr;
    int *Pq_ = nullptr;
    bool factored_ = false;
    // operators per factor index of the batched set: 1 (Av; omega_y by the
    // eigendecomposition of DY2) or 2 (Av, Ag: adjoint channels)
    int bops_ = 1;
    // Ag(0) for the mean flow when bops_ == 1, and the factor the mean-flow kernel uses
    double *Fg0_ = nullptr;
    int *Pg0_ = nullptr, *sys0_ = nullptr;
    const double *Fmean_ = nullptr;
    const int *Pmean_ = nullptr;
    // explicit inverses of the Av factors (bops_ == 1), the step's v solves,
    // and the scratch of their computation (xchunk_ matrices)
    double *Xb_ = nullptr, *xscratch_ = nullptr;
    int xchunk_ = 0;
    // the same inverses in mirror form for the mixed-precision step (NY <= 256),
    // and the rank-2 factors U, V of the odd part of Av they are built from
    double *Xm_ = nullptr, *UV_ = nullptr;
    // all Ag factors, made only when requested (get_factor) with bops_ == 1
    double *Fag_ = nullptr;
    int *Pag_ = nullptr;
    // scratch of the panel factorization (NY > 256): fchunk_ matrices x 2 operators
    double *fscratch_ = nullptr;
    int fchunk_ = 0;
    bool force_panel_ = false;
    bool mirror_solve_ = true;
    // batched factor tables: one factor per distinct k^2; jobs ({factor, system,
    // system or -1}: systems sharing a factor, two at a time), frep (factor -> a
    // system); fmap_host_ (system -> factor)
    int nfac_ = 0, njobs_ = 0;
    int4 *jobs_ = nullptr;
    int *frep_ = nullptr;
    std::vector<int> fmap_host_;
    // history ring phase: slot of the current level; one captured graph per phase
    int ring_ = 0;
    float2 *six32_ = nullptr;
    bool cur_stale_ = false;
    hipGraph_t graph_[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t gexec_[3] = {nullptr, nullptr, nullptr};
    long host_steps_ = 0;
    unsigned generation_ = 0;
    bool own_stream_ = true;
    // tangent: the base this channel linearizes about (its z buffer), or null
    const Channel *tangent_base_ = nullptr;
    std::unique_ptr<Channel> tangent_;
    // adjoint: the adjoint state (its factors hold the transposed form)
    std::unique_ptr<Channel> adjoint_;
    bool transposed_ = false;
    std::vector<State> ckpts_;
    State timing_;  // the state saved around time_kernel and autotune
};

FFTPlan make_plan(int L);

}  // namespace owt
