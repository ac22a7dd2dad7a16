/*
 * owt.h -- C ABI of the MI355X channel-flow time step.
 *
 * Replaces the reference solver's free-function surface (src/setup.hpp:84-114),
 * called by its driver in the order of src/main.cpp:63-87:
 *
 *   load_params + allocate_var + init_mat + init_fft   -> owt_create
 *   init_fields                                         -> owt_init_fields / owt_set_velocity
 *   time_Integration (x NT+1) and its stdout diagnostics -> owt_step, owt_wall_shear, owt_diag
 *   save_instant_field                                  -> owt_save (+ owt_checkpoint for restart)
 *   cleanup_fft + process exit                          -> owt_destroy
 *
 * Fields are exchanged in the reference's wave-space layout [NZ][NXH][NY]
 * (src/setup.hpp:58-69, y fastest) as interleaved complex doubles (re, im).
 * Every function returns 0 on success or a status code (OWT_E_*) and never
 * lets an exception cross the ABI; owt_last_error() gives the message.
 * A handle is one channel on one GPU with its own HIP stream; calls on one
 * handle are not thread-safe, distinct handles are independent.
 */
#ifndef OWT_H
#define OWT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OWT_OK = 0,
    OWT_E_CONFIG = 1,   /* case file missing, key missing or malformed (src/IO.cpp:36-39) */
    OWT_E_ARG = 2,      /* invalid argument or unsupported size */
    OWT_E_DEVICE = 3,   /* HIP runtime error, out of device memory, no device */
    OWT_E_SINGULAR = 4, /* zero pivot in an LU factorization (src/linalg.hpp:641-644) */
    OWT_E_IO = 5        /* file could not be written or read */
};

/* field ids for owt_get_field */
enum {
    OWT_U = 0, OWT_V = 1, OWT_W = 2,           /* velx, vely, velz */
    OWT_OMX = 3, OWT_OMY = 4, OWT_OMZ = 5,     /* vorx, vory, vorz */
    OWT_LMBX = 6, OWT_LMBY = 7, OWT_LMBZ = 8,  /* Lamb vector */
    OWT_HV = 9, OWT_HG = 10,                    /* nonlinear sources of the v and omega_y equations */
    OWT_P = 11  /* dynamic pressure p + |u|^2/2, computed on request (the reference has no pressure,
                   src/setup.hpp:58-79): (DY2 - k^2) P = i kx Lx + DY1 Ly + i kz Lz, wall rows
                   DY1 P = Ly + nu DY2 v, k = 0 pencil gauged to zero mean over y */
};

/* methods and precisions */
enum { OWT_PLAIN = 0, OWT_BATCHED = 1 };
enum { OWT_FP64 = 0, OWT_MIXED = 1 };

typedef struct owt_handle owt_handle;

typedef struct {
    int device;    /* HIP device ordinal */
    int precision; /* -1: from the case file ([owt] precision), else OWT_FP64 / OWT_MIXED */
    int method;    /* -1: from the case file ([owt] method), else OWT_PLAIN / OWT_BATCHED */
    int verbose;   /* print the reference's per-step friction line (src/solver.cpp:314) */
} owt_opts;

/* The reference's toml keys (src/IO.cpp:14-25) for callers without a case file. */
typedef struct {
    double nu, dt, dkx, dkz, alpha_mesh;
    int NX, NY, NZ;
    int precision, method;
} owt_params;

/* load_params + allocate_var + init_mat + init_fft (src/IO.cpp:8-74,
 * src/solver.cpp:15-216, src/FFT.cpp:11-34). opts may be NULL. */
int owt_create(const char *toml_path, const owt_opts *opts, owt_handle **out);
int owt_create_params(const owt_params *params, const owt_opts *opts, owt_handle **out);
/* cleanup_fft + frees (src/FFT.cpp:36-47) */
void owt_destroy(owt_handle *h);
/* message of the last failed call on h (or of the last failed owt_create* when h is NULL) */
const char *owt_last_error(const owt_handle *h);

/* sizes: NX, NY, NZ, NXH, NZH, NX2, NZ2 */
int owt_get_sizes(const owt_handle *h, int *sizes7);

/* init_fields (src/solver.cpp:329-426). mode 1: laminar profile plus the
 * reference's fixed (2, +-2) wave; mode 2: seeded divergence-free perturbation
 * of peak amplitude amp; mode 3: the laminar profile alone (seed, amp used in
 * mode 2 only). */
int owt_init_fields(owt_handle *h, int mode, uint64_t seed, double amp);

/* The run keys of the case file the handle was created from (src/IO.cpp:14-25
 * and the [owt] table): NT and save_interval of the driver loop
 * (src/main.cpp:69-87), and the initial field: init_mode 1..3 as for
 * owt_init_fields with seed and amp, or 4 to restart from the full-state
 * checkpoint `restart` (owt_restore). The reference's initial_condition key
 * (test-cases/Retau-180/case.toml:12) is read: 1 and 2 give mode 1, which is what the
 * reference computes for either (its init_fields never reads the key), 3
 * (renewal) gives mode 4. Handles made by owt_create_params report NT = 0,
 * save_interval = 1, mode 1. */
typedef struct {
    long NT, save_interval;
    int init_mode;
    uint64_t seed;
    double amp;
    char restart[1024];
} owt_run_params;
int owt_get_run_params(const owt_handle *h, owt_run_params *out);
/* injects u, v, w and runs the rest of init_fields (src/solver.cpp:379-425) */
int owt_set_velocity(owt_handle *h, const double *u, const double *v, const double *w);

/* nsteps calls of time_Integration (src/solver.cpp:541-627) */
int owt_step(owt_handle *h, int nsteps);
/* the latest friction line of src/solver.cpp:314: fx_fric, bottom, upper */
int owt_wall_shear(owt_handle *h, double *fx, double *bottom, double *upper);
/* up to max_rows most recent (fx_fric, bottom, upper) rows, oldest first;
 * row 0 is the evaluation of init_fields. Returns the row count (>= 0) or -code. */
int owt_diag(owt_handle *h, double *out, int max_rows);
/* number of completed steps (id_step) or -code */
long owt_steps_done(owt_handle *h);

/* current field in [NZ][NXH][NY] complex interleaved */
int owt_get_field(owt_handle *h, int field, double *out);

int owt_set_method(owt_handle *h, int method);
int owt_set_precision(owt_handle *h, int precision);

/* save_instant_field (src/IO.cpp:198-283): writes <dir>/F-<id_step, 8 digits>.H5
 * with the datasets velx_real, velx_imag, vely_real, vely_imag, velz_real,
 * velz_imag, each [NZ][NXH][NY] double (WRITE_3D, src/IO.cpp:118-157); the
 * reference writes to ./DATA. The serial HDF5 C library (1.10 or later) is
 * loaded at run time from $OWT_HDF5_LIB or libhdf5.so on the loader path;
 * OWT_E_IO when it cannot be loaded or the file cannot be written. */
int owt_save(owt_handle *h, const char *dir);
/* the same file with the dataset families the reference's writer has enabled
 * or commented out (src/IO.cpp:237-278), by mask: OWT_SAVE_VEL (velx_real ...),
 * OWT_SAVE_VOR (vorx_real ...), OWT_SAVE_LAMB (lmbx_real ...) in wave space;
 * OWT_SAVE_VEL_P (velx_p ...), OWT_SAVE_VOR_P, OWT_SAVE_LAMB_P in physical space,
 * [NZ2][NX2][NY] doubles (WRITE_3D_dealiased, src/IO.cpp:158-196).
 * owt_save(h, dir) is owt_save_fields(h, dir, OWT_SAVE_VEL). */
enum { OWT_SAVE_VEL = 1, OWT_SAVE_VOR = 2, OWT_SAVE_LAMB = 4, OWT_SAVE_VEL_P = 8, OWT_SAVE_VOR_P = 16, OWT_SAVE_LAMB_P = 32 };
int owt_save_fields(owt_handle *h, const char *dir, int what);

/* Plane statistics of the current field (no reference counterpart; the
 * reference prints the friction line only): profiles[9][NY] = U, V, W and the
 * plane averages <u'u'>, <v'v'>, <w'w'>, <u'v'>, <u'w'>, <v'w'> of the
 * fluctuations, by Parseval over wave space (the z-Nyquist modes excluded, as
 * the transforms to physical space drop them); re_tau = u_tau h / nu with
 * u_tau^2 = fx_fric h, h the half height. */
int owt_statistics(owt_handle *h, double *profiles, double *re_tau);

/* full state: all history levels, mean-flow and friction history, step counter */
int owt_checkpoint(owt_handle *h, const char *path);
int owt_restore(owt_handle *h, const char *path);

/* Fourier2Physical_dealiased / Physical2Fourier_dealiased (src/FFT.cpp:49-194)
 * on nf fields (FP64); physical layout [NZ2][NX2][NY] */
int owt_f2p(owt_handle *h, const double *wave, int nf, double *phys);
int owt_p2f(owt_handle *h, const double *phys, int nf, double *wave);
/* owt_f2p without any allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). wave_dev and phys_dev
 * are device buffers on h's GPU ([nf][NZ][NXH][NY] complex and
 * [nf][NZ2][NX2][NY] double); the z-padded intermediate uses a scratch array
 * the library holds on each GPU (64 MiB), so nf * NZ2 * NXH * NY * 16 bytes
 * must fit in it, else OWT_E_ARG. Runs on h's stream and returns when done.
 * Calls that overlap on one GPU share the scratch array and race. */
int owt_f2p_noalloc(owt_handle *h, const double *wave_dev, int nf, double *phys_dev);
/* Lamb vector of (u, v, w, omx, omy, omz) through the step's transform kernels
 * (src/solver.cpp:262-291) in the handle's precision; three output fields */
int owt_lamb(owt_handle *h, const double *six, double *three);

/* host-side operators, column-major N x N: 0..3 DY1..DY4 */
int owt_get_matrix(owt_handle *h, int which, double *out);

/* LU factor of the implicit operator op (0 Av, 1 Ag; src/solver.cpp:577-603)
 * of the +-kz system sys = mz*NXH + i (mz in [0, NZH]) as the batched method
 * stores it: lu N x N column-major in dgetrf's layout (unit-lower L and U,
 * rows in pivot order) and perm[k] = original row of pivot k. No reference
 * counterpart (the reference factors inside time_Integration); it exposes the
 * factors for validation. */
int owt_get_factor(owt_handle *h, int sys, int op, double *lu, int *perm);

/* Launch-shape autotuner (no reference counterpart): times every candidate
 * shape of each tunable launch on the handle's own data (reps launches each),
 * keeps the fastest and recaptures the step graphs. Knobs (OWT_TUNE_COUNT ints,
 * candidate index 0..3, 0 = default): solve jobs per block, the horizontal and
 * the nonlinear kernels (one shape each, the matrix-core kernels, so these two
 * slots accept any index and are not timed), tile shape of the z inverse, x
 * and z forward transforms. Every candidate computes the same values in the
 * same order, so results do not depend on the choice. The candidates run on
 * the handle's state, which is saved before and restored after (fields,
 * histories and counters are unchanged). Refused with a tangent attached.
 * report (may be NULL) receives a JSON summary of the timings. */
enum { OWT_TUNE_COUNT = 6 };
int owt_autotune(owt_handle *h, int reps, char *report, int report_len);
int owt_get_tune(owt_handle *h, int *knobs);
int owt_set_tune(owt_handle *h, const int *knobs);

/* Differentiation (FP64; no reference counterpart).
 *
 * Tangent: owt_tangent_init attaches to h a tangent state, the forward-mode
 * linearization of the step about h's trajectory, started from the direction
 * (du, dv, dw) of the initial velocity (wave layout) through the linearized
 * initial post-processing. h must hold an initial field with no steps taken.
 * owt_tangent_step(h, n) (or owt_step) then advances h and its tangent
 * together. owt_tangent_handle(h) is the tangent state as a borrowed handle for
 * owt_get_field, owt_diag (the tangent friction series) and the like; it is
 * valid until h is re-initialised or destroyed. */
int owt_tangent_init(owt_handle *h, const double *du, const double *dv, const double *dw);
owt_handle *owt_tangent_handle(owt_handle *h);
int owt_tangent_step(owt_handle *h, int nsteps);

/* Adjoint: the discrete adjoint (transpose of the tangent) of K steps from the
 * initial field h holds (no steps taken). objective OWT_OBJ_MEAN_FX: J = mean of
 * the friction force fx_fric (the mean wall shear) over steps first..last
 * (1 <= first <= last <= K). The forward sweep keeps a checkpoint of the state
 * every ckpt_every steps in device memory; the reverse sweep recomputes each
 * step's state from its checkpoint. Writes J and dJ/du, dJ/dv, dJ/dw of the
 * initial field ([NZ][NXH][NY] complex: d/dRe + i d/dIm of each entry), and
 * leaves h after step K. */
enum { OWT_OBJ_MEAN_FX = 0 };
int owt_adjoint_run(owt_handle *h, int K, int ckpt_every, int objective, int first, int last, double *J, double *grad_u,
                    double *grad_v, double *grad_w);
/* The same reverse sweep for the linear map L from the initial velocity to the
 * velocity after K steps (about h's trajectory): writes L^T w, w = (wu, wv, ww)
 * given in the wave layout, as grad_u, grad_v, grad_w (same convention as
 * owt_adjoint_run). With the tangent d of owt_tangent_init and its velocity
 * after K steps, L d, the two satisfy <L d, w> = <d, L^T w> with
 * <a, b> = sum Re(a) Re(b) + Im(a) Im(b) over the three fields. Leaves h after
 * step K. */
int owt_adjoint_apply(owt_handle *h, int K, int ckpt_every, const double *wu, const double *wv, const double *ww, double *grad_u,
                      double *grad_v, double *grad_w);

/* Ensembles: n independent channels on one GPU (sizes and parameters may
 * differ), stepped together. No reference counterpart (the reference runs one
 * channel per job). With the batched method each stage of the step is one
 * kernel launch over all members that run the same kernel instance (same
 * rows-per-lane class, transform length class, precision and tuned shape),
 * each block finding its member in a device table; members with equal
 * parameters share one copy of the operators. One graph per step holds these
 * launches, the members of each distinct shape as a branch of their own, so
 * their number grows with the distinct shapes, not with the member count. Member m is a
 * borrowed owt_handle for every single-channel call (fields, diagnostics,
 * checkpoint, save); owt_destroy on it does nothing, it lives until
 * owt_ensemble_destroy. A member stepped on its own leaves the ensemble's step
 * correct (members are then stepped one by one until their phases agree). */
typedef struct owt_ensemble owt_ensemble;
int owt_ensemble_create(int n, const owt_params *params, const owt_opts *opts, owt_ensemble **out);
void owt_ensemble_destroy(owt_ensemble *e);
/* message of the last failed call on e (or of the last failed create when e is NULL) */
const char *owt_ensemble_last_error(const owt_ensemble *e);
owt_handle *owt_ensemble_member(owt_ensemble *e, int m);
/* init_fields of every member, member m with seeds[m] (seeds NULL: seed m) */
int owt_ensemble_init_fields(owt_ensemble *e, int mode, const uint64_t *seeds, double amp);
/* nsteps steps of every member */
int owt_ensemble_step(owt_ensemble *e, int nsteps);
/* nsteps ensemble steps between two HIP events; elapsed milliseconds */
int owt_ensemble_time_steps(owt_ensemble *e, int nsteps, double *ms);
/* kernel launches in one grouped ensemble step (*n = 0 when the members are
 * stepped one by one: another method, or ring phases that differ) */
int owt_ensemble_launches(owt_ensemble *e, int *n);
/* launch-shape autotuner over the ensemble: the first member of each distinct
 * parameter set is tuned as by owt_autotune (reps launches per candidate) and
 * the members with the same parameters take its choices */
int owt_ensemble_autotune(owt_ensemble *e, int reps);

/* timing: 3 untimed steps (one per ring phase: the graphs are built), then
 * nsteps steps on the handle's stream between two HIP events; the elapsed
 * milliseconds. The state advances by nsteps + 3 steps, as by owt_step (inputs
 * stay resident on the device). Refused (OWT_E_ARG) with a tangent attached. */
int owt_time_steps(owt_handle *h, int nsteps, double *ms);
/* timing of one stage kernel launched reps times (after one untimed launch) on
 * the handle's stream at the current ring phase: 0 right-hand sides + implicit
 * solves, 1 horizontal, 2 mean flow, 3 z inverse, 4 x phase, 5 z forward,
 * 6 nonlinear, 7 LU factorization (the plain method refactors every step),
 * 8 the batched method's omega_y solves (nothing in the plain method). The
 * state is saved before and restored after, so fields and friction series are
 * unchanged. Returns the mean milliseconds per launch. Refused with a tangent
 * attached. */
int owt_time_kernel(owt_handle *h, int kernel, int reps, double *ms);

#ifdef __cplusplus
}
#endif

#endif /* OWT_H */
