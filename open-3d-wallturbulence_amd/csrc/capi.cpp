// extern "C" layer: status codes in place of the reference's exceptions and
// aborts; nothing throws across the ABI.
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>

#include "../../include/owt.h"
#include "channel.hpp"
#include "ensemble.hpp"

// A handle owns its channel, or borrows one member of an ensemble (own empty).
struct owt_handle {
    std::unique_ptr<owt::Channel> own;
    owt::Channel *ch = nullptr;
    std::string err;
    bool member = false;
    std::unique_ptr<owt_handle> tangent;  // borrowed view of ch->tangent()
};

struct owt_ensemble {
    std::unique_ptr<owt::Ensemble> e;
    std::vector<std::unique_ptr<owt_handle>> members;
    std::string err;
};

namespace {

std::mutex g_mu;
std::string g_create_err;

template <typename F>
int guard(owt_handle *h, F &&f) {
    if (!h) return OWT_E_ARG;
    try {
        f();
        return OWT_OK;
    } catch (const owt::Error &e) {
        h->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        h->err = "host out of memory";
        return OWT_E_DEVICE;
    } catch (const std::exception &e) {
        h->err = e.what();
        return OWT_E_DEVICE;
    }
}

int create(const owt::Params &p, const owt_opts *opts, owt_handle **out) {
    owt::Params q = p;
    int device = 0;
    bool verbose = false;
    if (opts) {
        device = opts->device;
        verbose = opts->verbose != 0;
        if (opts->precision >= 0) q.precision = opts->precision;
        if (opts->method >= 0) q.method = opts->method;
    }
    if (q.precision < 0 || q.precision > 1 || q.method < 0 || q.method > 1)
        throw owt::Error(OWT_E_ARG, "precision and method must be 0 or 1");
    auto h = std::make_unique<owt_handle>();
    h->own = std::make_unique<owt::Channel>(q, device, verbose);
    h->ch = h->own.get();
    *out = h.release();
    return OWT_OK;
}

owt::Params from_struct(const owt_params &s) {
    owt::Params p;
    p.nu = s.nu;
    p.dt = s.dt;
    p.dkx = s.dkx;
    p.dkz = s.dkz;
    p.alpha_mesh = s.alpha_mesh;
    p.NX = s.NX;
    p.NY = s.NY;
    p.NZ = s.NZ;
    p.precision = s.precision;
    p.method = s.method;
    return p;
}

template <typename F>
int eguard(owt_ensemble *e, F &&f) {
    if (!e) return OWT_E_ARG;
    try {
        f();
        return OWT_OK;
    } catch (const owt::Error &x) {
        e->err = x.what();
        return x.code;
    } catch (const std::bad_alloc &) {
        e->err = "host out of memory";
        return OWT_E_DEVICE;
    } catch (const std::exception &x) {
        e->err = x.what();
        return OWT_E_DEVICE;
    }
}

int create_guarded(const owt::Params *p, const char *path, const owt_opts *opts, owt_handle **out) {
    if (!out) return OWT_E_ARG;
    *out = nullptr;
    try {
        owt::Params q = path ? owt::load_params(path) : *p;
        if (!path) owt::derive_sizes(q);
        return create(q, opts, out);
    } catch (const owt::Error &e) {
        std::lock_guard<std::mutex> l(g_mu);
        g_create_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> l(g_mu);
        g_create_err = e.what();
        return OWT_E_DEVICE;
    }
}

}  // namespace

extern "C" {

int owt_create(const char *toml_path, const owt_opts *opts, owt_handle **out) {
    if (!toml_path) return OWT_E_ARG;
    return create_guarded(nullptr, toml_path, opts, out);
}

int owt_create_params(const owt_params *params, const owt_opts *opts, owt_handle **out) {
    if (!params) return OWT_E_ARG;
    const owt::Params p = from_struct(*params);
    return create_guarded(&p, nullptr, opts, out);
}

void owt_destroy(owt_handle *h) {
    if (!h || h->member) return;  // ensemble members live as long as their ensemble
    try {
        delete h;
    } catch (...) {
    }
}

int owt_autotune(owt_handle *h, int reps, char *report, int report_len) {
    return guard(h, [&] {
        const std::string r = h->ch->autotune(reps);
        if (report && report_len > 0) {
            std::strncpy(report, r.c_str(), (size_t)report_len - 1);
            report[report_len - 1] = 0;
        }
    });
}

int owt_get_tune(owt_handle *h, int *knobs) {
    if (!knobs) return OWT_E_ARG;
    return guard(h, [&] { h->ch->get_tune(knobs); });
}

int owt_set_tune(owt_handle *h, const int *knobs) {
    if (!knobs) return OWT_E_ARG;
    return guard(h, [&] { h->ch->set_tune(knobs); });
}

int owt_tangent_init(owt_handle *h, const double *du, const double *dv, const double *dw) {
    if (!du || !dv || !dw) return OWT_E_ARG;
    return guard(h, [&] { h->ch->tangent_init(du, dv, dw); });
}

owt_handle *owt_tangent_handle(owt_handle *h) {
    if (!h || !h->ch || !h->ch->tangent()) return nullptr;
    if (!h->tangent) {
        h->tangent = std::make_unique<owt_handle>();
        h->tangent->member = true;
    }
    h->tangent->ch = h->ch->tangent();
    return h->tangent.get();
}

int owt_tangent_step(owt_handle *h, int nsteps) {
    return guard(h, [&] {
        if (!h->ch->tangent()) throw owt::Error(OWT_E_ARG, "no tangent attached (owt_tangent_init)");
        h->ch->step(nsteps);
    });
}

int owt_adjoint_run(owt_handle *h, int K, int ckpt_every, int objective, int first, int last, double *J, double *grad_u,
                    double *grad_v, double *grad_w) {
    if (!J || !grad_u || !grad_v || !grad_w) return OWT_E_ARG;
    if (objective != OWT_OBJ_MEAN_FX) return OWT_E_ARG;
    return guard(h, [&] { *J = h->ch->adjoint_run(K, ckpt_every, first, last, grad_u, grad_v, grad_w); });
}

int owt_adjoint_apply(owt_handle *h, int K, int ckpt_every, const double *wu, const double *wv, const double *ww, double *grad_u,
                      double *grad_v, double *grad_w) {
    if (!wu || !wv || !ww || !grad_u || !grad_v || !grad_w) return OWT_E_ARG;
    return guard(h, [&] { h->ch->adjoint_run(K, ckpt_every, K, K, grad_u, grad_v, grad_w, wu, wv, ww); });
}

int owt_ensemble_create(int n, const owt_params *params, const owt_opts *opts, owt_ensemble **out) {
    if (!out) return OWT_E_ARG;
    *out = nullptr;
    if (n < 1 || !params) return OWT_E_ARG;
    try {
        std::vector<owt::Params> ps;
        for (int i = 0; i < n; ++i) {
            owt::Params p = from_struct(params[i]);
            owt::derive_sizes(p);
            if (opts && opts->precision >= 0) p.precision = opts->precision;
            if (opts && opts->method >= 0) p.method = opts->method;
            if (p.precision < 0 || p.precision > 1 || p.method < 0 || p.method > 1)
                throw owt::Error(OWT_E_ARG, "precision and method must be 0 or 1");
            ps.push_back(p);
        }
        auto e = std::make_unique<owt_ensemble>();
        e->e = std::make_unique<owt::Ensemble>(ps, opts ? opts->device : 0);
        for (int i = 0; i < n; ++i) {
            auto h = std::make_unique<owt_handle>();
            h->ch = &e->e->member(i);
            h->member = true;
            e->members.push_back(std::move(h));
        }
        *out = e.release();
        return OWT_OK;
    } catch (const owt::Error &x) {
        std::lock_guard<std::mutex> l(g_mu);
        g_create_err = x.what();
        return x.code;
    } catch (const std::exception &x) {
        std::lock_guard<std::mutex> l(g_mu);
        g_create_err = x.what();
        return OWT_E_DEVICE;
    }
}

void owt_ensemble_destroy(owt_ensemble *e) {
    try {
        delete e;
    } catch (...) {
    }
}

const char *owt_ensemble_last_error(const owt_ensemble *e) {
    if (e) return e->err.c_str();
    std::lock_guard<std::mutex> l(g_mu);
    return g_create_err.c_str();
}

owt_handle *owt_ensemble_member(owt_ensemble *e, int m) {
    if (!e || m < 0 || m >= (int)e->members.size()) return nullptr;
    return e->members[m].get();
}

int owt_ensemble_init_fields(owt_ensemble *e, int mode, const uint64_t *seeds, double amp) {
    return eguard(e, [&] { e->e->init_fields(mode, seeds, amp); });
}

int owt_ensemble_step(owt_ensemble *e, int nsteps) {
    return eguard(e, [&] { e->e->step(nsteps); });
}

int owt_ensemble_time_steps(owt_ensemble *e, int nsteps, double *ms) {
    if (!ms) return OWT_E_ARG;
    return eguard(e, [&] { *ms = e->e->time_steps(nsteps); });
}

int owt_ensemble_autotune(owt_ensemble *e, int reps) {
    return eguard(e, [&] { e->e->autotune(reps); });
}

int owt_ensemble_launches(owt_ensemble *e, int *n) {
    if (!n) return OWT_E_ARG;
    return eguard(e, [&] { *n = e->e->launches_per_step(); });
}

const char *owt_last_error(const owt_handle *h) {
    if (h) return h->err.c_str();
    std::lock_guard<std::mutex> l(g_mu);
    return g_create_err.c_str();
}

int owt_get_sizes(const owt_handle *h, int *s) {
    if (!h || !s) return OWT_E_ARG;
    const owt::Params &p = h->ch->params();
    s[0] = (int)p.NX;
    s[1] = (int)p.NY;
    s[2] = (int)p.NZ;
    s[3] = (int)p.NXH;
    s[4] = (int)p.NZH;
    s[5] = (int)p.NX2;
    s[6] = (int)p.NZ2;
    return OWT_OK;
}

int owt_get_run_params(const owt_handle *h, owt_run_params *out) {
    if (!h || !out) return OWT_E_ARG;
    const owt::Params &p = h->ch->params();
    out->NT = p.NT;
    out->save_interval = p.save_interval;
    out->init_mode = p.init_mode;
    out->seed = p.seed;
    out->amp = p.amp;
    std::memset(out->restart, 0, sizeof out->restart);
    std::strncpy(out->restart, p.restart.c_str(), sizeof out->restart - 1);
    return OWT_OK;
}

int owt_init_fields(owt_handle *h, int mode, uint64_t seed, double amp) {
    return guard(h, [&] { h->ch->init_fields(mode, seed, amp); });
}

int owt_set_velocity(owt_handle *h, const double *u, const double *v, const double *w) {
    if (!u || !v || !w) return OWT_E_ARG;
    return guard(h, [&] { h->ch->set_velocity(u, v, w); });
}

int owt_step(owt_handle *h, int nsteps) { return guard(h, [&] { h->ch->step(nsteps); }); }

int owt_wall_shear(owt_handle *h, double *fx, double *bottom, double *upper) {
    if (!fx || !bottom || !upper) return OWT_E_ARG;
    return guard(h, [&] { h->ch->wall_shear(fx, bottom, upper); });
}

int owt_diag(owt_handle *h, double *out, int max_rows) {
    if (!out || max_rows < 0) return -OWT_E_ARG;
    int rows = 0;
    int rc = guard(h, [&] { rows = h->ch->diag(out, max_rows); });
    return rc ? -rc : rows;
}

long owt_steps_done(owt_handle *h) {
    long n = 0;
    int rc = guard(h, [&] { n = h->ch->steps_done(); });
    return rc ? -rc : n;
}

int owt_get_field(owt_handle *h, int field, double *out) {
    if (!out) return OWT_E_ARG;
    return guard(h, [&] { h->ch->get_field(field, out); });
}

int owt_set_method(owt_handle *h, int method) { return guard(h, [&] { h->ch->set_method(method); }); }

int owt_set_precision(owt_handle *h, int precision) { return guard(h, [&] { h->ch->set_precision(precision); }); }

int owt_save(owt_handle *h, const char *dir) {
    if (!dir) return OWT_E_ARG;
    return guard(h, [&] { h->ch->save(dir); });
}

int owt_save_fields(owt_handle *h, const char *dir, int what) {
    if (!dir) return OWT_E_ARG;
    return guard(h, [&] { h->ch->save_fields(dir, what); });
}

int owt_statistics(owt_handle *h, double *profiles, double *re_tau) {
    if (!profiles || !re_tau) return OWT_E_ARG;
    return guard(h, [&] { h->ch->statistics(profiles, re_tau); });
}

int owt_checkpoint(owt_handle *h, const char *path) {
    if (!path) return OWT_E_ARG;
    return guard(h, [&] { h->ch->checkpoint(path); });
}

int owt_restore(owt_handle *h, const char *path) {
    if (!path) return OWT_E_ARG;
    return guard(h, [&] { h->ch->restore(path); });
}

int owt_f2p(owt_handle *h, const double *wave, int nf, double *phys) {
    if (!wave || !phys) return OWT_E_ARG;
    return guard(h, [&] { h->ch->f2p(wave, nf, phys); });
}

int owt_f2p_noalloc(owt_handle *h, const double *wave_dev, int nf, double *phys_dev) {
    if (!wave_dev || !phys_dev) return OWT_E_ARG;
    return guard(h, [&] { h->ch->f2p_noalloc(wave_dev, nf, phys_dev); });
}

int owt_p2f(owt_handle *h, const double *phys, int nf, double *wave) {
    if (!wave || !phys) return OWT_E_ARG;
    return guard(h, [&] { h->ch->p2f(phys, nf, wave); });
}

int owt_lamb(owt_handle *h, const double *six, double *three) {
    if (!six || !three) return OWT_E_ARG;
    return guard(h, [&] { h->ch->lamb(six, three); });
}

int owt_get_matrix(owt_handle *h, int which, double *out) {
    if (!out) return OWT_E_ARG;
    return guard(h, [&] {
        const owt::Operators &op = h->ch->ops();
        const std::vector<double> *m[] = {&op.DY1, &op.DY2, &op.DY3, &op.DY4};
        if (which < 0 || which > 3) throw owt::Error(OWT_E_ARG, "matrix id must be 0..3");
        std::memcpy(out, m[which]->data(), m[which]->size() * sizeof(double));
    });
}

int owt_get_factor(owt_handle *h, int sys, int op, double *lu, int *perm) {
    if (!lu || !perm) return OWT_E_ARG;
    return guard(h, [&] { h->ch->get_factor(sys, op, lu, perm); });
}

int owt_time_steps(owt_handle *h, int nsteps, double *ms) {
    if (!ms || nsteps < 1) return OWT_E_ARG;
    return guard(h, [&] { *ms = h->ch->time_steps(nsteps); });
}

int owt_time_kernel(owt_handle *h, int kernel, int reps, double *ms) {
    if (!ms || reps < 1) return OWT_E_ARG;
    if (kernel < 0 || kernel > 8) return OWT_E_ARG;
    return guard(h, [&] { *ms = h->ch->time_kernel(kernel, reps); });
}

}  // extern "C"
