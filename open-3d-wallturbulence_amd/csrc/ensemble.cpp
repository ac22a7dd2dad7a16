// Ensemble of independent channels (see ensemble.hpp).
#include "ensemble.hpp"

#include <cstring>
#include <map>

namespace owt {

// members whose operators are the same bit for bit
static bool same_operators(const Params &a, const Params &b) {
    return a.NX == b.NX && a.NY == b.NY && a.NZ == b.NZ && a.nu == b.nu && a.dt == b.dt && a.dkx == b.dkx &&
           a.dkz == b.dkz && a.alpha_mesh == b.alpha_mesh && a.precision == b.precision;
}

Ensemble::Ensemble(const std::vector<Params> &ps, int device) : device_(device) {
    if (ps.empty()) throw Error(2, "an ensemble needs at least one member");
    OWT_HIP(hipSetDevice(device));
    try {
        OWT_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        OWT_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        for (const Params &p : ps) m_.push_back(std::make_unique<Channel>(p, device, false));
    } catch (...) {
        m_.clear();
        if (fork_) (void)hipEventDestroy(fork_);
        if (s_) (void)hipStreamDestroy(s_);
        throw;
    }
}

Ensemble::~Ensemble() {
    (void)hipSetDevice(device_);
    if (s_) (void)hipStreamSynchronize(s_);
    drop_graphs();
    for (hipStream_t st : lane_s_) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (hipEvent_t ev : lane_ev_) (void)hipEventDestroy(ev);
    if (fork_) (void)hipEventDestroy(fork_);
    if (s_) (void)hipStreamDestroy(s_);
    m_.clear();
}

void Ensemble::drop_graphs() {
    for (int r = 0; r < 3; ++r) {
        if (gexec_[r]) (void)hipGraphExecDestroy(gexec_[r]);
        if (graph_[r]) (void)hipGraphDestroy(graph_[r]);
        if (table_[r]) (void)hipFree(table_[r]);
        gexec_[r] = nullptr;
        graph_[r] = nullptr;
        table_[r] = nullptr;
        launches_[r].clear();
    }
    gen_.clear();
}

// Members with the same parameters run the same kernel instances on the same
// sizes: the first of them is tuned (Channel::autotune) and the others take its
// choices, so that they stay in one lane.
void Ensemble::autotune(int reps) {
    OWT_HIP(hipSetDevice(device_));
    for (size_t i = 0; i < m_.size(); ++i) {
        size_t j = 0;
        while (j < i && !same_operators(m_[j]->params(), m_[i]->params())) ++j;
        if (j == i) {
            (void)m_[i]->autotune(reps);
        } else {
            int knobs[T_COUNT];
            m_[j]->get_tune(knobs);
            m_[i]->set_tune(knobs);
        }
    }
}

void Ensemble::init_fields(int mode, const uint64_t *seeds, double amp) {
    for (int i = 0; i < size(); ++i) m_[i]->init_fields(mode, seeds ? seeds[i] : (uint64_t)i, amp);
}

// The grouped step applies when every member runs the batched step with the
// omega_y eigendecomposition and nothing attached, and all members are at the
// same ring phase (they are, unless a member was stepped or restored on its
// own). Tables and graphs made before a member changed its kernels or buffers
// (precision, method) are dropped.
bool Ensemble::grouped_ready() {
    const int r = m_[0]->ring();
    for (auto &m : m_) {
        if (m->method() != 1 || m->ring() != r) return false;
        m->prepare_step();
        if (!m->ens_ready()) return false;
    }
    bool stale = gen_.size() != m_.size();
    for (size_t i = 0; !stale && i < m_.size(); ++i) stale = gen_[i] != m_[i]->generation();
    if (stale) {
        drop_graphs();
        for (auto &m : m_) gen_.push_back(m->generation());
    }
    return true;
}

// Member entries at ring phase `ring`, grouped per stage by kernel instance, in
// one device table: for each launch its entries, then for each launch the
// first block of each of its members.
void Ensemble::build_tables(int ring) {
    const int n = size();
    std::vector<EnsEntry> ent(n);
    for (int i = 0; i < n; ++i) {
        ent[i] = m_[i]->ens_entry(ring);
        for (int j = 0; j < i; ++j) {
            if (!same_operators(m_[j]->params(), m_[i]->params())) continue;
            // read the earlier member's copy of the same operators
            const EnsEntry &o = ent[j];
            Chan &c = ent[i].c;
            c.DY1 = o.c.DY1;
            c.DY2 = o.c.DY2;
            c.DY4 = o.c.DY4;
            c.DY1m = o.c.DY1m;
            c.DY2m = o.c.DY2m;
            c.hmir = o.c.hmir;
            c.Xm = o.c.Xm;
            c.kx = o.c.kx;
            c.kz = o.c.kz;
            c.qw = o.c.qw;
            c.twx = o.c.twx;
            c.twz = o.c.twz;
            c.twxf = o.c.twxf;
            c.twzf = o.c.twzf;
            c.hQ = o.c.hQ;
            c.hQi = o.c.hQi;
            c.hlre = o.c.hlre;
            c.hlim = o.c.hlim;
            ent[i].X = o.X;
            ent[i].jobs = o.jobs;
            ent[i].njobs = o.njobs;
            ent[i].Fg = o.Fg;
            ent[i].Pg = o.Pg;
            break;
        }
    }
    // lanes: members whose stages all run the same kernel instances, in order of
    // first appearance; each lane is one launch per stage
    std::vector<std::vector<int>> lanes;
    {
        std::map<std::vector<int>, int> lane_of;
        for (int i = 0; i < n; ++i) {
            std::vector<int> keys(ES_COUNT);
            for (int st = 0; st < ES_COUNT; ++st) keys[st] = ens_key(ent[i], st);
            auto it = lane_of.find(keys);
            if (it == lane_of.end()) {
                it = lane_of.emplace(keys, (int)lanes.size()).first;
                lanes.emplace_back();
            }
            lanes[it->second].push_back(i);
        }
    }
    nlanes_[ring] = (int)lanes.size();
    std::vector<Launch> &L = launches_[ring];
    L.clear();
    std::vector<EnsEntry> tab_e;
    std::vector<std::vector<int>> tab_f;
    for (int ln = 0; ln < (int)lanes.size(); ++ln) {
        const std::vector<int> &g = lanes[ln];
        for (int st = 0; st < ES_COUNT; ++st) {
            Launch l{};
            l.stage = st;
            l.lane = ln;
            l.nm = (int)g.size();
            l.rep = ent[g[0]];
            l.ents = tab_e.size();
            std::vector<int> first;
            int grid = 0, ny = 0;
            for (int i : g) {
                tab_e.push_back(ent[i]);
                first.push_back(grid);
                grid += ens_blocks(ent[i], st);
                ny = ent[i].c.NY > ny ? ent[i].c.NY : ny;
            }
            first.push_back(grid);
            l.grid = grid;
            l.ny_max = ny;
            l.first = tab_f.size();
            tab_f.push_back(first);
            L.push_back(l);
        }
    }
    // byte offsets: entries first, then the first-block arrays
    const size_t eb = tab_e.size() * sizeof(EnsEntry);
    std::vector<size_t> foff(tab_f.size());
    size_t fb = 0;
    for (size_t i = 0; i < tab_f.size(); ++i) {
        foff[i] = eb + fb;
        fb += tab_f[i].size() * sizeof(int);
    }
    std::vector<unsigned char> host(eb + fb);
    std::memcpy(host.data(), tab_e.data(), eb);
    for (size_t i = 0; i < tab_f.size(); ++i) std::memcpy(host.data() + foff[i], tab_f[i].data(), tab_f[i].size() * sizeof(int));
    for (Launch &l : L) {
        l.ents *= sizeof(EnsEntry);
        l.first = foff[l.first];
    }
    OWT_HIP(hipMalloc(&table_[ring], host.size()));
    OWT_HIP(hipMemcpy(table_[ring], host.data(), host.size(), hipMemcpyHostToDevice));
}

// Each lane's launches run on a stream of its own, forked from the ensemble's
// stream and joined back: the lanes are parallel branches of the graph, so one
// lane's bandwidth-bound stages overlap another's matrix-core stages.
void Ensemble::build_graph(int ring) {
    if (!table_[ring]) build_tables(ring);
    const unsigned char *t = (const unsigned char *)table_[ring];
    const int nl = nlanes_[ring];
    while ((int)lane_s_.size() < nl) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        OWT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        lane_s_.push_back(st);
        OWT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        lane_ev_.push_back(ev);
    }
    OWT_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        OWT_HIP(hipEventRecord(fork_, s_));
        for (int ln = 0; ln < nl; ++ln) OWT_HIP(hipStreamWaitEvent(lane_s_[ln], fork_, 0));
        for (const Launch &l : launches_[ring])
            launch_ens(l.stage, (const EnsEntry *)(t + l.ents), (const int *)(t + l.first), l.nm, l.grid, l.rep, l.ny_max,
                       lane_s_[l.lane]);
        for (int ln = 0; ln < nl; ++ln) {
            OWT_HIP(hipEventRecord(lane_ev_[ln], lane_s_[ln]));
            OWT_HIP(hipStreamWaitEvent(s_, lane_ev_[ln], 0));
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    OWT_HIP(hipStreamEndCapture(s_, &graph_[ring]));
    OWT_HIP(hipGraphInstantiate(&gexec_[ring], graph_[ring], nullptr, nullptr, 0));
}

void Ensemble::enqueue_step() {
    if (!grouped_ready()) {
        // members one by one, each on its own stream, then joined
        for (auto &m : m_) m->enqueue_step();
        for (auto &m : m_) m->sync();
        return;
    }
    const int r = m_[0]->ring();
    if (!gexec_[r]) build_graph(r);
    OWT_HIP(hipGraphLaunch(gexec_[r], s_));
    for (auto &m : m_) m->advance_ring();
}

int Ensemble::launches_per_step() {
    OWT_HIP(hipSetDevice(device_));
    for (auto &m : m_) m->sync();
    if (!grouped_ready()) return 0;
    const int r = m_[0]->ring();
    if (!table_[r]) build_tables(r);
    return (int)launches_[r].size();
}

void Ensemble::step(int n) {
    if (n < 0) throw Error(2, "step count must be non-negative");
    OWT_HIP(hipSetDevice(device_));
    // members' own queued work (initialisation, transfers) completes first
    for (auto &m : m_) m->sync();
    for (int k = 0; k < n; ++k) enqueue_step();
    OWT_HIP(hipStreamSynchronize(s_));
    for (auto &m : m_) m->finish();
}

double Ensemble::time_steps(int n) {
    if (n < 1) throw Error(2, "step count must be positive");
    // one step per ring phase builds the graphs outside the timed region
    step(3);
    hipEvent_t a = nullptr, b = nullptr;
    OWT_HIP(hipEventCreate(&a));
    OWT_HIP(hipEventCreate(&b));
    OWT_HIP(hipEventRecord(a, s_));
    for (int k = 0; k < n; ++k) enqueue_step();
    OWT_HIP(hipEventRecord(b, s_));
    OWT_HIP(hipEventSynchronize(b));
    float t = 0;
    OWT_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (auto &m : m_) m->finish();
    return t;
}

}  // namespace owt
