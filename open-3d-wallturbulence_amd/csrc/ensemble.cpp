// Ensemble of independent channels (see ensemble.hpp).
#include "ensemble.hpp"

namespace owt {

Ensemble::Ensemble(const std::vector<Params> &ps, int device) : device_(device) {
    if (ps.empty()) throw Error(2, "an ensemble needs at least one member");
    OWT_HIP(hipSetDevice(device));
    try {
        OWT_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
        OWT_HIP(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        for (const Params &p : ps) {
            m_.push_back(std::make_unique<Channel>(p, device, false));
            hipEvent_t e = nullptr;
            OWT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            join_.push_back(e);
        }
    } catch (...) {
        for (hipEvent_t e : join_) (void)hipEventDestroy(e);
        if (fork_) (void)hipEventDestroy(fork_);
        if (s_) (void)hipStreamDestroy(s_);
        throw;
    }
}

Ensemble::~Ensemble() {
    (void)hipSetDevice(device_);
    if (s_) (void)hipStreamSynchronize(s_);
    drop_graphs();
    for (hipEvent_t e : join_) (void)hipEventDestroy(e);
    if (fork_) (void)hipEventDestroy(fork_);
    if (s_) (void)hipStreamDestroy(s_);
    m_.clear();
}

void Ensemble::drop_graphs() {
    for (int r = 0; r < 3; ++r) {
        if (gexec_[r]) (void)hipGraphExecDestroy(gexec_[r]);
        if (graph_[r]) (void)hipGraphDestroy(graph_[r]);
        gexec_[r] = nullptr;
        graph_[r] = nullptr;
    }
    gen_.clear();
}

void Ensemble::init_fields(int mode, const uint64_t *seeds, double amp) {
    for (int i = 0; i < size(); ++i) m_[i]->init_fields(mode, seeds ? seeds[i] : (uint64_t)i, amp);
}

// The captured graph applies when every member uses the batched method and all
// members are at the same ring phase (they are, unless a member was stepped or
// restored on its own). Graphs captured before a member changed its kernels are
// dropped.
bool Ensemble::graph_ready() {
    const int r = m_[0]->ring();
    for (auto &m : m_)
        if (m->method() != 1 || m->ring() != r) return false;
    bool stale = gen_.size() != m_.size();
    for (size_t i = 0; !stale && i < m_.size(); ++i) stale = gen_[i] != m_[i]->generation();
    if (stale) {
        drop_graphs();
        for (auto &m : m_) gen_.push_back(m->generation());
    }
    return true;
}

void Ensemble::build_graph(int ring) {
    OWT_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
    try {
        OWT_HIP(hipEventRecord(fork_, s_));
        for (size_t i = 0; i < m_.size(); ++i) {
            Channel &m = *m_[i];
            OWT_HIP(hipStreamWaitEvent(m.stream(), fork_, 0));
            m.enqueue_body();
            OWT_HIP(hipEventRecord(join_[i], m.stream()));
        }
        for (hipEvent_t e : join_) OWT_HIP(hipStreamWaitEvent(s_, e, 0));
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
    if (!graph_ready()) {
        // members one by one, each on its own stream, then joined
        for (auto &m : m_) m->enqueue_step();
        for (auto &m : m_) m->sync();
        return;
    }
    for (auto &m : m_) m->prepare_step();
    const int r = m_[0]->ring();
    if (!gexec_[r]) build_graph(r);
    OWT_HIP(hipGraphLaunch(gexec_[r], s_));
    for (auto &m : m_) m->advance_ring();
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
