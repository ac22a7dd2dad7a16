// An ensemble of independent channels on one GPU, stepped together.
#pragma once
#include <memory>
#include <vector>

#include "channel.hpp"

namespace owt {

// Members may differ in every parameter (grid sizes included). The batched step
// of all members is captured into one graph per ring phase: a fork from the
// ensemble's stream into each member's stream, the member's step body there,
// and a join back, so members' kernels are independent graph branches.
class Ensemble {
  public:
    Ensemble(const std::vector<Params> &ps, int device);
    ~Ensemble();
    Ensemble(const Ensemble &) = delete;
    Ensemble &operator=(const Ensemble &) = delete;

    int size() const { return (int)m_.size(); }
    Channel &member(int i) { return *m_.at(i); }
    // init_fields of member m with seeds[m]
    void init_fields(int mode, const uint64_t *seeds, double amp);
    // n steps of every member
    void step(int n);
    // n steps between two events on the ensemble's stream (graphs built first); ms
    double time_steps(int n);

  private:
    bool graph_ready();
    void enqueue_step();
    void build_graph(int ring);
    void drop_graphs();

    int device_ = 0;
    hipStream_t s_ = nullptr;
    hipEvent_t fork_ = nullptr;
    std::vector<hipEvent_t> join_;
    std::vector<std::unique_ptr<Channel>> m_;
    std::vector<unsigned> gen_;  // members' generation() when the graphs were built
    hipGraph_t graph_[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t gexec_[3] = {nullptr, nullptr, nullptr};
};

}  // namespace owt
