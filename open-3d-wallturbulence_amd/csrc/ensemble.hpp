// An ensemble of independent channels on one GPU, stepped together.
#pragma once
#include <memory>
#include <vector>

#include "channel.hpp"

namespace owt {

// Members may differ in every parameter (grid sizes included). The members
// whose stages all run the same kernel instances (same rows-per-lane class,
// length class, precision, tuned shapes) form a lane, and the batched step of a
// lane is one launch per stage over all its members: each block finds its
// member in a table of member entries by its index. The number of launches per
// step therefore depends on the distinct shapes, not on the member count. Lanes
// run as parallel branches of the step's graph. Members
// with the same parameters read one copy of the operators (inverses, DY
// matrices, eigenbasis, twiddles): each member computes the same operators bit
// for bit, so the results equal the members stepped alone. One graph per ring
// phase holds the launches.
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
    // kernel launches of one grouped step (0 when the members are stepped one by one)
    int launches_per_step();
    // launch-shape autotuner: one member per distinct parameter set is tuned
    // and the others take its choices
    void autotune(int reps);

  private:
    // one grouped launch: entries and first blocks at these offsets of the
    // phase's device table
    struct Launch {
        int stage, lane, nm, grid, ny_max;
        size_t ents, first;  // offsets (bytes) into the table
        EnsEntry rep;
    };
    bool grouped_ready();
    void enqueue_step();
    void build_tables(int ring);
    void build_graph(int ring);
    void drop_graphs();

    int device_ = 0;
    hipStream_t s_ = nullptr;
    std::vector<std::unique_ptr<Channel>> m_;
    std::vector<unsigned> gen_;  // members' generation() when the graphs were built
    std::vector<Launch> launches_[3];
    int nlanes_[3] = {0, 0, 0};
    void *table_[3] = {nullptr, nullptr, nullptr};
    // the lanes' streams and join events, and the fork event
    std::vector<hipStream_t> lane_s_;
    std::vector<hipEvent_t> lane_ev_;
    hipEvent_t fork_ = nullptr;
    hipGraph_t graph_[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t gexec_[3] = {nullptr, nullptr, nullptr};
};

}  // namespace owt
