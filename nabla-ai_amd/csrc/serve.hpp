// serve.hpp — request loop over concurrent device-resident runs, and batched
// pressure solves (no counterpart in the reference, whose LocalJobRunner runs
// one subprocess per job).
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include <deque>

#include "octflow.hpp"
#include "solver.hpp"

namespace nbl {

class Server {
 public:
  Server(nbl_ctx* ctx, int slots, uint64_t budget);
  ~Server();
  int admit(const HostGrid& g, const nbl_flow_cfg& cfg, const nbl_solver_opts& opts);
  // A flow run on octree leaves (caller order, as nbl_octree_create takes
  // them); the server owns the octree, its multigrid hierarchy and the flow.
  int admitOctree(const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
                  const double origin[3], const double extent[3], int maxLevel, const nbl_octree_flow_cfg& cfg);
  nbl_run* run(int id);  // only while the run has no request in flight
  nbl_octree_flow* octreeFlow(int id);  // the same, for octree runs
  uint64_t submit(int id);
  void wait(uint64_t ticket, nbl_step_report* rep, double* latencyMs);
  void footprint(uint64_t* used, uint64_t* budget) const;

 private:
  struct Req {
    uint64_t ticket = 0;
    int run = -1;
    std::chrono::steady_clock::time_point arrival;
  };
  struct Slot {
    std::unique_ptr<nbl_run> run;           // uniform runs
    std::unique_ptr<nbl_octree> tree;       // octree runs: the leaves and
    std::unique_ptr<nbl_octree_flow> flow;  // their flow state
    uint64_t bytes = 0;
    bool busy = false;
    int pending = 0;
  };
  struct Done {
    nbl_step_report rep{};
    double latencyMs = 0.0;
    std::string error;
  };
  void work();
  int insert(std::unique_ptr<Slot> s, uint64_t foot);
  Slot& idle(int id);

  nbl_ctx* ctx_;
  uint64_t budget_, used_ = 0;
  mutable std::mutex mu_;
  std::mutex admitMu_;
  std::condition_variable cv_, doneCv_;
  std::vector<std::unique_ptr<Slot>> runs_;
  std::deque<Req> queue_;
  std::unordered_map<uint64_t, Done> done_;
  std::unordered_set<uint64_t> live_;  // submitted, not yet waited for
  std::vector<std::thread> workers_;
  uint64_t nextTicket_ = 1;
  bool stop_ = false;
};

void solveBatch(nbl_poisson* const* ps, const double* const* b, double* const* x, int n, int concurrency,
                nbl_poisson_report* reps);
}  // namespace nbl
struct nbl_octree;
namespace nbl {
void octreeSolveBatch(nbl_octree* const* os, const double* const* b, double* const* x, int n, int concurrency,
                      double tol, int maxIters, nbl_poisson_report* reps, bool mg);

}  // namespace nbl

struct nbl_server {
  nbl_ctx* ctx;
  std::unique_ptr<nbl::Server> impl;
};
