// serve.hip — device-resident request loop over many concurrent runs, and
// batched pressure solves.
//
// Each admitted run keeps its fields in HBM and owns a HIP stream.  A pool of
// worker threads (the "slots") takes step requests in arrival order; a
// request whose run is busy waits while later requests for idle runs go
// ahead, so work from different runs overlaps on the device as soon as a slot
// frees up (continuous batching at the stream level).  A step's arithmetic
// does not depend on what else is in flight: results are the same alone and
// in the mix.  Admission measures the run's device footprint and refuses
// runs that would exceed the server's HBM budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nbl_api.h"
#include "octree.hpp"
#include "serve.hpp"

namespace nbl {

using Clock = std::chrono::steady_clock;

Server::Server(nbl_ctx* ctx, int slots, uint64_t budget) : ctx_(ctx), budget_(budget) {
  if (slots < 1 || slots > 64) throw std::invalid_argument("server: slots must be in [1, 64]");
  for (int w = 0; w < slots; ++w) workers_.emplace_back([this] { work(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& t : workers_) t.join();
}

int Server::insert(std::unique_ptr<Slot> s, uint64_t foot) {
  std::lock_guard<std::mutex> lk(mu_);
  if (used_ + foot > budget_) throw CapacityError("server: the run's device footprint exceeds the HBM budget");
  used_ += foot;
  s->bytes = foot;
  runs_.push_back(std::move(s));
  return int(runs_.size()) - 1;
}

int Server::admit(const HostGrid& g, const nbl_flow_cfg& cfg, const nbl_solver_opts& opts) {
  std::lock_guard<std::mutex> lk(admitMu_);
  const uint64_t before = ctx_->bytes.load();
  auto s = std::make_unique<Slot>();
  s->run = std::make_unique<nbl_run>();
  s->run->ctx = ctx_;
  s->run->impl = std::make_unique<RunDev>(ctx_, g, cfg, opts);
  return insert(std::move(s), ctx_->bytes.load() - before);
}

int Server::admitOctree(const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
                        const double origin[3], const double extent[3], int maxLevel,
                        const nbl_octree_flow_cfg& cfg) {
  std::lock_guard<std::mutex> lk(admitMu_);
  const uint64_t before = ctx_->bytes.load();
  auto s = std::make_unique<Slot>();
  s->tree = std::make_unique<nbl_octree>();
  s->tree->ctx = ctx_;
  s->tree->impl = std::make_unique<OctreeDev>(ctx_, keys, levels, masks, n, origin, extent, maxLevel);
  s->flow = std::make_unique<nbl_octree_flow>();
  s->flow->ctx = ctx_;
  s->flow->tree = s->tree.get();
  s->flow->impl = std::make_unique<OctFlowDev>(ctx_, *s->tree->impl, cfg);
  // the multigrid hierarchy (and its FP32 copies) would otherwise be built by
  // the first served step, outside the measured footprint
  s->tree->impl->mgLevels(nullptr, 0, nullptr);
  return insert(std::move(s), ctx_->bytes.load() - before);
}

Server::Slot& Server::idle(int id) {
  if (id < 0 || id >= int(runs_.size())) throw std::invalid_argument("server: unknown run id");
  Slot& s = *runs_[size_t(id)];
  if (s.busy || s.pending) throw std::runtime_error("server: run has requests in flight");
  return s;
}

nbl_run* Server::run(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  Slot& s = idle(id);
  if (!s.run) throw std::invalid_argument("server: run is an octree run");
  return s.run.get();
}

nbl_octree_flow* Server::octreeFlow(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  Slot& s = idle(id);
  if (!s.flow) throw std::invalid_argument("server: run is a uniform-grid run");
  return s.flow.get();
}

namespace {
// An octree step in the uniform report's fields: no drag or momentum residual;
// mass_error = inflow - outflow; cfl the configured number the step size obeys.
nbl_step_report octreeReport(const nbl_octree_step_report& r, double cfl, uint64_t leaves) {
  nbl_step_report o{};
  o.step = r.step;
  o.t = r.t;
  o.dt = r.dt;
  o.cfl = cfl;
  o.divergence_max = r.divergence_max;
  o.mass_error = r.inflow - r.outflow;
  o.poisson_iters = r.poisson_iters;
  o.poisson_residual = r.poisson_residual;
  o.cell_count = leaves;
  o.accepted = 1;
  return o;
}
}  // namespace

uint64_t Server::submit(int id) {
  std::lock_guard<std::mutex> lk(mu_);
  if (id < 0 || id >= int(runs_.size())) throw std::invalid_argument("server: unknown run id");
  const uint64_t t = nextTicket_++;
  queue_.push_back(Req{t, id, Clock::now()});
  live_.insert(t);
  ++runs_[size_t(id)]->pending;
  cv_.notify_all();
  return t;
}

void Server::wait(uint64_t ticket, nbl_step_report* rep, double* latencyMs) {
  std::unique_lock<std::mutex> lk(mu_);
  if (!live_.count(ticket)) throw std::invalid_argument("server: unknown or already collected ticket");
  doneCv_.wait(lk, [&] { return done_.count(ticket) != 0; });
  Done d = done_[ticket];
  done_.erase(ticket);
  live_.erase(ticket);
  lk.unlock();
  if (!d.error.empty()) throw std::runtime_error(d.error);
  if (rep) *rep = d.rep;
  if (latencyMs) *latencyMs = d.latencyMs;
}

void Server::footprint(uint64_t* used, uint64_t* budget) const {
  std::lock_guard<std::mutex> lk(mu_);
  if (used) *used = used_;
  if (budget) *budget = budget_;
}

void Server::work() {
  (void)hipSetDevice(ctx_->device);
  for (;;) {
    Req req{};
    Slot* slot = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      // oldest request whose run is idle; drains the queue before stopping
      auto pick = [&]() -> bool {
        for (auto it = queue_.begin(); it != queue_.end(); ++it)
          if (!runs_[size_t(it->run)]->busy) {
            req = *it;
            queue_.erase(it);
            return true;
          }
        return false;
      };
      cv_.wait(lk, [&] { return pick() || (stop_ && queue_.empty()); });
      if (req.ticket == 0) return;
      slot = runs_[size_t(req.run)].get();
      slot->busy = true;
    }
    Done d;
    try {
      if (slot->flow)
        d.rep = octreeReport(slot->flow->impl->step(0.0), slot->flow->impl->cfl(), slot->tree->impl->size());
      else
        d.rep = slot->run->impl->step(0.0);
    } catch (const std::exception& e) {
      d.error = e.what();
    }
    d.latencyMs = std::chrono::duration<double, std::milli>(Clock::now() - req.arrival).count();
    {
      std::lock_guard<std::mutex> lk(mu_);
      slot->busy = false;
      --slot->pending;
      done_[req.ticket] = d;
    }
    cv_.notify_all();
    doneCv_.notify_all();
  }
}

// Batched solves: `concurrency` host threads, each driving solves on its own
// solver's stream, so independent problems overlap on the device.  Staging
// buffers are allocated before the threads start (a device free would
// synchronise the device and serialise the batch).
void solveBatch(nbl_poisson* const* ps, const double* const* b, double* const* x, int n, int concurrency,
                nbl_poisson_report* reps) {
  if (n <= 0) return;
  if (concurrency < 1) throw std::invalid_argument("batch: concurrency must be >= 1");
  for (int q = 0; q < n; ++q)
    if (!ps[q] || !b[q] || !x[q]) throw std::invalid_argument("batch: NULL solver or vector");
  std::vector<DevBuf> db, dx;
  db.reserve(size_t(n));
  dx.reserve(size_t(n));
  for (int q = 0; q < n; ++q) {
    db.emplace_back(ps[q]->ctx, ps[q]->grid.cells() * 8);
    dx.emplace_back(ps[q]->ctx, ps[q]->grid.cells() * 8);
  }
  std::mutex mu;
  int next = 0;
  std::string err;
  auto body = [&] {
    for (;;) {
      int q;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (next >= n || !err.empty()) return;
        q = next++;
      }
      try {
        nbl_poisson* p = ps[q];
        NBL_HIP(hipSetDevice(p->ctx->device));
        const size_t cells = p->grid.cells();
        NBL_HIP(hipMemcpyAsync(db[size_t(q)].as<void>(), b[q], cells * 8, hipMemcpyHostToDevice, p->stream));
        reps[q] = p->impl->solve(db[size_t(q)].as<double>(), dx[size_t(q)].as<double>(), p->stream);
        NBL_HIP(hipMemcpyAsync(x[q], dx[size_t(q)].as<void>(), cells * 8, hipMemcpyDeviceToHost, p->stream));
        NBL_HIP(hipStreamSynchronize(p->stream));
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        if (err.empty()) err = e.what();
      }
    }
  };
  std::vector<std::thread> th;
  const int nt = std::min(concurrency, n);
  for (int t = 0; t < nt; ++t) th.emplace_back(body);
  for (auto& t : th) t.join();
  if (!err.empty()) throw std::runtime_error(err);
}

// Batched octree solves, as solveBatch: each octree's own stream, host vectors
// in caller order; Jacobi-PCG or (mg) the multigrid-preconditioned CG.
void octreeSolveBatch(nbl_octree* const* os, const double* const* b, double* const* x, int n, int concurrency,
                      double tol, int maxIters, nbl_poisson_report* reps, bool mg) {
  if (n <= 0) return;
  if (concurrency < 1) throw std::invalid_argument("batch: concurrency must be >= 1");
  for (int q = 0; q < n; ++q)
    if (!os[q] || !b[q] || !x[q]) throw std::invalid_argument("batch: NULL octree or vector");
  std::mutex mu;
  int next = 0;
  std::string err;
  auto body = [&] {
    for (;;) {
      int q;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (next >= n || !err.empty()) return;
        q = next++;
      }
      try {
        NBL_HIP(hipSetDevice(os[q]->ctx->device));
        reps[q] = mg ? os[q]->impl->solveMg(b[q], x[q], tol, maxIters) : os[q]->impl->solve(b[q], x[q], tol, maxIters);
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(mu);
        if (err.empty()) err = e.what();
      }
    }
  };
  std::vector<std::thread> th;
  const int nt = std::min(concurrency, n);
  for (int t = 0; t < nt; ++t) th.emplace_back(body);
  for (auto& t : th) t.join();
  if (!err.empty()) throw std::runtime_error(err);
}

}  // namespace nbl
