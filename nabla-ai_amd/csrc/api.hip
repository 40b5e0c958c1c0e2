// api.hip — the extern "C" boundary (include/nbl_api.h).  Every entry point
// catches C++ exceptions and turns them into a status code plus a
// thread-local message; nothing throws across the ABI.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbl_api.h"
#include "octflow.hpp"
#include "octree.hpp"
#include "serve.hpp"
#include "solver.hpp"
#include "tree.hpp"

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return NBL_OK;
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return NBL_ERR_INVALID;
  } catch (const nbl::HipError& e) {
    g_err = e.what();
    return NBL_ERR_DEVICE;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return NBL_ERR_CAPACITY;
  } catch (const nbl::CapacityError& e) {
    g_err = e.what();
    return NBL_ERR_CAPACITY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return NBL_ERR_STATE;
  } catch (...) {
    g_err = "unknown error";
    return NBL_ERR_STATE;
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw std::invalid_argument(what);
}

void useDevice(nbl_ctx* ctx) { NBL_HIP(hipSetDevice(ctx->device)); }

nbl_solver_opts optsOr(const nbl_solver_opts* o) {
  nbl_solver_opts d;
  nbl_solver_opts_default(&d);
  return o ? *o : d;
}

// Staging of nbl_poisson_solve_noalloc: b's and x's device copies, each up to
// kNoAllocCells FP64 values (2 x 16 MiB), held by every GPU the library runs on.
constexpr long long kNoAllocCells = 1LL << 21;
__device__ __align__(256) double scratch[2 * kNoAllocCells];

// Device address of `scratch` on device `dev` (looked up once per device, with
// that device current: each GPU holds its own copy at its own address).
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
double* g_scratchPtr[kMaxDevices] = {};

double* noAllocScratch(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    int cur = -1;
    NBL_HIP(hipGetDevice(&cur));
    if (cur != dev) throw std::logic_error("noAllocScratch: device is not current");
    void* p = nullptr;
    NBL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw nbl::HipError("hipGetSymbolAddress(scratch) returned NULL");
    g_scratchPtr[dev] = static_cast<double*>(p);
  }
  return g_scratchPtr[dev];
}

}  // namespace

extern "C" {

void nbl_flow_cfg_default(nbl_flow_cfg* c) {  // FlowConfig{} (ns_solver.hpp:14-32)
  if (!c) return;
  c->reynolds = 100.0;
  c->ref_length = 1.0;
  c->velocity_scale = 1.0;
  c->density = 1.0;
  c->cfl = 0.8;
  c->body_force[0] = c->body_force[1] = c->body_force[2] = 0.0;
  c->convection = NBL_CONV_CENTRAL;
  c->poisson_tol = 1e-9;
  c->poisson_max_iters = 5000;
  c->tol_continuity = 1e-7;
  c->tol_mass = 1e-8;
  c->ref_area = 0.0;
}

void nbl_solver_opts_default(nbl_solver_opts* o) {
  if (!o) return;
  o->precision = NBL_PREC_FP64;
  o->pre_smooth = 2;
  o->post_smooth = 2;
  o->check_every = 1;
  o->coarse = NBL_COARSE_DIRECT;
  o->cells_per_thread = 0;
  o->loops = 0;
  o->rhs_planes = 0;
  o->vcycle = 0;
  o->launch_order = 0;
  o->iteration_graph = 0;
}

const char* nbl_last_error(void) { return g_err.c_str(); }
const char* nbl_version(void) { return "nabla-mi355x 0.1.0 (gfx950)"; }

int nbl_ctx_create(int device, nbl_ctx** out) {
  return guarded([&] {
    need(out != nullptr, "out is NULL");
    int n = 0;
    NBL_HIP(hipGetDeviceCount(&n));
    need(device >= 0 && device < n, "device index out of range");
    NBL_HIP(hipSetDevice(device));
    auto* c = new nbl_ctx();
    c->device = device;
    *out = c;
  });
}

int nbl_ctx_destroy(nbl_ctx* ctx) {
  return guarded([&] { delete ctx; });
}

int nbl_ctx_bytes_in_use(nbl_ctx* ctx, uint64_t* bytes) {
  return guarded([&] {
    need(ctx && bytes, "NULL argument");
    *bytes = ctx->bytes.load();
  });
}

int nbl_ctx_probe_copy(nbl_ctx* ctx, uint64_t bytes, int reps, double* ms) {
  return guarded([&] {
    need(ctx && ms && bytes > 0 && reps > 0, "bad probe arguments");
    useDevice(ctx);
    nbl::DevBuf a(ctx, bytes), b(ctx, bytes);
    hipStream_t s;
    NBL_HIP(hipStreamCreate(&s));
    hipEvent_t e0, e1;
    NBL_HIP(hipEventCreate(&e0));
    NBL_HIP(hipEventCreate(&e1));
    NBL_HIP(hipMemsetAsync(a.as<void>(), 0, bytes, s));
    NBL_HIP(hipMemcpyAsync(b.as<void>(), a.as<void>(), bytes, hipMemcpyDeviceToDevice, s));
    NBL_HIP(hipEventRecord(e0, s));
    for (int q = 0; q < reps; ++q)
      NBL_HIP(hipMemcpyAsync(b.as<void>(), a.as<void>(), bytes, hipMemcpyDeviceToDevice, s));
    NBL_HIP(hipEventRecord(e1, s));
    NBL_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    NBL_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
  });
}

// ---- runs -----------------------------------------------------------------
int nbl_run_create(nbl_ctx* ctx, const nbl_grid_desc* grid, const nbl_flow_cfg* cfg,
                   const nbl_solver_opts* opts, nbl_run** out) {
  return guarded([&] {
    need(ctx && grid && cfg && out, "NULL argument");
    useDevice(ctx);
    auto r = std::make_unique<nbl_run>();
    r->ctx = ctx;
    r->impl = std::make_unique<nbl::RunDev>(ctx, nbl::HostGrid::from(grid), *cfg, optsOr(opts));
    *out = r.release();
  });
}

int nbl_run_destroy(nbl_run* run) {
  return guarded([&] { delete run; });
}

int nbl_run_initialize(nbl_run* run) {
  return guarded([&] {
    need(run, "run is NULL");
    useDevice(run->ctx);
    run->impl->initialize();
  });
}

int nbl_run_set_state(nbl_run* run, const double* u, size_t nu, const double* v, size_t nv, const double* w,
                      size_t nw, const double* p, size_t np, int step, double t, double dt_last) {
  return guarded([&] {
    need(run, "run is NULL");
    useDevice(run->ctx);
    run->impl->setState(u, nu, v, nv, w, nw, p, np, step, t, dt_last);
  });
}

int nbl_run_get_state(nbl_run* run, double* u, double* v, double* w, double* p, int* step, double* t,
                      double* dt_last) {
  return guarded([&] {
    need(run, "run is NULL");
    useDevice(run->ctx);
    run->impl->getState(u, v, w, p, step, t, dt_last);
  });
}

int nbl_run_compute_dt(nbl_run* run, double* dt) {
  return guarded([&] {
    need(run && dt, "NULL argument");
    useDevice(run->ctx);
    *dt = run->impl->computeDt();
  });
}

int nbl_run_step(nbl_run* run, double dt, nbl_step_report* rep) {
  return guarded([&] {
    need(run, "run is NULL");
    useDevice(run->ctx);
    const nbl_step_report r = run->impl->step(dt);
    if (rep) *rep = r;
  });
}

int nbl_run_step_many(nbl_run* run, int nsteps, nbl_step_report* reps) {
  return guarded([&] {
    need(run, "run is NULL");
    need(nsteps >= 0, "nsteps must be >= 0");
    useDevice(run->ctx);
    for (int s = 0; s < nsteps; ++s) {
      const nbl_step_report r = run->impl->step(0.0);
      if (reps) reps[s] = r;
    }
  });
}

int nbl_run_device_state(nbl_run* run, double** u, double** v, double** w, double** p) {
  return guarded([&] {
    need(run, "run is NULL");
    if (u) *u = run->impl->u();
    if (v) *v = run->impl->v();
    if (w) *w = run->impl->w();
    if (p) *p = run->impl->p();
  });
}

int nbl_run_rhs(nbl_run* run, double* ru, double* rv, double* rw) {
  return guarded([&] {
    need(run && ru && rv && rw, "NULL argument");
    useDevice(run->ctx);
    run->impl->rhs(ru, rv, rw);
  });
}

int nbl_run_cell_velocity(nbl_run* run, int component, double* out) {
  return guarded([&] {
    need(run && out, "NULL argument");
    useDevice(run->ctx);
    run->impl->cellVelocity(component, out);
  });
}

int nbl_run_set_pressure_solver(nbl_run* run, nbl_host_poisson_fn solve, void* user) {
  return guarded([&] {
    need(run, "run is NULL");
    run->impl->setHostSolver(solve, user);
  });
}

int nbl_run_last_solves(nbl_run* run, double* ms, int* iters, int* vcycles) {
  return guarded([&] {
    need(run, "run is NULL");
    run->impl->lastSolves(ms, iters, vcycles);
  });
}

// ---- pressure plugin ---------------------------------------------------------
int nbl_poisson_create(nbl_ctx* ctx, const nbl_grid_desc* grid, int kind, double tol, int max_iters,
                       const nbl_solver_opts* opts, nbl_poisson** out) {
  return guarded([&] {
    need(ctx && grid && out, "NULL argument");
    useDevice(ctx);
    auto ps = std::make_unique<nbl_poisson>();
    ps->ctx = ctx;
    ps->grid = nbl::HostGrid::from(grid);
    ps->impl = std::make_unique<nbl::PoissonDev>(ctx, ps->grid, kind, tol, max_iters, optsOr(opts));
    NBL_HIP(hipStreamCreate(&ps->stream));
    *out = ps.release();
  });
}

int nbl_poisson_destroy(nbl_poisson* ps) {
  return guarded([&] {
    if (!ps) return;
    if (ps->stream) {
      (void)hipStreamSynchronize(ps->stream);
      (void)hipStreamDestroy(ps->stream);
    }
    delete ps;
  });
}

namespace {
// The single-problem entry points refuse a batch (its vectors hold every problem).
void needSingle(const nbl_poisson* ps) {
  need(!ps->impl->batched(), "this entry point takes a single-problem solver; use nbl_poisson_batch_solve");
}
}  // namespace

int nbl_poisson_batch_create(nbl_ctx* ctx, const nbl_grid_desc* grid, int problems, double tol, int max_iters,
                             const nbl_solver_opts* opts, nbl_poisson** out) {
  return guarded([&] {
    need(ctx && grid && out, "NULL argument");
    need(problems >= 1, "a batch holds at least one problem");
    useDevice(ctx);
    auto ps = std::make_unique<nbl_poisson>();
    ps->ctx = ctx;
    ps->grid = nbl::HostGrid::from(grid);
    ps->impl = std::make_unique<nbl::PoissonDev>(ctx, ps->grid, 0, tol, max_iters, optsOr(opts), problems);
    NBL_HIP(hipStreamCreate(&ps->stream));
    *out = ps.release();
  });
}

int nbl_poisson_batch_size(nbl_poisson* ps, int* problems) {
  return guarded([&] {
    need(ps && problems, "NULL argument");
    *problems = ps->impl->batched() ? ps->impl->problems() : 0;
  });
}

int nbl_poisson_batch_solve_device(nbl_poisson* ps, const double* d_b, double* d_x, nbl_poisson_report* reps) {
  return guarded([&] {
    need(ps && d_b && d_x, "NULL argument");
    need(ps->impl->batched(), "not a batch: use nbl_poisson_solve_device");
    useDevice(ps->ctx);
    ps->impl->solve(d_b, d_x, ps->stream);
    NBL_HIP(hipStreamSynchronize(ps->stream));
    if (reps) std::copy(ps->impl->reports().begin(), ps->impl->reports().end(), reps);
  });
}

namespace {
struct Staging {  // device copies of host vectors for the host-buffer entry points
  nbl::DevBuf a, b;
  Staging(nbl_ctx* ctx, size_t n) : a(ctx, n * 8), b(ctx, n * 8) {}
};
}  // namespace

int nbl_poisson_batch_solve(nbl_poisson* ps, const double* const* b, double* const* x, nbl_poisson_report* reps) {
  return guarded([&] {
    need(ps && b && x, "NULL argument");
    need(ps->impl->batched(), "not a batch: use nbl_poisson_solve");
    const int nb = ps->impl->problems();
    for (int q = 0; q < nb; ++q) need(b[q] && x[q], "NULL problem vector");
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    Staging st(ps->ctx, n * size_t(nb));
    for (int q = 0; q < nb; ++q)
      NBL_HIP(hipMemcpyAsync(st.a.as<double>() + size_t(q) * n, b[q], n * 8, hipMemcpyHostToDevice, ps->stream));
    ps->impl->solve(st.a.as<double>(), st.b.as<double>(), ps->stream);
    for (int q = 0; q < nb; ++q)
      NBL_HIP(hipMemcpyAsync(x[q], st.b.as<double>() + size_t(q) * n, n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
    if (reps) std::copy(ps->impl->reports().begin(), ps->impl->reports().end(), reps);
  });
}

int nbl_poisson_solve(nbl_poisson* ps, const double* b, double* x, nbl_poisson_report* rep) {
  return guarded([&] {
    need(ps && b && x, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    Staging st(ps->ctx, n);
    NBL_HIP(hipMemcpyAsync(st.a.as<void>(), b, n * 8, hipMemcpyHostToDevice, ps->stream));
    const nbl_poisson_report r = ps->impl->solve(st.a.as<double>(), st.b.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(x, st.b.as<void>(), n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
    if (rep) *rep = r;
  });
}

// The same solve with the device copies of b and x in `scratch` (a file-scope
// device array each GPU holds for the life of the process) instead of two
// allocations per call.  The host array name is only a stand-in: kernels and
// copies use the address hipGetSymbolAddress returns for the current device,
// looked up once per device (noAllocScratch).  b's copy is written whole by
// the upload and x's whole by the solve's last kernel, so nothing has to be
// cleared between calls; calls that overlap on one GPU share the array and
// race.
int nbl_poisson_solve_noalloc(nbl_poisson* ps, const double* b, double* x, nbl_poisson_report* rep) {
  return guarded([&] {
    need(ps && b && x, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    if (n > size_t(kNoAllocCells))
      throw nbl::CapacityError("nbl_poisson_solve_noalloc: " + std::to_string(n) + " cells exceed the " +
                               std::to_string(kNoAllocCells) + "-cell scratch; use nbl_poisson_solve");
    double* s = noAllocScratch(ps->ctx->device);
    double* db = s;
    double* dx = s + kNoAllocCells;
    NBL_HIP(hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, ps->stream));
    const nbl_poisson_report r = ps->impl->solve(db, dx, ps->stream);
    NBL_HIP(hipMemcpyAsync(x, dx, n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
    if (rep) *rep = r;
  });
}

int nbl_poisson_solve_device(nbl_poisson* ps, const double* d_b, double* d_x, nbl_poisson_report* rep) {
  return guarded([&] {
    need(ps && d_b && d_x, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const nbl_poisson_report r = ps->impl->solve(d_b, d_x, ps->stream);
    NBL_HIP(hipStreamSynchronize(ps->stream));
    if (rep) *rep = r;
  });
}

int nbl_poisson_precondition(nbl_poisson* ps, const double* r, double* z) {
  return guarded([&] {
    need(ps && r && z, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    Staging st(ps->ctx, n);
    NBL_HIP(hipMemcpyAsync(st.a.as<void>(), r, n * 8, hipMemcpyHostToDevice, ps->stream));
    ps->impl->precondition(st.a.as<double>(), st.b.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(z, st.b.as<void>(), n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_apply(nbl_poisson* ps, const double* x, double* out) {
  return guarded([&] {
    need(ps && x && out, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    Staging st(ps->ctx, n);
    NBL_HIP(hipMemcpyAsync(st.a.as<void>(), x, n * 8, hipMemcpyHostToDevice, ps->stream));
    nbl::launch_apply(ps->impl->level(0), st.a.as<double>(), st.b.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(out, st.b.as<void>(), n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_diagonal(nbl_poisson* ps, double* diag) {
  return guarded([&] {
    need(ps && diag, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    nbl::DevBuf d(ps->ctx, n * 8);
    nbl::launch_diag(ps->impl->level(0), d.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(diag, d.as<void>(), n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_levels(nbl_poisson* ps, int* nlevels, int* dims) {
  return guarded([&] {
    need(ps && nlevels, "NULL argument");
    *nlevels = ps->impl->nlevels();
    if (dims)
      for (int l = 0; l < ps->impl->nlevels(); ++l) {
        const nbl::MgLevel& L = ps->impl->level(l);
        dims[3 * l] = L.nx;
        dims[3 * l + 1] = L.ny;
        dims[3 * l + 2] = L.nz;
      }
  });
}

int nbl_poisson_smooth_once(nbl_poisson* ps, const double* b, double* x) {
  return guarded([&] {
    need(ps && b && x, "NULL argument");
    needSingle(ps);
    need(ps->impl->kind() == 0, "smoother exists for the mg-pcg solver only");
    useDevice(ps->ctx);
    const size_t n = ps->grid.cells();
    Staging st(ps->ctx, n);
    nbl::DevBuf xo(ps->ctx, n * 8);
    NBL_HIP(hipMemcpyAsync(st.a.as<void>(), b, n * 8, hipMemcpyHostToDevice, ps->stream));
    NBL_HIP(hipMemcpyAsync(st.b.as<void>(), x, n * 8, hipMemcpyHostToDevice, ps->stream));
    nbl::launch_mg_sweep<double>(ps->impl->level(0), st.a.as<double>(), nbl::MeanRef{}, st.b.as<double>(),
                                 nbl::MeanRef{}, false, nullptr, nbl::MeanRef{}, nullptr, xo.as<double>(), false,
                                 ps->impl->red(), nullptr, ps->stream);
    NBL_HIP(hipMemcpyAsync(x, xo.as<void>(), n * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_restrict(nbl_poisson* ps, const double* rf, double* rc) {
  return guarded([&] {
    need(ps && rf && rc, "NULL argument");
    needSingle(ps);
    need(ps->impl->nlevels() >= 2, "grid has a single multigrid level");
    useDevice(ps->ctx);
    const nbl::MgLevel& L = ps->impl->level(0);
    const nbl::MgLevel& Lc = ps->impl->level(1);
    nbl::DevBuf a(ps->ctx, size_t(L.n) * 8), c(ps->ctx, size_t(Lc.n) * 8);
    NBL_HIP(hipMemcpyAsync(a.as<void>(), rf, size_t(L.n) * 8, hipMemcpyHostToDevice, ps->stream));
    nbl::launch_restrict_plain(L, Lc, a.as<double>(), c.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(rc, c.as<void>(), size_t(Lc.n) * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_prolong(nbl_poisson* ps, const double* ec, double* xf) {
  return guarded([&] {
    need(ps && ec && xf, "NULL argument");
    needSingle(ps);
    need(ps->impl->nlevels() >= 2, "grid has a single multigrid level");
    useDevice(ps->ctx);
    const nbl::MgLevel& L = ps->impl->level(0);
    const nbl::MgLevel& Lc = ps->impl->level(1);
    nbl::DevBuf a(ps->ctx, size_t(L.n) * 8), c(ps->ctx, size_t(Lc.n) * 8);
    NBL_HIP(hipMemcpyAsync(a.as<void>(), xf, size_t(L.n) * 8, hipMemcpyHostToDevice, ps->stream));
    NBL_HIP(hipMemcpyAsync(c.as<void>(), ec, size_t(Lc.n) * 8, hipMemcpyHostToDevice, ps->stream));
    nbl::launch_prolong_plain(L, Lc, c.as<double>(), a.as<double>(), ps->stream);
    NBL_HIP(hipMemcpyAsync(xf, a.as<void>(), size_t(L.n) * 8, hipMemcpyDeviceToHost, ps->stream));
    NBL_HIP(hipStreamSynchronize(ps->stream));
  });
}

int nbl_poisson_last_timing(nbl_poisson* ps, double* ms, int* vcycles) {
  return guarded([&] {
    need(ps, "NULL argument");
    if (ms) *ms = ps->impl->lastMs();
    if (vcycles) *vcycles = ps->impl->lastVcycles();
  });
}

int nbl_poisson_profile(nbl_poisson* ps, int kernel, int reps, double* ms, double* bytes) {
  return guarded([&] {
    need(ps && ms && bytes, "NULL argument");
    needSingle(ps);
    useDevice(ps->ctx);
    ps->impl->profile(kernel, reps, ps->stream, ms, bytes);
  });
}

// ---- octree construction ----------------------------------------------------------
int nbl_tree_create(const double origin[3], const double extent[3], int max_level, nbl_tree** out) {
  return guarded([&] {
    need(origin && extent && out, "NULL argument");
    *out = new nbl_tree(origin, extent, max_level);
  });
}
int nbl_tree_destroy(nbl_tree* t) {
  return guarded([&] { delete t; });
}
int nbl_tree_count(nbl_tree* t, size_t* n) {
  return guarded([&] {
    need(t && n, "NULL argument");
    *n = t->impl.size();
  });
}
int nbl_tree_refine_uniform(nbl_tree* t, int times) {
  return guarded([&] {
    need(t, "tree is NULL");
    need(times >= 0, "times must be >= 0");
    t->impl.refineUniform(times);
  });
}
int nbl_tree_refine_flagged(nbl_tree* t, const uint8_t* flags, size_t n) {
  return guarded([&] {
    need(t && flags, "NULL argument");
    t->impl.refineFlagged(flags, n);
  });
}
int nbl_tree_is_balanced(nbl_tree* t, int* balanced) {
  return guarded([&] {
    need(t && balanced, "NULL argument");
    *balanced = t->impl.balanced() ? 1 : 0;
  });
}
int nbl_tree_leaves(nbl_tree* t, uint64_t* keys, uint8_t* levels, double* centers) {
  return guarded([&] {
    need(t, "tree is NULL");
    t->impl.leaves(keys, levels, centers);
  });
}

// ---- octree ---------------------------------------------------------------------
int nbl_octree_create(nbl_ctx* ctx, const uint64_t* keys, const uint8_t* levels, const uint8_t* masks, size_t n,
                      const double origin[3], const double extent[3], int max_level, nbl_octree** out) {
  return guarded([&] {
    need(ctx && keys && levels && masks && origin && extent && out, "NULL argument");
    useDevice(ctx);
    auto o = std::make_unique<nbl_octree>();
    o->ctx = ctx;
    o->impl = std::make_unique<nbl::OctreeDev>(ctx, keys, levels, masks, n, origin, extent, max_level);
    *out = o.release();
  });
}

int nbl_octree_destroy(nbl_octree* oc) {
  return guarded([&] { delete oc; });
}

int nbl_octree_count(nbl_octree* oc, size_t* n, size_t* nnz) {
  return guarded([&] {
    need(oc, "octree is NULL");
    if (n) *n = oc->impl->size();
    if (nnz) *nnz = oc->impl->nnz();
  });
}

int nbl_octree_coef_classes(nbl_octree* oc, int* ncls) {
  return guarded([&] {
    need(oc && ncls, "NULL argument");
    *ncls = oc->impl->coefClasses();
  });
}

int nbl_octree_permutation(nbl_octree* oc, uint64_t* perm) {
  return guarded([&] {
    need(oc && perm, "NULL argument");
    useDevice(oc->ctx);
    oc->impl->permutation(perm);
  });
}

int nbl_octree_face_table(nbl_octree* oc, uint8_t* count, int32_t* nbr) {
  return guarded([&] {
    need(oc, "octree is NULL");
    useDevice(oc->ctx);
    oc->impl->faceTable(count, nbr);
  });
}

int nbl_octree_apply(nbl_octree* oc, const double* x, double* out) {
  return guarded([&] {
    need(oc && x && out, "NULL argument");
    useDevice(oc->ctx);
    oc->impl->apply(x, out);
  });
}

int nbl_octree_diagonal(nbl_octree* oc, double* diag) {
  return guarded([&] {
    need(oc && diag, "NULL argument");
    useDevice(oc->ctx);
    oc->impl->diagonal(diag);
  });
}

int nbl_octree_poisson_solve(nbl_octree* oc, const double* b, double* x, double tol, int max_iters,
                             nbl_poisson_report* rep) {
  return guarded([&] {
    need(oc && b && x, "NULL argument");
    useDevice(oc->ctx);
    const nbl_poisson_report r = oc->impl->solve(b, x, tol, max_iters);
    if (rep) *rep = r;
  });
}

int nbl_octree_transport(nbl_octree* oc, const double* u, const double* v, const double* w, double* phi,
                         double dt, int steps, double* integral) {
  return guarded([&] {
    need(oc && u && v && w && phi, "NULL argument");
    useDevice(oc->ctx);
    const double I = oc->impl->transport(u, v, w, phi, dt, steps);
    if (integral) *integral = I;
  });
}

int nbl_octree_sensors(nbl_octree* oc, const double* u, const double* v, const double* w, const double* p,
                       double nu, double* speed_prev, double* resid_prev, const nbl_sensor_fields* out) {
  return guarded([&] {
    need(oc && u && v && w && p && speed_prev && resid_prev, "NULL argument");
    useDevice(oc->ctx);
    const size_t n = oc->impl->size();
    std::vector<double> vals(out ? 8 * n : 0);
    std::vector<uint8_t> fl(out ? 3 * n : 0);
    oc->impl->sensors(u, v, w, p, nu, speed_prev, resid_prev, out ? vals.data() : nullptr,
                      out ? fl.data() : nullptr);
    if (!out) return;
    double* const dst[8] = {out->vorticity_mag, out->strain_mag, out->q_criterion, out->grad_p_mag,
                            out->grad_u_mag,    out->y_plus,     out->tau_wall,    out->trunc_err};
    for (int f = 0; f < 8; ++f)
      if (dst[f]) std::memcpy(dst[f], vals.data() + size_t(f) * n, n * sizeof(double));
    uint8_t* const fdst[3] = {out->near_wall, out->reverse_flow, out->resid_stagnation};
    for (int f = 0; f < 3; ++f)
      if (fdst[f]) std::memcpy(fdst[f], fl.data() + size_t(f) * n, n);
  });
}

int nbl_octree_mg_solve(nbl_octree* oc, const double* b, double* x, double tol, int max_iters,
                        nbl_poisson_report* rep) {
  return guarded([&] {
    need(oc && b && x, "NULL argument");
    useDevice(oc->ctx);
    const nbl_poisson_report r = oc->impl->solveMg(b, x, tol, max_iters);
    if (rep) *rep = r;
  });
}

int nbl_octree_mg_solve_from(nbl_octree* oc, const double* b, double* x, double tol, int max_iters,
                             nbl_poisson_report* rep) {
  return guarded([&] {
    need(oc && b && x, "NULL argument");
    useDevice(oc->ctx);
    const nbl_poisson_report r = oc->impl->solveMg(b, x, tol, max_iters, true);
    if (rep) *rep = r;
  });
}

int nbl_octree_mg_levels(nbl_octree* oc, long long* counts, int cap, int* levels, double* build_ms) {
  return guarded([&] {
    need(oc && levels, "NULL argument");
    need(cap >= 0 && (cap == 0 || counts), "counts must hold cap entries");
    useDevice(oc->ctx);
    *levels = oc->impl->mgLevels(counts, cap, build_ms);
  });
}

int nbl_octree_set_precision(nbl_octree* oc, int precision) {
  return guarded([&] {
    need(oc, "NULL argument");
    need(precision >= NBL_PREC_FP64 && precision <= NBL_PREC_MIXED_BF16,
         "precision must be NBL_PREC_FP64, NBL_PREC_MIXED or NBL_PREC_MIXED_BF16");
    useDevice(oc->ctx);
    oc->impl->setMgPrecision(precision);
  });
}

void nbl_octree_flow_cfg_default(nbl_octree_flow_cfg* c) {
  if (!c) return;
  c->nu = 1.0 / 500.0;
  c->cfl = 0.5;
  c->inlet_speed = 1.0;
  c->inlet_profile = 0;
  const int faces[6] = {NBL_OCT_INLET, NBL_OCT_OUTLET, NBL_OCT_WALL, NBL_OCT_WALL, NBL_OCT_SLIP, NBL_OCT_SLIP};
  for (int d = 0; d < 6; ++d) c->face[d] = faces[d];
  c->convection = 1;
  c->poisson_tol = 1e-8;
  c->poisson_max_iters = 500;
  c->precision = NBL_PREC_FP64;
}

int nbl_octree_flow_create(nbl_octree* oc, const nbl_octree_flow_cfg* cfg, nbl_octree_flow** out) {
  return guarded([&] {
    need(oc && cfg && out, "NULL argument");
    need(cfg->precision >= NBL_PREC_FP64 && cfg->precision <= NBL_PREC_MIXED_BF16, "unknown precision");
    useDevice(oc->ctx);
    auto f = std::make_unique<nbl_octree_flow>();
    f->ctx = oc->ctx;
    f->tree = oc;
    f->impl = std::make_unique<nbl::OctFlowDev>(oc->ctx, *oc->impl, *cfg);
    *out = f.release();
  });
}

int nbl_octree_flow_destroy(nbl_octree_flow* f) {
  return guarded([&] {
    if (!f) return;
    useDevice(f->ctx);
    delete f;
  });
}

int nbl_octree_flow_set_state(nbl_octree_flow* f, const double* u, const double* v, const double* w,
                              const double* p) {
  return guarded([&] {
    need(f && u && v && w, "NULL argument");
    useDevice(f->ctx);
    f->impl->setState(u, v, w, p);
  });
}

int nbl_octree_flow_get_state(nbl_octree_flow* f, double* u, double* v, double* w, double* p) {
  return guarded([&] {
    need(f, "NULL argument");
    useDevice(f->ctx);
    f->impl->getState(u, v, w, p);
  });
}

int nbl_octree_flow_step(nbl_octree_flow* f, double dt, nbl_octree_step_report* rep) {
  return guarded([&] {
    need(f, "NULL argument");
    useDevice(f->ctx);
    const nbl_octree_step_report r = f->impl->step(dt);
    if (rep) *rep = r;
  });
}

int nbl_octree_flow_divergence(nbl_octree_flow* f, double* div) {
  return guarded([&] {
    need(f && div, "NULL argument");
    useDevice(f->ctx);
    f->impl->divergence(div);
  });
}

int nbl_octree_profile(nbl_octree* oc, int kernel, int reps, double* ms, double* bytes) {
  return guarded([&] {
    need(oc && ms && bytes, "NULL argument");
    useDevice(oc->ctx);
    oc->impl->profile(kernel, reps, ms, bytes);
  });
}

int nbl_octree_last_timing(nbl_octree* oc, double* setup_ms, double* solve_ms) {
  return guarded([&] {
    need(oc, "NULL argument");
    if (setup_ms) *setup_ms = oc->impl->setupMs();
    if (solve_ms) *solve_ms = oc->impl->lastSolveMs();
  });
}

int nbl_run_sweep_timing(nbl_run* run, int enable, double* avg_ms, long long* launches, double* bytes) {
  return guarded([&] {
    need(run, "run is NULL");
    if (enable >= 0) run->impl->poisson().setSweepTiming(enable != 0);
    run->impl->poisson().sweepTiming(avg_ms, launches, bytes);
  });
}

// ---- serving and batched solves -------------------------------------------
int nbl_server_create(nbl_ctx* ctx, int slots, uint64_t hbm_budget, nbl_server** out) {
  return guarded([&] {
    need(ctx && out, "NULL argument");
    useDevice(ctx);
    auto s = std::make_unique<nbl_server>();
    s->ctx = ctx;
    s->impl = std::make_unique<nbl::Server>(ctx, slots, hbm_budget);
    *out = s.release();
  });
}

int nbl_server_destroy(nbl_server* srv) {
  return guarded([&] { delete srv; });
}

int nbl_server_admit(nbl_server* srv, const nbl_grid_desc* grid, const nbl_flow_cfg* cfg,
                     const nbl_solver_opts* opts, int* run_id) {
  return guarded([&] {
    need(srv && grid && cfg && run_id, "NULL argument");
    useDevice(srv->ctx);
    *run_id = srv->impl->admit(nbl::HostGrid::from(grid), *cfg, optsOr(opts));
  });
}

int nbl_server_run(nbl_server* srv, int run_id, nbl_run** run) {
  return guarded([&] {
    need(srv && run, "NULL argument");
    *run = srv->impl->run(run_id);
  });
}

int nbl_server_admit_octree(nbl_server* srv, const uint64_t* keys, const uint8_t* levels, const uint8_t* masks,
                            size_t n, const double origin[3], const double extent[3], int max_level,
                            const nbl_octree_flow_cfg* cfg, int* run_id) {
  return guarded([&] {
    need(srv && keys && levels && masks && origin && extent && cfg && run_id, "NULL argument");
    need(cfg->precision >= NBL_PREC_FP64 && cfg->precision <= NBL_PREC_MIXED_BF16, "unknown precision");
    useDevice(srv->ctx);
    *run_id = srv->impl->admitOctree(keys, levels, masks, n, origin, extent, max_level, *cfg);
  });
}

int nbl_server_octree_flow(nbl_server* srv, int run_id, nbl_octree_flow** flow) {
  return guarded([&] {
    need(srv && flow, "NULL argument");
    *flow = srv->impl->octreeFlow(run_id);
  });
}

int nbl_server_submit(nbl_server* srv, int run_id, uint64_t* ticket) {
  return guarded([&] {
    need(srv && ticket, "NULL argument");
    *ticket = srv->impl->submit(run_id);
  });
}

int nbl_server_wait(nbl_server* srv, uint64_t ticket, nbl_step_report* rep, double* latency_ms) {
  return guarded([&] {
    need(srv, "server is NULL");
    srv->impl->wait(ticket, rep, latency_ms);
  });
}

int nbl_server_footprint(nbl_server* srv, uint64_t* used, uint64_t* budget) {
  return guarded([&] {
    need(srv, "server is NULL");
    srv->impl->footprint(used, budget);
  });
}

int nbl_poisson_solve_batch(nbl_poisson* const* ps, const double* const* b, double* const* x, int n,
                            int concurrency, nbl_poisson_report* reps) {
  return guarded([&] {
    need(ps && b && x && reps && n >= 0, "NULL argument");
    for (int q = 0; q < n; ++q) {
      need(ps[q] != nullptr, "NULL solver");
      needSingle(ps[q]);
    }
    if (n > 0) useDevice(ps[0]->ctx);
    nbl::solveBatch(ps, b, x, n, concurrency, reps);
  });
}

int nbl_octree_solve_batch(nbl_octree* const* os, const double* const* b, double* const* x, int n,
                           int concurrency, double tol, int max_iters, nbl_poisson_report* reps) {
  return guarded([&] {
    need(os && b && x && reps && n >= 0, "NULL argument");
    if (n > 0) useDevice(os[0]->ctx);
    nbl::octreeSolveBatch(os, b, x, n, concurrency, tol, max_iters, reps, false);
  });
}

int nbl_octree_mg_solve_batch(nbl_octree* const* os, const double* const* b, double* const* x, int n,
                              int concurrency, double tol, int max_iters, nbl_poisson_report* reps) {
  return guarded([&] {
    need(os && b && x && reps && n >= 0, "NULL argument");
    if (n > 0) useDevice(os[0]->ctx);
    nbl::octreeSolveBatch(os, b, x, n, concurrency, tol, max_iters, reps, true);
  });
}

}  // extern "C"
