// Host layer over the C-ABI: device context, the Poisson plugin and NSSolver.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>

#include "nabla/flow/device.hpp"
#include "nabla/flow/ns_solver.hpp"
#include "nabla/flow/poisson.hpp"

namespace nabla::flow {

nbl_ctx* deviceContext() {
  static std::once_flag once;
  static nbl_ctx* ctx = nullptr;
  std::call_once(once, [] {
    const char* e = std::getenv("NABLA_DEVICE");
    nblCheck(nbl_ctx_create(e ? std::atoi(e) : 0, &ctx), "device context");
  });
  return ctx;
}

namespace {
// Grid identity for the solver cache: dims, spacings, boundary kinds, solid cells.
std::uint64_t signature(const UniformGrid& g) {
  std::uint64_t h = 1469598103934665603ull;
  auto mix = [&](std::uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  mix(std::uint64_t(g.nx));
  mix(std::uint64_t(g.ny));
  mix(std::uint64_t(g.nz));
  for (double d : {g.dx, g.dy, g.dz}) {
    std::uint64_t b;
    std::memcpy(&b, &d, 8);
    mix(b);
  }
  for (const auto& f : g.face) mix(std::uint64_t(f.type));
  for (std::size_t i = 0; i < g.solid.size(); ++i)
    if (g.solid[i]) mix(i);
  return h;
}
}  // namespace

DevicePoisson::DevicePoisson(int kind, double tol, int maxIters, int pre, int post, Precision prec)
    : kind_(kind), tol_(tol), maxIters_(maxIters), pre_(pre), post_(post), prec_(prec) {}

DevicePoisson::~DevicePoisson() {
  if (ps_) nbl_poisson_destroy(ps_);
}

nbl_solver_opts DevicePoisson::opts() const {
  nbl_solver_opts o;
  nbl_solver_opts_default(&o);
  o.precision = static_cast<int>(prec_);
  o.pre_smooth = pre_;
  o.post_smooth = post_;
  return o;
}

PoissonSolver::Report DevicePoisson::solve(const UniformGrid& grid, const std::vector<double>& b,
                                           std::vector<double>& x) {
  if (b.size() != grid.cellCount()) throw std::invalid_argument("PoissonSolver::solve: b size mismatch");
  UniformGrid g = grid;
  if (g.solid.empty()) g.allocate();
  const std::uint64_t sig = signature(g);
  if (!ps_ || sig != sig_) {
    if (ps_) nbl_poisson_destroy(ps_);
    ps_ = nullptr;
    const nbl_grid_desc d = g.desc();
    const nbl_solver_opts o = opts();
    nblCheck(nbl_poisson_create(deviceContext(), &d, kind_, tol_, maxIters_, &o, &ps_), "poisson create");
    sig_ = sig;
  }
  x.assign(grid.cellCount(), 0.0);
  nbl_poisson_report r{};
  nblCheck(nbl_poisson_solve(ps_, b.data(), x.data(), &r), "poisson solve");
  return Report{r.iterations, r.residual, r.converged != 0};
}

NSSolver::NSSolver(UniformGrid grid, FlowConfig config, std::shared_ptr<PoissonSolver> poisson)
    : grid_(std::move(grid)), cfg_(config) {
  nu_ = cfg_.velocityScale * cfg_.refLength / cfg_.reynolds;
  if (grid_.solid.empty()) grid_.allocate();
  nbl_flow_cfg c{};
  nbl_flow_cfg_default(&c);
  c.reynolds = cfg_.reynolds;
  c.ref_length = cfg_.refLength;
  c.velocity_scale = cfg_.velocityScale;
  c.density = cfg_.density;
  c.cfl = cfg_.cfl;
  c.body_force[0] = cfg_.bodyForce.x;
  c.body_force[1] = cfg_.bodyForce.y;
  c.body_force[2] = cfg_.bodyForce.z;
  c.convection = static_cast<int>(cfg_.convection);
  c.poisson_tol = cfg_.poissonTol;
  c.poisson_max_iters = cfg_.poissonMaxIters;
  c.tol_continuity = cfg_.tolContinuity;
  c.tol_mass = cfg_.tolMass;
  c.ref_area = cfg_.refArea;
  nbl_solver_opts o;
  nbl_solver_opts_default(&o);
  if (poisson) {
    auto* dp = dynamic_cast<DevicePoisson*>(poisson.get());
    if (dp && dp->name() == "mg-pcg") {
      c.poisson_tol = dp->tol();
      c.poisson_max_iters = dp->maxIters();
      o = dp->opts();
    } else {
      hostPoisson_ = std::move(poisson);
    }
  }
  const nbl_grid_desc d = grid_.desc();
  nblCheck(nbl_run_create(deviceContext(), &d, &c, &o, &run_), "NSSolver");
  if (hostPoisson_) nblCheck(nbl_run_set_pressure_solver(run_, &NSSolver::hostSolve, this), "NSSolver");
}

// The projection's pressure solve through a PoissonSolver outside the device
// path (called back by the run with the right-hand side on the host).
int NSSolver::hostSolve(void* self, const double* b, double* x, std::size_t n, nbl_poisson_report* rep) {
  auto* s = static_cast<NSSolver*>(self);
  try {
    const std::vector<double> bv(b, b + n);
    std::vector<double> xv;
    const PoissonSolver::Report r = s->hostPoisson_->solve(s->grid_, bv, xv);
    if (xv.size() != n) return 1;
    std::copy(xv.begin(), xv.end(), x);
    rep->iterations = r.iterations;
    rep->residual = r.residual;
    rep->converged = r.converged ? 1 : 0;
    return 0;
  } catch (...) {
    return 1;  // the run reports the failure; nothing is thrown across the C-ABI
  }
}

// CgPoisson's statics on a temporary device solver for the grid.
namespace {
struct TmpCg {
  nbl_poisson* ps = nullptr;
  explicit TmpCg(const UniformGrid& grid) {
    UniformGrid g = grid;
    if (g.solid.empty()) g.allocate();
    const nbl_grid_desc d = g.desc();
    nblCheck(nbl_poisson_create(deviceContext(), &d, 1, 1e-10, 1, nullptr, &ps), "CgPoisson");
  }
  ~TmpCg() {
    if (ps) nbl_poisson_destroy(ps);
  }
};
}  // namespace

void HipCgPoisson::applyA(const UniformGrid& grid, const std::vector<double>& x, std::vector<double>& out) {
  if (x.size() != grid.cellCount()) throw std::invalid_argument("CgPoisson::applyA: x size mismatch");
  TmpCg t(grid);
  out.assign(grid.cellCount(), 0.0);
  nblCheck(nbl_poisson_apply(t.ps, x.data(), out.data()), "CgPoisson::applyA");
}

bool HipCgPoisson::hasDirichlet(const UniformGrid& grid) {  // poisson.cpp:93-100
  for (const FaceBC& f : grid.face)
    if (f.type == BoundaryType::Outlet) return true;
  return false;
}

void HipCgPoisson::buildDiagonal(const UniformGrid& grid, std::vector<double>& diag) {
  TmpCg t(grid);
  diag.assign(grid.cellCount(), 1.0);
  nblCheck(nbl_poisson_diagonal(t.ps, diag.data()), "CgPoisson::buildDiagonal");
}

NSSolver::~NSSolver() {
  if (run_) nbl_run_destroy(run_);
}

void NSSolver::initialize() {
  nblCheck(nbl_run_initialize(run_), "initialize");
  step_ = 0;
  t_ = 0.0;
  dtLast_ = 0.0;
  fresh_ = false;
}

static StepReport toReport(const nbl_step_report& r) {
  StepReport s;
  s.step = r.step;
  s.t = r.t;
  s.dt = r.dt;
  s.cfl = r.cfl;
  s.momentumResidual = r.momentum_residual;
  s.continuityResidual = r.continuity_residual;
  s.divergenceMax = r.divergence_max;
  s.massError = r.mass_error;
  s.poissonIters = r.poisson_iters;
  s.poissonResidual = r.poisson_residual;
  s.cd = r.cd;
  s.cl = r.cl;
  s.cellCount = std::size_t(r.cell_count);
  s.accepted = r.accepted != 0;
  return s;
}

StepReport NSSolver::step() { return step(0.0); }

StepReport NSSolver::step(double dt) {
  nbl_step_report r{};
  nblCheck(nbl_run_step(run_, dt, &r), "step");
  step_ = r.step;
  t_ = r.t;
  dtLast_ = r.dt;
  fresh_ = false;
  return toReport(r);
}

void NSSolver::fetch() const {
  if (fresh_) return;
  const std::size_t nx = grid_.nx, ny = grid_.ny, nz = grid_.nz;
  u_.resize((nx + 1) * ny * nz);
  v_.resize(nx * (ny + 1) * nz);
  w_.resize(nx * ny * (nz + 1));
  p_.resize(nx * ny * nz);
  nblCheck(nbl_run_get_state(run_, u_.data(), v_.data(), w_.data(), p_.data(), nullptr, nullptr, nullptr),
           "state");
  fresh_ = true;
}

const std::vector<double>& NSSolver::uFace() const {
  fetch();
  return u_;
}
const std::vector<double>& NSSolver::vFace() const {
  fetch();
  return v_;
}
const std::vector<double>& NSSolver::wFace() const {
  fetch();
  return w_;
}
const std::vector<double>& NSSolver::pressure() const {
  fetch();
  return p_;
}

double NSSolver::cellU(int i, int j, int k) const {
  fetch();
  const std::size_t a = std::size_t(i) + std::size_t(grid_.nx + 1) * (j + std::size_t(grid_.ny) * k);
  return 0.5 * (u_[a] + u_[a + 1]);
}
double NSSolver::cellV(int i, int j, int k) const {
  fetch();
  const std::size_t a = std::size_t(i) + std::size_t(grid_.nx) * (j + std::size_t(grid_.ny + 1) * k);
  return 0.5 * (v_[a] + v_[a + grid_.nx]);
}
double NSSolver::cellW(int i, int j, int k) const {
  if (grid_.nz == 1) return 0.0;
  fetch();
  // w faces: nx * ny per plane, plane k + 1 is the face above cell (i, j, k)
  const std::size_t a = grid_.cidx(i, j, k);
  return 0.5 * (w_[a] + w_[a + std::size_t(grid_.nx) * grid_.ny]);
}

std::vector<double> NSSolver::cellVelocity(int component) const {
  std::vector<double> out(grid_.cellCount());
  nblCheck(nbl_run_cell_velocity(run_, component, out.data()), "cellVelocity");
  return out;
}

void NSSolver::setState(std::vector<double> u, std::vector<double> v, std::vector<double> w,
                        std::vector<double> p, int step, double t, double dtLast) {
  const int rc = nbl_run_set_state(run_, u.data(), u.size(), v.data(), v.size(), w.data(), w.size(), p.data(),
                                   p.size(), step, t, dtLast);
  if (rc == NBL_ERR_INVALID) throw std::runtime_error("NSSolver::setState: size mismatch");
  nblCheck(rc, "setState");
  step_ = step;
  t_ = t;
  dtLast_ = dtLast;
  u_ = std::move(u);
  v_ = std::move(v);
  w_ = std::move(w);
  p_ = std::move(p);
  fresh_ = true;
}

double NSSolver::computeDt() const {
  double dt = 0.0;
  nblCheck(nbl_run_compute_dt(run_, &dt), "computeDt");
  return dt;
}

}  // namespace nabla::flow
