// ref_shim.cpp — TEST INFRASTRUCTURE ONLY.
//
// extern "C" wrappers over the reference's public C++ API, compiled together
// with the reference's own hot-path sources (see oracle/Makefile) into
// oracle/_ref/libnabla_ref.so.  Used only to pin the oracle restatement and as
// the "reference" CPU baseline; nothing in the product links it.  Function
// signatures mirror oracle.h with a ref_ prefix.
#include <cstring>
#include <memory>
#include <vector>

#include "nabla/adaptive/adaptive.hpp"
#include "nabla/flow/ns_solver.hpp"
#include "nabla/flow/octree_poisson.hpp"
#include "nabla/flow/poisson.hpp"
#include "nabla/mesh/morton.hpp"
#include "nabla/mesh/octree.hpp"
#include "nbl_api.h"

using namespace nabla;

namespace {

flow::UniformGrid toGrid(const nbl_grid_desc* d) {
  flow::UniformGrid g;
  g.nx = d->nx;
  g.ny = d->ny;
  g.nz = d->nz;
  g.origin = {d->origin[0], d->origin[1], d->origin[2]};
  g.dx = d->dx;
  g.dy = d->dy;
  g.dz = d->dz;
  for (int f = 0; f < 6; ++f) {
    const nbl_face_bc& s = d->face[f];
    flow::FaceBC& t = g.face[f];
    t.type = static_cast<flow::BoundaryType>(s.type);
    t.wallVelocity = {s.wall_velocity[0], s.wall_velocity[1], s.wall_velocity[2]};
    t.profile = static_cast<flow::InletProfile>(s.profile);
    t.inletSpeed = s.inlet_speed;
    t.inletAxis = s.inlet_axis;
  }
  g.allocate();
  if (d->solid) std::memcpy(g.solid.data(), d->solid, g.cellCount());
  return g;
}

flow::FlowConfig toCfg(const nbl_flow_cfg* c) {
  flow::FlowConfig f;
  f.reynolds = c->reynolds;
  f.refLength = c->ref_length;
  f.velocityScale = c->velocity_scale;
  f.density = c->density;
  f.cfl = c->cfl;
  f.bodyForce = {c->body_force[0], c->body_force[1], c->body_force[2]};
  f.convection = static_cast<flow::ConvectionScheme>(c->convection);
  f.poissonTol = c->poisson_tol;
  f.poissonMaxIters = c->poisson_max_iters;
  f.tolContinuity = c->tol_continuity;
  f.tolMass = c->tol_mass;
  f.refArea = c->ref_area;
  return f;
}

void toRep(const flow::StepReport& r, nbl_step_report* o) {
  o->step = r.step;
  o->t = r.t;
  o->dt = r.dt;
  o->cfl = r.cfl;
  o->momentum_residual = r.momentumResidual;
  o->continuity_residual = r.continuityResidual;
  o->divergence_max = r.divergenceMax;
  o->mass_error = r.massError;
  o->poisson_iters = r.poissonIters;
  o->poisson_residual = r.poissonResidual;
  o->cd = r.cd;
  o->cl = r.cl;
  o->cell_count = r.cellCount;
  o->accepted = r.accepted ? 1 : 0;
}

template <class R>
void toPrep(const R& r, nbl_poisson_report* o) {
  if (!o) return;
  o->iterations = r.iterations;
  o->residual = r.residual;
  o->converged = r.converged ? 1 : 0;
}

}  // namespace

extern "C" {

void* ref_ns_create(const nbl_grid_desc* d, const nbl_flow_cfg* c) {
  return new flow::NSSolver(toGrid(d), toCfg(c));
}
void ref_ns_destroy(void* h) { delete static_cast<flow::NSSolver*>(h); }
void ref_ns_initialize(void* h) { static_cast<flow::NSSolver*>(h)->initialize(); }
int ref_ns_set_state(void* h, const double* u, size_t nu, const double* v, size_t nv, const double* w,
                     size_t nw, const double* p, size_t np, int step, double t, double dt_last) {
  try {
    static_cast<flow::NSSolver*>(h)->setState(std::vector<double>(u, u + nu), std::vector<double>(v, v + nv),
                                              std::vector<double>(w, w + nw), std::vector<double>(p, p + np),
                                              step, t, dt_last);
    return 0;
  } catch (...) {
    return -1;
  }
}
void ref_ns_get_state(void* h, double* u, double* v, double* w, double* p, int* step, double* t,
                      double* dt_last) {
  const auto& s = *static_cast<flow::NSSolver*>(h);
  if (u) std::memcpy(u, s.uFace().data(), s.uFace().size() * 8);
  if (v) std::memcpy(v, s.vFace().data(), s.vFace().size() * 8);
  if (w) std::memcpy(w, s.wFace().data(), s.wFace().size() * 8);
  if (p) std::memcpy(p, s.pressure().data(), s.pressure().size() * 8);
  if (step) *step = s.stepIndex();
  if (t) *t = s.time();
  if (dt_last) *dt_last = s.dtLast();
}
double ref_ns_compute_dt(void* h) { return static_cast<flow::NSSolver*>(h)->computeDt(); }
void ref_ns_step(void* h, double dt, nbl_step_report* rep) {
  auto& s = *static_cast<flow::NSSolver*>(h);
  const flow::StepReport r = dt > 0.0 ? s.step(dt) : s.step();
  if (rep) toRep(r, rep);
}
void ref_ns_cell_velocity(void* h, int comp, double* out) {
  const auto v = static_cast<flow::NSSolver*>(h)->cellVelocity(comp);
  std::memcpy(out, v.data(), v.size() * 8);
}

void ref_apply_a(const nbl_grid_desc* d, const double* x, double* out) {
  const flow::UniformGrid g = toGrid(d);
  std::vector<double> xx(x, x + g.cellCount()), o;
  flow::CgPoisson::applyA(g, xx, o);
  std::memcpy(out, o.data(), o.size() * 8);
}
void ref_diagonal(const nbl_grid_desc* d, double* diag) {
  const flow::UniformGrid g = toGrid(d);
  std::vector<double> o;
  flow::CgPoisson::buildDiagonal(g, o);
  std::memcpy(diag, o.data(), o.size() * 8);
}
void ref_mg_solve(const nbl_grid_desc* d, const double* b, double* x, double tol, int max_cycles,
                  nbl_poisson_report* rep) {
  const flow::UniformGrid g = toGrid(d);
  flow::MgPoisson mg(tol, max_cycles);
  std::vector<double> bb(b, b + g.cellCount()), xx;
  const auto r = mg.solve(g, bb, xx);
  std::memcpy(x, xx.data(), xx.size() * 8);
  toPrep(r, rep);
}
void ref_cg_solve(const nbl_grid_desc* d, const double* b, double* x, double tol, int max_iters,
                  nbl_poisson_report* rep) {
  const flow::UniformGrid g = toGrid(d);
  flow::CgPoisson cg(tol, max_iters);
  std::vector<double> bb(b, b + g.cellCount()), xx;
  const auto r = cg.solve(g, bb, xx);
  std::memcpy(x, xx.data(), xx.size() * 8);
  toPrep(r, rep);
}

uint64_t ref_morton_encode(uint32_t x, uint32_t y, uint32_t z) { return mesh::mortonEncode(x, y, z); }
void ref_morton_decode(uint64_t key, uint32_t* x, uint32_t* y, uint32_t* z) {
  mesh::mortonDecode(key, *x, *y, *z);
}
void* ref_oct_create(const double o[3], const double e[3], int max_level) {
  try {
    return new mesh::Octree(mesh::Vec3{o[0], o[1], o[2]}, mesh::Vec3{e[0], e[1], e[2]}, max_level);
  } catch (...) {
    return nullptr;
  }
}
void ref_oct_destroy(void* h) { delete static_cast<mesh::Octree*>(h); }
size_t ref_oct_count(void* h) { return static_cast<mesh::Octree*>(h)->cellCount(); }
void ref_oct_refine_uniform(void* h, int times) { static_cast<mesh::Octree*>(h)->refineUniform(times); }
void ref_oct_refine_box(void* h, const double lo[3], const double hi[3], int times) {
  static_cast<mesh::Octree*>(h)->refineBox(mesh::Vec3{lo[0], lo[1], lo[2]}, mesh::Vec3{hi[0], hi[1], hi[2]},
                                           times);
}
int ref_oct_refine_key(void* h, uint64_t key) {
  try {
    static_cast<mesh::Octree*>(h)->refineByMorton(key);
    return 0;
  } catch (...) {
    return -1;
  }
}
int ref_oct_level_of(void* h, uint64_t key) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  const size_t i = t.find(key);
  return i == mesh::Octree::npos ? -1 : t.level(i);
}
void ref_oct_set_mask(void* h, size_t i, uint8_t m) {
  static_cast<mesh::Octree*>(h)->setMask(i, static_cast<mesh::CellMask>(m));
}
void ref_oct_dump(void* h, uint64_t* keys, uint8_t* levels, uint8_t* masks) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  for (size_t i = 0; i < t.cellCount(); ++i) {
    if (keys) keys[i] = t.morton(i);
    if (levels) levels[i] = static_cast<uint8_t>(t.level(i));
    if (masks) masks[i] = static_cast<uint8_t>(t.mask(i));
  }
}
int ref_oct_face_neighbors(void* h, size_t i, int dir, uint64_t* out) {
  const auto v = static_cast<mesh::Octree*>(h)->faceNeighbors(i, dir);
  for (size_t m = 0; m < v.size(); ++m) out[m] = v[m];
  return static_cast<int>(v.size());
}
// All face neighbours at once: cnt[6i+d] neighbours of leaf i across face d,
// their leaf indices in nbr[24i+4d ..], unused slots -1.
void ref_oct_face_table(void* h, uint8_t* cnt, int64_t* nbr) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  for (size_t i = 0; i < t.cellCount(); ++i)
    for (int d = 0; d < 6; ++d) {
      const auto v = t.faceNeighbors(i, d);
      cnt[6 * i + d] = static_cast<uint8_t>(v.size());
      for (int m = 0; m < 4; ++m) nbr[24 * i + 4 * d + m] = m < int(v.size()) ? int64_t(v[m]) : -1;
    }
}
int ref_oct_is_balanced(void* h) { return static_cast<mesh::Octree*>(h)->isBalanced() ? 1 : 0; }
void ref_oct_centers(void* h, double* xyz) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  for (size_t i = 0; i < t.cellCount(); ++i) {
    const mesh::Vec3 c = t.cellCenter(i);
    xyz[3 * i] = c.x;
    xyz[3 * i + 1] = c.y;
    xyz[3 * i + 2] = c.z;
  }
}
void ref_oct_apply(void* h, const double* x, double* out) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  std::vector<double> xx(x, x + t.cellCount()), o;
  flow::OctreePoisson::applyA(t, xx, o);
  std::memcpy(out, o.data(), o.size() * 8);
}
void ref_oct_diagonal(void* h, double* diag) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  std::vector<double> o;
  flow::OctreePoisson::buildDiagonal(t, o);
  std::memcpy(diag, o.data(), o.size() * 8);
}
void ref_oct_solve(void* h, const double* b, double* x, double tol, int max_iters, nbl_poisson_report* rep) {
  const auto& t = *static_cast<mesh::Octree*>(h);
  flow::OctreePoisson ps(tol, max_iters);
  std::vector<double> bb(b, b + t.cellCount()), xx;
  const auto r = ps.solve(t, bb, xx);
  std::memcpy(x, xx.data(), xx.size() * 8);
  toPrep(r, rep);
}

double ref_oct_transport(void* h, const double* u, const double* v, const double* w, double* phi, double dt) {
  auto& t = *static_cast<mesh::Octree*>(h);
  const size_t n = t.cellCount();
  t.u().assign(u, u + n);
  t.v().assign(v, v + n);
  t.w().assign(w, w + n);
  t.registerScalar("phi");
  t.scalar("phi").assign(phi, phi + n);
  const double I = adaptive::octreeTransportStep(t, "phi", dt);
  std::memcpy(phi, t.scalar("phi").data(), n * 8);
  return I;
}

// PhysicalSensorEvaluator::evaluate on the tree's leaves: out holds the eight
// real sensors field-major (vorticity, strain, Q, |grad p|, |grad u|, y+,
// tau_wall, truncation estimate), flags the near-wall, reverse-flow and
// residual-stagnation flags; speed_prev / resid_prev are read and updated.
void ref_oct_sensors(void* h, const double* u, const double* v, const double* w, const double* p, double nu,
                     double* speedPrev, double* residPrev, double* out, uint8_t* flags) {
  auto& t = *static_cast<mesh::Octree*>(h);
  const size_t n = t.cellCount();
  t.u().assign(u, u + n);
  t.v().assign(v, v + n);
  t.w().assign(w, w + n);
  t.p().assign(p, p + n);
  t.registerScalar("speed_prev");
  t.registerScalar("resid_prev");
  t.scalar("speed_prev").assign(speedPrev, speedPrev + n);
  t.scalar("resid_prev").assign(residPrev, residPrev + n);
  const auto rows = adaptive::PhysicalSensorEvaluator(nu).evaluate(t);
  for (size_t i = 0; i < n; ++i) {
    const auto& r = rows[i];
    const double vals[8] = {r.vorticityMag, r.strainMag, r.qCriterion, r.gradPMag,
                            r.gradUMag,     r.yPlus,     r.tauWall,    r.truncErr};
    for (int f = 0; f < 8; ++f) out[size_t(f) * n + i] = vals[f];
    flags[i] = r.nearWall ? 1 : 0;
    flags[n + i] = r.reverseFlow ? 1 : 0;
    flags[2 * n + i] = r.residStagnation ? 1 : 0;
  }
  std::memcpy(speedPrev, t.scalar("speed_prev").data(), n * 8);
  std::memcpy(residPrev, t.scalar("resid_prev").data(), n * 8);
}
}  // extern "C"
