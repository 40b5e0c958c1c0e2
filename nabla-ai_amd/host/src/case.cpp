// Case builders, case.json loading, the run loop and the three output formats
// (reference core/src/flow/case.cpp): diagnostics.jsonl lines, NBFCK01
// checkpoints and ASCII VTU snapshots are written byte-for-byte in the
// reference's layout.  Optional keys beyond the reference's set
// ("poisson_tol", "poisson_max_iters", "precision", "init", the
// "taylor-green" and "cube-channel" case types) default to reference
// behaviour when absent.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <memory>
#include <stdexcept>

#include "json.hpp"
#include "nabla/flow/case.hpp"
#include "nabla/flow/device.hpp"

#ifndef NBL_GIT_SHA
#define NBL_GIT_SHA "unknown"
#endif

namespace nabla::flow {
namespace {

constexpr const char* kVersion = "0.1.0";

void allFaces(UniformGrid& g, BoundaryType t) {
  for (auto& f : g.face) f.type = t;
}

using Init = CaseInit::Kind;
using InitSpec = CaseInit;

std::uint64_t splitmix(std::uint64_t& s) {
  std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

CaseSpec makeLidCavity(int n, double reynolds, double lidVelocity) {  // case.cpp:25-50
  CaseSpec s;
  s.name = "lid-cavity";
  UniformGrid& g = s.grid;
  g.nx = n;
  g.ny = n;
  g.nz = 1;
  g.dx = g.dy = g.dz = 1.0 / n;
  allFaces(g, BoundaryType::Wall);
  g.face[4].type = g.face[5].type = BoundaryType::Periodic;
  g.face[3].wallVelocity = {lidVelocity, 0.0, 0.0};
  g.allocate();
  s.flow.reynolds = reynolds;
  s.flow.refLength = 1.0;
  s.flow.velocityScale = lidVelocity;
  s.flow.cfl = 0.8;
  s.flow.convection = ConvectionScheme::Central;
  s.run.maxSteps = 60000;
  s.run.steadyTol = 1e-6;
  return s;
}

CaseSpec makeChannel(int nx, int ny, double reynolds, double bulkVelocity) {  // case.cpp:52-84
  CaseSpec s;
  s.name = "channel";
  UniformGrid& g = s.grid;
  const double ly = 1.0, lx = 6.0 * ly;
  g.nx = nx;
  g.ny = ny;
  g.nz = 1;
  g.dx = lx / nx;
  g.dy = ly / ny;
  g.dz = g.dy;
  g.face[0].type = BoundaryType::Inlet;
  g.face[0].profile = InletProfile::Parabolic;
  g.face[0].inletSpeed = bulkVelocity;
  g.face[0].inletAxis = 0;
  g.face[1].type = BoundaryType::Outlet;
  g.face[2].type = g.face[3].type = BoundaryType::Wall;
  g.face[4].type = g.face[5].type = BoundaryType::Periodic;
  g.allocate();
  s.flow.reynolds = reynolds;
  s.flow.refLength = ly;
  s.flow.velocityScale = bulkVelocity;
  s.flow.cfl = 0.8;
  s.flow.convection = ConvectionScheme::Central;
  s.run.maxSteps = 40000;
  s.run.steadyTol = 1e-6;
  return s;
}

static void markCube(UniformGrid& g, double x0, double x1, double y0, double y1, double z0, double z1) {
  for (int k = 0; k < g.nz; ++k)
    for (int j = 0; j < g.ny; ++j)
      for (int i = 0; i < g.nx; ++i) {
        const V3 c = g.cellCenter(i, j, k);
        if (c.x >= x0 && c.x <= x1 && c.y >= y0 && c.y <= y1 && c.z >= z0 && c.z <= z1) g.solid[g.cidx(i, j, k)] = 1;
      }
}

CaseSpec makeWallMountedCube(int resolution, double reynolds) {  // case.cpp:86-134
  CaseSpec s;
  s.name = "wall-mounted-cube";
  const double h = 1.0;
  UniformGrid& g = s.grid;
  g.dx = g.dy = g.dz = h / resolution;
  g.nx = int(std::lround(14.0 * resolution));
  g.ny = int(std::lround(3.0 * resolution));
  g.nz = int(std::lround(6.4 * resolution));
  g.face[0].type = BoundaryType::Inlet;
  g.face[0].profile = InletProfile::Uniform;
  g.face[0].inletSpeed = 1.0;
  g.face[0].inletAxis = 0;
  g.face[1].type = BoundaryType::Outlet;
  g.face[2].type = g.face[3].type = BoundaryType::Wall;
  g.face[4].type = g.face[5].type = BoundaryType::Periodic;
  g.allocate();
  const double zc = 3.2 * h;
  markCube(g, 3.0 * h, 4.0 * h, 0.0, h, zc - 0.5 * h, zc + 0.5 * h);
  s.flow.reynolds = reynolds;
  s.flow.refLength = h;
  s.flow.velocityScale = 1.0;
  s.flow.refArea = h * h;
  s.flow.cfl = 0.7;
  s.flow.convection = ConvectionScheme::Weno5;
  s.run.maxSteps = 200;
  s.run.snapshotEvery = 50;
  s.run.checkpointEvery = 50;
  return s;
}

// Taylor-Green box (n^3, L = 2 pi, all periodic) and the cube channel
// (nx x ny x nz, h = 1 with `res` cells per h): the measured workloads.
static CaseSpec makeTaylorGreen(int n, double reynolds) {
  CaseSpec s;
  s.name = "taylor-green";
  UniformGrid& g = s.grid;
  g.nx = g.ny = g.nz = n;
  g.dx = g.dy = g.dz = 2.0 * M_PI / n;
  allFaces(g, BoundaryType::Periodic);
  g.allocate();
  s.flow.reynolds = reynolds;
  s.flow.cfl = 0.5;
  s.flow.convection = ConvectionScheme::Weno5;
  s.flow.poissonTol = 1e-8;
  s.run.maxSteps = 100;
  return s;
}

static CaseSpec makeCubeChannel(int nx, int ny, int nz, int res, double reynolds) {
  CaseSpec s;
  s.name = "cube-channel";
  const double h = 1.0;
  UniformGrid& g = s.grid;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.dx = g.dy = g.dz = h / res;
  g.face[0].type = BoundaryType::Inlet;
  g.face[0].inletSpeed = 1.0;
  g.face[1].type = BoundaryType::Outlet;
  g.face[2].type = g.face[3].type = BoundaryType::Wall;
  g.face[4].type = g.face[5].type = BoundaryType::Periodic;
  g.allocate();
  const double zc = 0.5 * nz * g.dz;
  markCube(g, 3.0 * h, 4.0 * h, 0.0, h, zc - 0.5 * h, zc + 0.5 * h);
  s.flow.reynolds = reynolds;
  s.flow.refLength = h;
  s.flow.refArea = h * h;
  s.flow.cfl = 0.7;
  s.flow.convection = ConvectionScheme::Weno5;
  s.flow.poissonTol = 1e-8;
  s.run.maxSteps = 500;
  return s;
}

CaseSpec loadCase(const std::string& jsonPath) {  // case.cpp:136-170
  const json::Value j = json::parseFile(jsonPath);
  const std::string type = j.string("type", "lid-cavity");
  const double Re = j.number("reynolds", 100.0);
  const int res = j.integer("resolution", 64);
  CaseSpec s;
  if (type == "lid-cavity") {
    s = makeLidCavity(res, Re, j.number("lid_velocity", 1.0));
  } else if (type == "channel" || type == "poiseuille") {
    s = makeChannel(j.integer("nx", 6 * res), res, Re, j.number("bulk_velocity", 1.0));
  } else if (type == "wall-mounted-cube") {
    s = makeWallMountedCube(res, Re);
  } else if (type == "taylor-green") {
    s = makeTaylorGreen(res, Re);
    s.init.kind = Init::TaylorGreen;
  } else if (type == "cube-channel") {
    s = makeCubeChannel(j.integer("nx", 12 * res), j.integer("ny", 4 * res), j.integer("nz", 4 * res), res, Re);
    s.init.kind = Init::Perturbed;
    s.init.amp = 0.01;
    s.init.seed = 20240601;
    if (const json::Value* pz = j.find("perturbation")) {
      s.init.amp = pz->number("amp", s.init.amp);
      s.init.seed = std::uint64_t(pz->number("seed", double(s.init.seed)));
    }
  } else {
    throw std::runtime_error("loadCase: unknown case type '" + type + "'");
  }
  s.name = j.string("name", type);
  s.flow.cfl = j.number("cfl", s.flow.cfl);
  if (const json::Value* c = j.find("convection")) {
    if (c->kind == json::Value::Kind::String && c->str == "weno5") s.flow.convection = ConvectionScheme::Weno5;
    else if (c->kind == json::Value::Kind::String && c->str == "central") s.flow.convection = ConvectionScheme::Central;
  }
  s.flow.poissonTol = j.number("poisson_tol", s.flow.poissonTol);
  s.flow.poissonMaxIters = j.integer("poisson_max_iters", s.flow.poissonMaxIters);
  const std::string prec = j.string("precision", "fp64");
  if (prec == "mixed") s.precision = Precision::Mixed;
  else if (prec == "mixed-bf16") s.precision = Precision::MixedBF16;
  else if (prec != "fp64") throw std::runtime_error("loadCase: unknown precision '" + prec + "'");
  s.run.runDir = j.string("run_dir", "run_" + type);
  s.run.maxSteps = j.integer("max_steps", s.run.maxSteps);
  s.run.snapshotEvery = j.integer("snapshot_every", s.run.snapshotEvery);
  {
    const std::string vf = j.string("vtu_format", "ascii");
    if (vf == "binary") s.run.vtuFormat = VtuFormat::Binary;
    else if (vf != "ascii") throw std::invalid_argument("vtu_format must be \"ascii\" or \"binary\"");
  }
  s.run.checkpointEvery = j.integer("checkpoint_every", s.run.checkpointEvery);
  s.run.stopTime = j.number("stop_time", s.run.stopTime);
  s.run.steadyTol = j.number("steady_tol", s.run.steadyTol);
  s.run.restartFrom = j.string("restart_from", "");
  return s;
}

// ---- NBFCK01 checkpoint (case.cpp:247-323) ---------------------------------
namespace {
const char kMagic[8] = {'N', 'B', 'F', 'C', 'K', '0', '1', '\0'};
template <class T>
void put(std::ofstream& f, const T& v) {
  f.write(reinterpret_cast<const char*>(&v), sizeof(T));
}
template <class T>
void get(std::ifstream& f, T& v) {
  f.read(reinterpret_cast<char*>(&v), sizeof(T));
}
void putVec(std::ofstream& f, const std::vector<double>& a) {
  const std::uint64_t n = a.size();
  put(f, n);
  if (n) f.write(reinterpret_cast<const char*>(a.data()), std::streamsize(n * 8));
}
std::vector<double> getVec(std::ifstream& f) {
  std::uint64_t n = 0;
  get(f, n);
  std::vector<double> a(n);
  if (n) f.read(reinterpret_cast<char*>(a.data()), std::streamsize(n * 8));
  return a;
}
}  // namespace

void writeCheckpoint(const NSSolver& solver, const std::string& path) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("writeCheckpoint: cannot open " + path);
  f.write(kMagic, 8);
  const UniformGrid& g = solver.grid();
  put(f, g.nx);
  put(f, g.ny);
  put(f, g.nz);
  const int step = solver.stepIndex();
  const double t = solver.time(), dt = solver.dtLast();
  put(f, step);
  put(f, t);
  put(f, dt);
  putVec(f, solver.uFace());
  putVec(f, solver.vFace());
  putVec(f, solver.wFace());
  putVec(f, solver.pressure());
}

void readCheckpoint(NSSolver& solver, const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("readCheckpoint: cannot open " + path);
  char magic[8];
  f.read(magic, 8);
  if (std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error("readCheckpoint: bad magic in " + path);
  int nx, ny, nz, step;
  double t, dt;
  get(f, nx);
  get(f, ny);
  get(f, nz);
  const UniformGrid& g = solver.grid();
  if (nx != g.nx || ny != g.ny || nz != g.nz) throw std::runtime_error("readCheckpoint: grid dimension mismatch");
  get(f, step);
  get(f, t);
  get(f, dt);
  std::vector<double> u = getVec(f), v = getVec(f), w = getVec(f), p = getVec(f);
  solver.setState(std::move(u), std::move(v), std::move(w), std::move(p), step, t, dt);
}

// ---- run loop (case.cpp:332-401) --------------------------------------------
namespace {
void appendMeta(std::ofstream& f) {
  f << "{\"event\":\"meta\",\"solver\":\"nabla_solve\",\"solver_version\":\"" << kVersion << "\",\"git_sha\":\""
    << NBL_GIT_SHA << "\"}\n";
}
void appendDiagnostics(std::ofstream& f, const StepReport& r, bool accepted) {
  f << "{\"step\":" << r.step << ",\"t\":" << r.t << ",\"dt\":" << r.dt << ",\"cfl\":" << r.cfl
    << ",\"momentum_residual\":" << r.momentumResidual << ",\"continuity_residual\":" << r.continuityResidual
    << ",\"divergence_max\":" << r.divergenceMax << ",\"mass_error\":" << r.massError << ",\"cd\":" << r.cd
    << ",\"cl\":" << r.cl << ",\"poisson_iters\":" << r.poissonIters << ",\"poisson_residual\":" << r.poissonResidual
    << ",\"cells\":" << r.cellCount << ",\"accepted\":" << (accepted ? "true" : "false") << "}\n";
}

void applyInit(NSSolver& s, const InitSpec& in) {
  const UniformGrid& g = s.grid();
  const std::size_t nx = g.nx, ny = g.ny, nz = g.nz;
  std::vector<double> u((nx + 1) * ny * nz, 0.0), v(nx * (ny + 1) * nz, 0.0), w(nx * ny * (nz + 1), 0.0),
      p(nx * ny * nz, 0.0);
  if (in.kind == Init::TaylorGreen) {
    for (std::size_t k = 0; k < nz; ++k)
      for (std::size_t j = 0; j < ny; ++j)
        for (std::size_t i = 0; i <= nx; ++i) {
          const double x = double(i % nx) * g.dx, y = (j + 0.5) * g.dy, z = (k + 0.5) * g.dz;
          u[i + (nx + 1) * (j + ny * k)] = std::sin(x) * std::cos(y) * std::cos(z);
        }
    for (std::size_t k = 0; k < nz; ++k)
      for (std::size_t j = 0; j <= ny; ++j)
        for (std::size_t i = 0; i < nx; ++i) {
          const double x = (i + 0.5) * g.dx, y = double(j % ny) * g.dy, z = (k + 0.5) * g.dz;
          v[i + nx * (j + (ny + 1) * k)] = -std::cos(x) * std::sin(y) * std::cos(z);
        }
  } else {
    std::uint64_t st = in.seed;
    for (std::size_t k = 0; k < nz; ++k)
      for (std::size_t j = 0; j < ny; ++j) {
        for (std::size_t i = 0; i <= nx; ++i) u[i + (nx + 1) * (j + ny * k)] = 1.0;
        for (std::size_t i = 1; i < nx; ++i) {
          const double xi = 2.0 * double(splitmix(st) >> 11) * 0x1.0p-53 - 1.0;
          u[i + (nx + 1) * (j + ny * k)] += in.amp * xi;
        }
        for (std::size_t i = 0; i <= nx; ++i) {
          const bool touch = (i > 0 && g.isSolid(int(i) - 1, int(j), int(k))) || (i < nx && g.isSolid(int(i), int(j), int(k)));
          if (touch) u[i + (nx + 1) * (j + ny * k)] = 0.0;
        }
        u[nx + (nx + 1) * (j + ny * k)] = u[nx - 1 + (nx + 1) * (j + ny * k)];
      }
  }
  s.setState(std::move(u), std::move(v), std::move(w), std::move(p), 0, 0.0, 0.0);
}
}  // namespace

RunSummary runCase(const CaseSpec& spec) {
  const InitSpec& init = spec.init;
  auto poisson = std::make_shared<HipMgPoisson>(spec.flow.poissonTol, spec.flow.poissonMaxIters, 2, 2, spec.precision);
  NSSolver solver(spec.grid, spec.flow, poisson);
  solver.initialize();
  if (init.kind != Init::Default) applyInit(solver, init);
  std::filesystem::create_directories(spec.run.runDir);
  const bool restart = !spec.run.restartFrom.empty();
  if (restart) readCheckpoint(solver, spec.run.restartFrom);
  const std::string diagPath = spec.run.runDir + "/diagnostics.jsonl";
  std::ofstream diag(diagPath, restart ? std::ios::app : std::ios::trunc);
  if (!restart) appendMeta(diag);
  RunSummary sum;
  sum.runDir = spec.run.runDir;
  sum.cells = solver.grid().cellCount();
  StepReport last;
  // snapshots are formatted and written on a thread of their own while the
  // steps go on; the host copy of the fields is taken at the step itself
  SnapshotWriter vtu(spec.run.vtuFormat);
  for (int s = 0; s < spec.run.maxSteps; ++s) {
    if (spec.run.cancel && *spec.run.cancel) {
      vtu.drain();
      throw RunCancelled(last.step);
    }
    last = solver.step();
    appendDiagnostics(diag, last, last.accepted);
    diag.flush();
    if (spec.run.snapshotEvery > 0 && last.step % spec.run.snapshotEvery == 0)
      vtu.submit(captureVtu(solver), spec.run.runDir + "/snapshot_" + std::to_string(last.step) + ".vtu");
    if (spec.run.checkpointEvery > 0 && last.step % spec.run.checkpointEvery == 0)
      writeCheckpoint(solver, spec.run.runDir + "/checkpoint_" + std::to_string(last.step) + ".ckpt");
    if (spec.run.stopTime > 0.0 && solver.time() >= spec.run.stopTime) break;
    if (spec.run.steadyTol > 0.0 && last.step > 20 && last.momentumResidual < spec.run.steadyTol) {
      sum.steady = true;
      break;
    }
  }
  vtu.submit(captureVtu(solver), spec.run.runDir + "/final.vtu");
  vtu.drain();
  writeCheckpoint(solver, spec.run.runDir + "/final.ckpt");
  sum.steps = last.step;
  sum.time = last.t;
  sum.momentumResidual = last.momentumResidual;
  sum.continuityResidual = last.continuityResidual;
  sum.massError = last.massError;
  sum.cd = last.cd;
  sum.cl = last.cl;
  return sum;
}

}  // namespace nabla::flow
