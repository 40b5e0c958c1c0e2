// Case description, run driver and the reference's file formats
// (core/include/nabla/flow/case.hpp:12-57): case.json in; diagnostics.jsonl,
// NBFCK01 checkpoints and VTU snapshots out, byte-compatible.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "nabla/flow/grid.hpp"
#include "nabla/flow/ns_solver.hpp"

namespace nabla::flow {

// VTU encoding of snapshots ("vtu_format": "ascii", the reference's, or
// "binary": VTK appended raw data, the same arrays).
enum class VtuFormat { Ascii, Binary };

struct RunOptions {
  std::string runDir = "run";
  int maxSteps = 20000;
  double stopTime = 0.0;
  double steadyTol = 0.0;
  int snapshotEvery = 0;
  int checkpointEvery = 0;
  std::string restartFrom;
  bool quiet = false;
  VtuFormat vtuFormat = VtuFormat::Ascii;
  // Set to non-zero by another thread to stop the run between steps (the
  // in-process job runner's cancel); runCase then throws RunCancelled.
  const volatile int* cancel = nullptr;
};

struct RunCancelled : std::runtime_error {
  explicit RunCancelled(int step) : std::runtime_error("cancelled after step " + std::to_string(step)) {}
};

// Initial state of the measured case types, applied through setState after
// initialize(): the analytic Taylor-Green field, or the channel's uniform
// stream with a seeded uniform perturbation of the interior u-faces.
struct CaseInit {
  enum class Kind { Default, TaylorGreen, Perturbed };
  Kind kind = Kind::Default;
  double amp = 0.0;
  std::uint64_t seed = 0;
};

struct CaseSpec {
  std::string name = "case";
  UniformGrid grid;
  FlowConfig flow;
  RunOptions run;
  Precision precision = Precision::FP64;  // optional "precision" key
  CaseInit init;                          // "taylor-green" / "cube-channel" initial states
};

struct RunSummary {
  int steps = 0;
  double time = 0.0;
  double momentumResidual = 0.0;
  double continuityResidual = 0.0;
  double massError = 0.0;
  double cd = 0.0;
  double cl = 0.0;
  std::size_t cells = 0;
  bool steady = false;
  std::string runDir;
};

CaseSpec makeLidCavity(int n, double reynolds, double lidVelocity = 1.0);
CaseSpec makeChannel(int nx, int ny, double reynolds, double bulkVelocity = 1.0);
CaseSpec makeWallMountedCube(int resolution, double reynolds);

CaseSpec loadCase(const std::string& jsonPath);
RunSummary runCase(const CaseSpec& spec);

void writeGridVtu(const NSSolver& solver, const std::string& path);

// The cell fields one VTU snapshot prints, copied to the host.
struct VtuSnapshot {
  UniformGrid grid;
  std::vector<double> u, v, w, p;
};
VtuSnapshot captureVtu(const NSSolver& solver);
void writeVtu(const VtuSnapshot& s, const std::string& path, VtuFormat fmt);

// Writes submitted snapshots in order on a thread of its own; drain() waits
// for all of them and rethrows the first write error.
class SnapshotWriter {
 public:
  explicit SnapshotWriter(VtuFormat fmt);
  ~SnapshotWriter();
  SnapshotWriter(const SnapshotWriter&) = delete;
  SnapshotWriter& operator=(const SnapshotWriter&) = delete;
  void submit(VtuSnapshot s, std::string path);
  void drain();

 private:
  struct Impl;
  std::unique_ptr<Impl> impl_;
};
void writeCheckpoint(const NSSolver& solver, const std::string& path);
void readCheckpoint(NSSolver& solver, const std::string& path);

}  // namespace nabla::flow
