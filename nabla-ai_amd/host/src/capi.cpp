// C entry of the host layer for in-process job runners: one nabla_solve run
// (core/apps/nabla_solve/main.cpp:131-175) on the calling thread, with a
// cancel flag in place of the subprocess's SIGTERM.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <sstream>
#include <string>

#include "nabla/flow/case.hpp"
#include "nbl_host.h"

namespace {
void putMsg(char* msg, size_t len, const std::string& s) {
  if (!msg || len == 0) return;
  const size_t n = std::min(len - 1, s.size());
  std::memcpy(msg, s.data(), n);
  msg[n] = '\0';
}
}  // namespace

extern "C" int nbl_case_run(const char* case_json, const char* restart, const volatile int* cancel, char* msg,
                            size_t msg_len) {
  try {
    if (!case_json || !*case_json) throw std::invalid_argument("run requires a case JSON path");
    nabla::flow::CaseSpec spec = nabla::flow::loadCase(case_json);
    if (restart && *restart) spec.run.restartFrom = restart;
    spec.run.cancel = cancel;
    const nabla::flow::RunSummary s = nabla::flow::runCase(spec);
    std::ostringstream o;  // the CLI's summary
    o << "Nabla AI run \xE2\x80\x94 case " << spec.name << '\n'
      << "  cells         : " << s.cells << '\n'
      << "  steps         : " << s.steps << "  (t = " << s.time << ")\n"
      << "  momentum res  : " << s.momentumResidual << '\n'
      << "  continuity    : " << s.continuityResidual << '\n'
      << "  mass error    : " << s.massError << '\n'
      << "  Cd, Cl        : " << s.cd << ", " << s.cl << '\n'
      << "  steady        : " << (s.steady ? "yes" : "no") << '\n'
      << "  run dir       : " << s.runDir << "  (diagnostics.jsonl, *.vtu, *.ckpt)\n";
    putMsg(msg, msg_len, o.str());
    return 0;
  } catch (const nabla::flow::RunCancelled& ex) {
    putMsg(msg, msg_len, std::string(ex.what()) + "\n");
    return -15;
  } catch (const std::exception& ex) {
    putMsg(msg, msg_len, std::string("error: ") + ex.what() + "\n");
    return 1;
  }
}
