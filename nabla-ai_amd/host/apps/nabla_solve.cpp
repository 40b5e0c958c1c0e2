// nabla_solve — the CLI process boundary (reference core/apps/nabla_solve/main.cpp:131-175):
//   nabla_solve run <case.json> [--restart <checkpoint>]
// exit 0 on success, 1 with "error: ..." on stderr, 2 on usage errors.
#include <exception>
#include <iostream>
#include <string>
#include <string_view>

#include "nabla/flow/case.hpp"

namespace {

int runCli(int argc, char** argv) {
  std::string caseFile, restart;
  for (int i = 2; i < argc; ++i) {
    const std::string_view a = argv[i];
    if (a == "--restart") {
      if (i + 1 >= argc) {
        std::cerr << "error: missing value for --restart\n";
        return 2;
      }
      restart = argv[++i];
    } else if (!a.empty() && a.front() == '-') {
      std::cerr << "error: unknown run option '" << a << "'\n";
      return 2;
    } else {
      caseFile = std::string(a);
    }
  }
  if (caseFile.empty()) {
    std::cerr << "error: run requires a case JSON path\n"
                 "usage: nabla_solve run <case.json> [--restart <checkpoint>]\n";
    return 2;
  }
  try {
    nabla::flow::CaseSpec spec = nabla::flow::loadCase(caseFile);
    if (!restart.empty()) spec.run.restartFrom = restart;
    const nabla::flow::RunSummary s = nabla::flow::runCase(spec);
    std::cout << "Nabla AI run \xE2\x80\x94 case " << spec.name << '\n'
              << "  cells         : " << s.cells << '\n'
              << "  steps         : " << s.steps << "  (t = " << s.time << ")\n"
              << "  momentum res  : " << s.momentumResidual << '\n'
              << "  continuity    : " << s.continuityResidual << '\n'
              << "  mass error    : " << s.massError << '\n'
              << "  Cd, Cl        : " << s.cd << ", " << s.cl << '\n'
              << "  steady        : " << (s.steady ? "yes" : "no") << '\n'
              << "  run dir       : " << s.runDir << "  (diagnostics.jsonl, *.vtu, *.ckpt)\n";
    return 0;
  } catch (const std::exception& ex) {
    std::cerr << "error: " << ex.what() << '\n';
    return 1;
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc >= 2 && std::string_view(argv[1]) == "run") return runCli(argc, argv);
  if (argc >= 2 && (std::string_view(argv[1]) == "--version" || std::string_view(argv[1]) == "version")) {
    std::cout << "nabla_solve (MI355X gfx950 build) 0.1.0\n";
    return 0;
  }
  std::cerr << "usage: nabla_solve run <case.json> [--restart <checkpoint>]\n";
  return 2;
}
