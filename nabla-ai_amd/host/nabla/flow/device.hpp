// Shared device context for the host layer and the C-ABI status -> exception
// mapping (the reference throws std::runtime_error / std::invalid_argument).
#pragma once

#include <stdexcept>
#include <string>

#include "nbl_api.h"

namespace nabla::flow {

inline void nblCheck(int rc, const char* what) {
  if (rc == NBL_OK) return;
  const std::string msg = std::string(what) + ": " + nbl_last_error();
  if (rc == NBL_ERR_INVALID) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// Process-wide context on the device selected by NABLA_DEVICE (default 0).
nbl_ctx* deviceContext();

}  // namespace nabla::flow
