// The C ABI's handles shared between its source files.
#pragma once

#include <memory>

#include "model.hpp"
#include "sl3dpp_amd.h"

// An Earth handle: shared, so the operators made from it keep it alive
// after the caller destroys the handle.
struct sl3_earth {
  std::shared_ptr<const sl3amd::Earth> earth;
};
