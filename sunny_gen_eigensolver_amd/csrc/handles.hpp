// Opaque handle types of the C ABI (include/sunny_geneig.h) and the helpers
// the ABI's translation units share.
#pragma once

#include "sunny_geneig.h"

#include "serving.hpp"
#include "solver.hpp"
#include "varlen.hpp"

#include <memory>
#include <string>

struct sunny_geneig_solver {
    std::unique_ptr<sgeig::Solver> impl;
    sunny_geneig_kernel_mode_t mode = SUNNY_GENEIG_KERNEL_UNIFIED;
};

struct sunny_geneig_varlen_solver {
    std::unique_ptr<sgeig::VarlenSolver> impl;
};

struct sunny_geneig_session {
    std::unique_ptr<sgeig::ServingSession> impl;
};

namespace sgeig {
// The calling thread's sunny_geneig_last_error() message.
void api_set_error(const std::string& m);
void api_clear_error();
// Resolve the no-allocation launch's device scratch for `device` (which must
// be current); once per device, later calls return at once. Failures are
// remembered and reported by the entry point.
void noalloc_setup(int device);
}  // namespace sgeig
