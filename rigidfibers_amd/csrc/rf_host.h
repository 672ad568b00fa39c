// rf_host.h -- host helpers shared by the C-ABI translation units
// (rf_capi.hip, rf_serve.hip).  Not part of the public C ABI.
#pragma once

#include <string>

#include "../../include/rigidfibers.h"
#include "rf_internal.h"

namespace rf {

// Records msg as the calling thread's rf_last_error() and returns code.
rf_status set_error(rf_status code, const std::string& msg);
// The checks rf_create applies to a configuration.
rf_status check_config(const rf_config* c);
// Constants of the operator (slenderness terms, eigenvalues, force) for the kernels.
void physical_constants(const rf_config& c, PhysConsts* pc);

}  // namespace rf
