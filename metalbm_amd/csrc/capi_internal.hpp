// capi_internal.hpp -- helpers shared by the C-ABI translation units.
#pragma once

#include <string>

#include "../../include/metalbm_amd.h"

namespace lbm {

// Records `msg` as the handle's last error (or the creation error when h is
// NULL) and returns `code`.
int record_error(lbm_handle* h, int code, const std::string& msg);

}  // namespace lbm
