// rh_api_internal.hpp -- library-internal hooks of rh_api.cpp used by the
// other translation units of the C ABI (not part of include/radhip.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/radhip.h"

namespace rh_internal {

// Records msg as this thread's rh_last_error() text and returns st.
rh_status fail(rh_status st, const std::string& msg);
// Device the model lives on.
int model_device(const rh_model* m);
// Makes the device current and clears errors left pending by other callers.
hipError_t enter_device(int device);
// rh_field_dev with the partial sums in the caller's `work` buffer of
// work_cap doubles: the element splits are reduced until the partials fit;
// RH_ERR_ARG, before any launch, when they cannot.  Allocates nothing.
rh_status field_dev_into(rh_model* m, const double* d_M, const double* d_pts, int64_t np, double* d_B,
                         double* d_H, rh_precision prec, double* work, int64_t work_cap);

}  // namespace rh_internal
