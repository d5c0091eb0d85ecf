// Device 3-D Laplace basis solver (ig_solve.hip), called by the engine's
// ig_solve3d after argument checks.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "ion_gym_amd.h"

namespace ig {
// Runs the solve on stream st; returns 0 or an IG_E_* code with err set.
int solve3d_run(hipStream_t st, const ig_solve3d_desc *d, double *phi, ig_solve3d_result *res,
                std::string &err);
} // namespace ig
