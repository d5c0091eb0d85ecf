// Device Laplace basis solvers (ig_solve.hip), called by the engine's
// ig_solve3d and ig_solve2d after argument checks.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "ion_gym_amd.h"

namespace ig {
// Runs the solve on stream st; returns 0 or an IG_E_* code with err set.
int solve3d_run(hipStream_t st, const ig_solve3d_desc *d, double *phi, ig_solve3d_result *res,
                std::string &err);
// The 2-D solve (ig_solve2d) on stream st, the same conventions.
int solve2d_run(hipStream_t st, const ig_solve2d_desc *d, double *phi, ig_solve3d_result *res,
                std::string &err);
} // namespace ig
