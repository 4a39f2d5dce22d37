// geo_launch.hpp -- launcher of the general-hexahedra stage kernel
// (geo_kernel.hpp), whose instances live in geo_inst.hip.
#pragma once
#include <hip/hip_runtime.h>

namespace nkc {

struct GeoArgs;

// Orders (N + 1) the general kernel is built for.
constexpr int GEO_MIN_N1 = 4, GEO_MAX_N1 = 12;

// Launch one stage (mode: GMODE_RK, GMODE_RESIDUAL, GMODE_CURL) over grid
// workgroups of GEO_BS lanes; false if the order is not built.
bool launch_geo(const GeoArgs &a, int n1, int mode, int grid, hipStream_t s);

}  // namespace nkc
