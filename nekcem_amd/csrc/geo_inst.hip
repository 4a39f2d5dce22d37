// geo_inst.hip -- instances of the general-hexahedra stage kernel for
// N = 3..11, a translation unit of their own.
#include "geo_kernel.hpp"
#include "geo_launch.hpp"

namespace nkc {

namespace {
template <int N1>
void launch_geo_n(const GeoArgs &a, int mode, int grid, hipStream_t s) {
    if (mode == GMODE_RK) hipLaunchKernelGGL((geo_stage_kernel<N1, GMODE_RK>), dim3(grid), dim3(GEO_BS), 0, s, a);
    else if (mode == GMODE_RESIDUAL)
        hipLaunchKernelGGL((geo_stage_kernel<N1, GMODE_RESIDUAL>), dim3(grid), dim3(GEO_BS), 0, s, a);
    else hipLaunchKernelGGL((geo_stage_kernel<N1, GMODE_CURL>), dim3(grid), dim3(GEO_BS), 0, s, a);
}
}  // namespace

bool launch_geo(const GeoArgs &a, int n1, int mode, int grid, hipStream_t s) {
    switch (n1) {
    case 4: launch_geo_n<4>(a, mode, grid, s); return true;
    case 5: launch_geo_n<5>(a, mode, grid, s); return true;
    case 6: launch_geo_n<6>(a, mode, grid, s); return true;
    case 7: launch_geo_n<7>(a, mode, grid, s); return true;
    case 8: launch_geo_n<8>(a, mode, grid, s); return true;
    case 9: launch_geo_n<9>(a, mode, grid, s); return true;
    case 10: launch_geo_n<10>(a, mode, grid, s); return true;
    case 11: launch_geo_n<11>(a, mode, grid, s); return true;
    case 12: launch_geo_n<12>(a, mode, grid, s); return true;
    default: return false;
    }
}

}  // namespace nkc
