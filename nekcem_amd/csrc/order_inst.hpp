// order_inst.hpp -- the stage and adjoint kernel instances of one polynomial
// order.  Included by order_n<N1>.hip with NKC_N1 = N + 1 defined.
#pragma once
#ifndef NKC_N1
#error "define NKC_N1 before including order_inst.hpp"
#endif
#include "order_launch.hpp"

namespace nkc {
namespace {

// Kernel instance by feature set: FEAT = 2 with a PML, an incident field or
// a graphene sheet; 1 with a source, Drude or Lorentz region; 0 otherwise.
template <int N1, int MAT, int MODE, int PACK, int OCC, int EPW>
void launch_t(const StageArgs &a, int grid, hipStream_t s) {
    const dim3 bs(Shape<N1, PACK>::BS * EPW);
    if (a.pml_off || a.inc_mask || a.gr_off)
        hipLaunchKernelGGL((stage_kernel<N1, MAT, MODE, PACK, OCC, 2, EPW>), dim3(grid), bs, 0, s, a);
    else if (a.src_off || a.drude_off || a.lor_off)
        hipLaunchKernelGGL((stage_kernel<N1, MAT, MODE, PACK, OCC, 1, EPW>), dim3(grid), bs, 0, s, a);
    else
        hipLaunchKernelGGL((stage_kernel<N1, MAT, MODE, PACK, OCC, 0, EPW>), dim3(grid), bs, 0, s, a);
}

// The residual test entry always runs layout 0.
template <int N1, int MAT, int PACK, int OCC, int EPW>
void launch_m(const StageArgs &a, int mode, int grid, hipStream_t s) {
    if (mode == MODE_RK) launch_t<N1, MAT, MODE_RK, PACK, OCC, EPW>(a, grid, s);
    else launch_t<N1, MAT, MODE_RESIDUAL, 0, 0, 1>(a, grid, s);
}

template <int N1, int PACK, int OCC, int EPW>
void launch_p(const StageArgs &a, int mat, int mode, int grid, hipStream_t s) {
    if (mat == MAT_ELEM) launch_m<N1, MAT_ELEM, PACK, OCC, EPW>(a, mode, grid, s);
    else launch_m<N1, MAT_NODE, PACK, OCC, EPW>(a, mode, grid, s);
}

// The six-role kernel (layouts 6, 7): RK stages with at most a source,
// Drude or Lorentz region; anything else runs layout 0.
template <int N1, int OCC>
void launch_s6(const StageArgs &a, int mat, int mode, int grid, hipStream_t s) {
    if (mode != MODE_RK || a.pml_off || a.inc_mask || a.gr_off) {
        launch_p<N1, 0, 0, 1>(a, mat, mode, grid, s);
        return;
    }
    const dim3 bs(Shape6<N1>::BS);
    const bool feat = a.src_off || a.drude_off || a.lor_off;
    if (mat == MAT_ELEM) {
        if (feat) hipLaunchKernelGGL((stage6_kernel<N1, MAT_ELEM, MODE_RK, OCC, 1>), dim3(grid), bs, 0, s, a);
        else hipLaunchKernelGGL((stage6_kernel<N1, MAT_ELEM, MODE_RK, OCC, 0>), dim3(grid), bs, 0, s, a);
    } else {
        if (feat) hipLaunchKernelGGL((stage6_kernel<N1, MAT_NODE, MODE_RK, OCC, 1>), dim3(grid), bs, 0, s, a);
        else hipLaunchKernelGGL((stage6_kernel<N1, MAT_NODE, MODE_RK, OCC, 0>), dim3(grid), bs, 0, s, a);
    }
}

template <int N1>
void launch_n(const StageArgs &a, int mat, int mode, int layout, int grid, hipStream_t s) {
    switch (layout) {
    case 1: launch_p<N1, 1, 0, 1>(a, mat, mode, grid, s); break;
    case 2: launch_p<N1, 0, 1, 1>(a, mat, mode, grid, s); break;
    case 3: launch_p<N1, 1, 1, 1>(a, mat, mode, grid, s); break;
    case 4:
        if constexpr (N1 <= EPW_MAX_N1) launch_p<N1, 1, 0, 2>(a, mat, mode, grid, s);
        break;
    case 5:
        if constexpr (N1 <= EPW_MAX_N1) launch_p<N1, 1, 1, 2>(a, mat, mode, grid, s);
        break;
    case 6:
        if constexpr (N1 >= S6_MIN_N1) launch_s6<N1, 0>(a, mat, mode, grid, s);
        break;
    case 7:
        if constexpr (N1 >= S6_MIN_N1) launch_s6<N1, 1>(a, mat, mode, grid, s);
        break;
    default: launch_p<N1, 0, 0, 1>(a, mat, mode, grid, s); break;
    }
}

template <int N1, int MAT>
void launch_adj_t(const AdjArgs &a, int grid, hipStream_t s) {
    if (a.rs)
        hipLaunchKernelGGL((adjoint_kernel<N1, MAT, 1>), dim3(grid), dim3(Shape<N1>::BS), 0, s, a);
    else
        hipLaunchKernelGGL((adjoint_kernel<N1, MAT, 0>), dim3(grid), dim3(Shape<N1>::BS), 0, s, a);
}

}  // namespace

#define NKC_CAT2(a, b) a##b
#define NKC_CAT(a, b) NKC_CAT2(a, b)

void NKC_CAT(launch_stage_n, NKC_N1)(const StageArgs &a, int mat, int mode, int layout, int grid, hipStream_t s) {
    launch_n<NKC_N1>(a, mat, mode, layout, grid, s);
}

void NKC_CAT(launch_adjoint_n, NKC_N1)(const AdjArgs &a, int mat, int grid, hipStream_t s) {
    if (mat == MAT_ELEM) launch_adj_t<NKC_N1, MAT_ELEM>(a, grid, s);
    else launch_adj_t<NKC_N1, MAT_NODE>(a, grid, s);
}

#undef NKC_CAT
#undef NKC_CAT2

}  // namespace nkc
