// order_launch.hpp -- per-order launchers of the stage and adjoint kernels.
//
// Every polynomial order's kernel instances live in a translation unit of
// their own (order_n<N1>.hip, which includes order_inst.hpp), so the
// instances of the seven orders compile in parallel.  The host code calls
// them through these plain functions.
#pragma once
#include <hip/hip_runtime.h>

#include "adjoint_kernel.hpp"
#include "stage6_kernel.hpp"
#include "stage_kernel.hpp"

namespace nkc {

// Orders that have the two-element-per-workgroup layouts 4 and 5.
constexpr int EPW_MAX_N1 = 6;
// Orders that have the six-role layouts 6 and 7.
constexpr int S6_MIN_N1 = 8;

#define NKC_DECL_ORDER(n)                                                                          \
    void launch_stage_n##n(const StageArgs &a, int mat, int mode, int layout, int grid, hipStream_t s); \
    void launch_adjoint_n##n(const AdjArgs &a, int mat, int grid, hipStream_t s);
NKC_DECL_ORDER(4)
NKC_DECL_ORDER(5)
NKC_DECL_ORDER(6)
NKC_DECL_ORDER(7)
NKC_DECL_ORDER(8)
NKC_DECL_ORDER(9)
NKC_DECL_ORDER(10)
#undef NKC_DECL_ORDER

}  // namespace nkc
