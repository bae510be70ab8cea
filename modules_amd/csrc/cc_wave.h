// cc_wave.h -- wavefront-level reductions for gfx950 (wave64) built on DPP row operations.
//
// A 64-lane sum runs as six DPP steps (quad swaps, half-row and row mirrors, then the two GFX9 row
// broadcasts) with no LDS traffic; lane 63 ends with the total and readlane makes it wavefront-uniform.
// The summation order is fixed, so the result has the same bits in every run.
#pragma once
#include <hip/hip_runtime.h>

namespace ccgpu {

constexpr int kDppQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kDppRowHalfMirror = 0x141; // lane i <- 7 - i within each 8 lanes
constexpr int kDppRowMirror = 0x140;     // lane i <- 15 - i within each row of 16
constexpr int kDppRowBcast15 = 0x142;    // lane 15 of each row to the next row
constexpr int kDppRowBcast31 = 0x143;    // lane 31 to rows 2 and 3

// v moved by the DPP pattern CTRL; lanes of rows outside ROW_MASK get `old`
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_move(double v, double old)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// value of lane `lane` (wavefront-uniform index) in every lane
__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// sum over the 64 lanes (all lanes must be active); every lane gets the total
__device__ __forceinline__ double wave_sum_dpp(double v)
{
    v += dpp_move<kDppQuadXor1>(v, 0.0);
    v += dpp_move<kDppQuadXor2>(v, 0.0);
    v += dpp_move<kDppRowHalfMirror>(v, 0.0);
    v += dpp_move<kDppRowMirror>(v, 0.0);
    v += dpp_move<kDppRowBcast15, 0xA>(v, 0.0);
    v += dpp_move<kDppRowBcast31, 0xC>(v, 0.0);
    return readlane_d(v, 63);
}

// maximum over the 64 lanes (all lanes active); every lane gets it
__device__ __forceinline__ double wave_max_dpp(double v)
{
    const double lo = -__builtin_inf();
    v = fmax(v, dpp_move<kDppQuadXor1>(v, lo));
    v = fmax(v, dpp_move<kDppQuadXor2>(v, lo));
    v = fmax(v, dpp_move<kDppRowHalfMirror>(v, lo));
    v = fmax(v, dpp_move<kDppRowMirror>(v, lo));
    v = fmax(v, dpp_move<kDppRowBcast15, 0xA>(v, lo));
    v = fmax(v, dpp_move<kDppRowBcast31, 0xC>(v, lo));
    return readlane_d(v, 63);
}

}  // namespace ccgpu
