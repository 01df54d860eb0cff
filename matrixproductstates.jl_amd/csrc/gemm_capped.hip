// Complex GEMM on a capped grid: fewer workgroups than output blocks.
//
// A launch whose grid would exceed the caller's workgroup cap (Gemm::max_wg)
// runs here instead: grid.x workgroups per split walk the output blocks
// w = tile + tiles * z (z = batch index x lockstep copy) with stride
// grid.x.  Each block runs the same tile body as the one-block-per-workgroup
// kernels of gemm.hip, so the results are bitwise those of the uncapped
// launch.  The RK4 sweeps cap the GEMMs of their environment-building stream
// so that those run beside the main stream's QR without taking every CU.
// Kept in its own translation unit: the loop changes the register allocation
// of the tile body, and the uncapped kernels keep theirs.
#include "gemm_body.h"

namespace mpship {

template <bool F64, int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__global__ __launch_bounds__(NT, 2) void zgemm_capped_kernel(GemmParams p) {
  const int ntile = p.tiles_m * p.tiles_n;
  const int64_t total = (int64_t)ntile * p.zcount;
  for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
    zgemm_body<F64, TM, TA, TB, TC, TWO, KF>(p, (int)(w % ntile), (int)(w / ntile));
    if (w + gridDim.x < total) __syncthreads();   // the next block's tiles reuse the epilogue's LDS
  }
}

template <bool F64, int TM, typename TA, typename TB, typename TC>
static void capped_kf(const GemmParams& p, dim3 grid, hipStream_t s) {
  const bool two = p.am.n0 || p.ak.n0 || p.bk.n0 || p.bn.n0;
  const bool kf = (p.K % BK) == 0;
  if (two) {
    if (kf)
      MPS_LAUNCH((zgemm_capped_kernel<F64, TM, TA, TB, TC, true, true>), grid, dim3(NT), 0, s, p);
    else
      MPS_LAUNCH((zgemm_capped_kernel<F64, TM, TA, TB, TC, true, false>), grid, dim3(NT), 0, s, p);
  } else {
    if (kf)
      MPS_LAUNCH((zgemm_capped_kernel<F64, TM, TA, TB, TC, false, true>), grid, dim3(NT), 0, s, p);
    else
      MPS_LAUNCH((zgemm_capped_kernel<F64, TM, TA, TB, TC, false, false>), grid, dim3(NT), 0, s, p);
  }
}

template <typename TA, typename TB, typename TC>
static void capped_cmp(const GemmParams& p, Compute cmp, int tm, dim3 grid, hipStream_t s) {
  if (cmp == CMP_F64) {
    if (tm == 32)
      capped_kf<true, 32, TA, TB, TC>(p, grid, s);
    else
      capped_kf<true, 64, TA, TB, TC>(p, grid, s);
  } else {
    if (tm == 32)
      capped_kf<false, 32, TA, TB, TC>(p, grid, s);
    else
      capped_kf<false, 64, TA, TB, TC>(p, grid, s);
  }
}

template <typename TA, typename TB>
static void capped_tc(const GemmParams& p, DType tc, Compute cmp, int tm, dim3 grid, hipStream_t s) {
  if (tc == F64)
    capped_cmp<TA, TB, double2>(p, cmp, tm, grid, s);
  else
    capped_cmp<TA, TB, float2>(p, cmp, tm, grid, s);
}

void launch_gemm_capped(const GemmParams& p, DType ta, DType tb, DType tc, Compute cmp, int tm, dim3 grid,
                        hipStream_t s) {
  REQUIRE(grid.x >= 1 && grid.z == 1, "capped GEMM: grid.x workgroups per split, z = 1");
  if (ta == F64 && tb == F64)
    capped_tc<double2, double2>(p, tc, cmp, tm, grid, s);
  else if (ta == F64 && tb == F32)
    capped_tc<double2, float2>(p, tc, cmp, tm, grid, s);
  else if (ta == F32 && tb == F64)
    capped_tc<float2, double2>(p, tc, cmp, tm, grid, s);
  else
    capped_tc<float2, float2>(p, tc, cmp, tm, grid, s);
}

}  // namespace mpship
