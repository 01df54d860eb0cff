// Complex GEMM on CDNA4 MFMA with two-level index maps.
//
// One kernel serves every contraction of the hot path: each operand index
// (m, n, k) is a flattened group of reference tensor indices whose memory
// offset is given by an IdxMap, so the permutedims/reshape chains of the
// reference (mps_base.jl) cost nothing.
//
//   CMP_F64     complex FP64, 3 real MFMA per complex MAC (Gauss/3M:
//               P1 = Ar Br, P2 = Ai Bi, P3 = (Ar+Ai)(Br+Bi);
//               Re = P1 - P2, Im = P3 - P1 - P2), v_mfma_f64_16x16x4_f64,
//               FP64 accumulation
//   CMP_F32_3M  complex FP32, the same 3M products on v_mfma_f32_16x16x4_f32,
//               FP32 accumulation.
// 3M rounding is normwise: |C - fl(C)| <= c K eps |A| |B| elementwise in the
// moduli of whole complex rows/columns (not componentwise in Re/Im), with c
// about 3 against 2 for the four-product form; an imaginary part much
// smaller than |A||B| carries an absolute, not relative, error of that size.
//
// Workgroup tile TM x TM (complex; TM = 64 or 32), K step 16, 256 threads =
// 4 waves of 64, each wave a TM/2 x TM/2 block of MFMA 16x16 tiles (2x2 or
// 1x1).  The 32 tile serves small outputs (a few dozen 64 tiles would leave
// most CUs idle, and splitting a short K costs more in partials than it
// gains).  Deep-K products are split over blockIdx.y (split-K) into a
// partial buffer and reduced by a second kernel in a fixed order, so results
// are bitwise reproducible run to run.
#include "gemm_body.h"

#include <algorithm>

namespace mpship {

// complex FP64 (v_mfma_f64_16x16x4_f64, FP64 accumulation): one output
// block (tile blockIdx.x, copy/batch blockIdx.z) per workgroup
template <int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__global__ __launch_bounds__(NT, 2) void zgemm_f64_kernel(GemmParams p) {
  zgemm_body<true, TM, TA, TB, TC, TWO, KF>(p, blockIdx.x, blockIdx.z);
}

// complex FP32 (v_mfma_f32_16x16x4_f32, FP32 accumulation)
template <int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__global__ __launch_bounds__(NT, 2) void zgemm_f32_3m_kernel(GemmParams p) {
  zgemm_body<false, TM, TA, TB, TC, TWO, KF>(p, blockIdx.x, blockIdx.z);
}

// ------------------------------------------------------- split-K reduction
template <typename TC>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(GemmParams p) {
  const int64_t bo = (int64_t)blockIdx.y * p.bst;   // lockstep copy
  const int64_t MN = (int64_t)p.M * p.N;
  const int64_t total = MN * p.batch;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(idx / MN);
    const int64_t e = idx - (int64_t)b * MN;
    const int col = (int)(e / p.M), row = (int)(e - (int64_t)col * p.M);
    // partials summed in split order; eight loads in flight per batch
    double sr = 0.0, si = 0.0;
    const double2* P = boff((const double2*)p.P, bo) + (int64_t)b * MN + e;
    const int64_t ss = (int64_t)p.batch * MN;
    int s = 0;
    for (; s + 8 <= p.splits; s += 8) {
      double2 v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = P[(s + q) * ss];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        sr += v[q].x;
        si += v[q].y;
      }
    }
    for (; s < p.splits; ++s) {
      const double2 v = P[s * ss];
      sr += v.x;
      si += v.y;
    }
    TC* C = boff((TC*)p.C, bo) + (int64_t)b * p.c_bs + moff(p.cm, row) + moff(p.cn, col);
    double re = p.ar * sr - p.ai * si;
    double im = p.ar * si + p.ai * sr;
    if (p.br != 0.0 || p.bi != 0.0) {
      double orr, oi;
      ldc(C, orr, oi);
      re += p.br * orr - p.bi * oi;
      im += p.br * oi + p.bi * orr;
    }
    re += (row == col) ? p.dg : 0.0;
    stc(C, re, im);
  }
}

// ------------------------------------------------------------------ host
static DevMap dm(const IdxMap& m) { return DevMap{m.n0, m.s0, m.s1}; }

static bool single_unit(const IdxMap& m) { return m.n0 == 0 && m.s0 == 1; }

struct GemmPlan {
  int tm, tiles_m, tiles_n, splits, kchunk;
};

// ------------------------------------------------------- tuned plans
// Written by mpsamd/autotune.py from timings on the MI355X: per GEMM shape of
// the sweeps (M, N, K, batch), per power-of-two class of lockstep copies and
// per arithmetic, the tile and split-K depth that ran fastest.  A table
// compiled into the library keeps the plan, and so the summation order, a
// fixed function of the shape.
struct TunedPlan {
  int64_t M, N, K, batch;
  int nb_class, cmp, tm, splits;
};
static const TunedPlan kTuned[] = {
#include "gemm_tuned.inc"
    {0, 0, 0, 0, -1, -1, 0, 0}};

int nb_class(int64_t nb) {
  int c = 0;
  while (nb > 1) {
    nb >>= 1;
    ++c;
  }
  return c;
}

static const TunedPlan* tuned_plan(const Gemm& g, int64_t nbp, Compute cmp) {
  const int cl = nb_class(nbp);
  for (const TunedPlan& t : kTuned)
    if (t.M == g.M && t.N == g.N && t.K == g.K && t.batch == g.batch && t.nb_class == cl && t.cmp == (int)cmp)
      return &t;
  return nullptr;
}

static GemmPlan make_plan(const Gemm& g, int num_cus, Compute cmp) {
  GemmPlan pl;
  const int64_t nbq = g.bat.plan_nb > 0 ? g.bat.plan_nb : g.bat.nb;
  const TunedPlan* tp = g.force_tm > 0 ? nullptr : (g.tunable ? tuned_plan(g, nbq, cmp) : nullptr);
  if (g.force_tm > 0 || tp) {
    pl.tm = g.force_tm > 0 ? g.force_tm : tp->tm;
    pl.tiles_m = (int)((g.M + pl.tm - 1) / pl.tm);
    pl.tiles_n = (int)((g.N + pl.tm - 1) / pl.tm);
    int splits = std::max(1, g.force_tm > 0 ? g.force_splits : tp->splits);
    if (g.max_splits > 0 && splits > g.max_splits) splits = g.max_splits;
    int64_t kchunk = (g.K + splits - 1) / splits;
    kchunk = std::max<int64_t>(((kchunk + BK - 1) / BK) * BK, BK);
    splits = (int)std::max<int64_t>((g.K + kchunk - 1) / kchunk, 1);
    pl.splits = splits;
    pl.kchunk = (int)kchunk;
    return pl;
  }
  // 64 tiles unless they would leave more than half of the CUs idle (every
  // lockstep copy is a block of the grid too)
  const int64_t nbp = g.bat.plan_nb > 0 ? g.bat.plan_nb : g.bat.nb;
  const int64_t t64 = ((g.M + 63) / 64) * ((g.N + 63) / 64) * g.batch * nbp;
  pl.tm = (2 * t64 < num_cus) ? 32 : 64;
  pl.tiles_m = (int)((g.M + pl.tm - 1) / pl.tm);
  pl.tiles_n = (int)((g.N + pl.tm - 1) / pl.tm);
  // Split K while the tiles leave CUs idle, up to about 4 x num_cus
  // workgroups (two fit per CU at a time; measured on the skinny long-K
  // products, 1152x128x1792 and 1792x128x1152, the deeper split still wins
  // over its partials pass); a grid with a tile for every CU is not split.
  // Chunks are at least 32 deep for 64 tiles and 128 deep for 32 tiles (a
  // quarter of the MFMA work per K step).
  const int64_t blocks = (int64_t)pl.tiles_m * pl.tiles_n * g.batch * nbp;
  const int64_t target = 4LL * num_cus;
  const int64_t min_chunk = (pl.tm == 64) ? 2 * BK : 8 * BK;
  int splits = 1;
  if (blocks < num_cus && g.K >= 2 * min_chunk) {
    splits = (int)((target + blocks - 1) / blocks);
    const int64_t max_splits = (g.K + min_chunk - 1) / min_chunk;
    if (splits > max_splits) splits = (int)max_splits;
    if (splits > 32) splits = 32;
    if (splits < 1) splits = 1;
  }
  if (g.max_splits > 0 && splits > g.max_splits) splits = g.max_splits;
  int64_t kchunk = (g.K + splits - 1) / splits;
  kchunk = std::max<int64_t>(((kchunk + BK - 1) / BK) * BK, BK);   // K = 0: one empty chunk
  splits = (int)((g.K + kchunk - 1) / kchunk);
  if (splits < 1) splits = 1;
  pl.splits = splits;
  pl.kchunk = (int)(kchunk > 0 ? kchunk : BK);
  return pl;
}

void gemm_plan_info(const Gemm& g, Compute cmp, int num_cus, int* tm, int* splits) {
  const GemmPlan pl = make_plan(g, num_cus, cmp);
  *tm = pl.tm;
  *splits = pl.splits;
}

// split-K partials of one lockstep copy
size_t gemm_workspace_bytes(const Gemm& g, Compute cmp, int num_cus) {
  if (g.M == 0 || g.N == 0) return 0;
  GemmPlan pl = make_plan(g, num_cus, cmp);
  if (pl.splits <= 1) return 0;
  return (size_t)pl.splits * g.batch * g.M * g.N * sizeof(double2);
}

template <bool F64, int TM, typename TA, typename TB, typename TC>
static void launch_kf(const GemmParams& p, bool two, bool kf, dim3 grid, hipStream_t s) {
  // K a multiple of BK (every step full): the clamp- and mask-free loaders
#define MPS_GEMM_LAUNCH(TWO_, KF_)                                                                    \
  do {                                                                                            \
    if constexpr (F64)                                                                            \
      MPS_LAUNCH((zgemm_f64_kernel<TM, TA, TB, TC, TWO_, KF_>), grid, dim3(NT), 0, s, p);   \
    else                                                                                          \
      MPS_LAUNCH((zgemm_f32_3m_kernel<TM, TA, TB, TC, TWO_, KF_>), grid, dim3(NT), 0, s, p); \
  } while (0)
  if (two) {
    if (kf) MPS_GEMM_LAUNCH(true, true); else MPS_GEMM_LAUNCH(true, false);
  } else {
    if (kf) MPS_GEMM_LAUNCH(false, true); else MPS_GEMM_LAUNCH(false, false);
  }
#undef MPS_GEMM_LAUNCH
}

template <typename TA, typename TB, typename TC>
static void launch_typed(const GemmParams& p, Compute cmp, int tm, dim3 grid, hipStream_t s) {
  const bool two = p.am.n0 || p.ak.n0 || p.bk.n0 || p.bn.n0;
  const bool kf = (p.K % BK) == 0;
  if (cmp == CMP_F64) {
    if (tm == 32)
      launch_kf<true, 32, TA, TB, TC>(p, two, kf, grid, s);
    else
      launch_kf<true, 64, TA, TB, TC>(p, two, kf, grid, s);
  } else {
    if (tm == 32)
      launch_kf<false, 32, TA, TB, TC>(p, two, kf, grid, s);
    else
      launch_kf<false, 64, TA, TB, TC>(p, two, kf, grid, s);
  }
}

template <typename TA, typename TB>
static void launch_tc(const GemmParams& p, DType tc, Compute cmp, int tm, dim3 grid, hipStream_t s) {
  if (tc == F64)
    launch_typed<TA, TB, double2>(p, cmp, tm, grid, s);
  else
    launch_typed<TA, TB, float2>(p, cmp, tm, grid, s);
}

void launch_gemm(const Gemm& g, Compute cmp, void* ws, size_t ws_bytes, int num_cus, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  REQUIRE(g.M < (1LL << 31) && g.N < (1LL << 31) && g.K < (1LL << 31), "GEMM too large");
  GemmPlan pl = make_plan(g, num_cus, cmp);
  GemmParams p;
  p.M = (int)g.M;
  p.N = (int)g.N;
  p.K = (int)g.K;
  p.tiles_m = pl.tiles_m;
  p.tiles_n = pl.tiles_n;
  p.splits = pl.splits;
  p.kchunk = pl.kchunk;
  p.A = g.A; p.am = dm(g.am); p.ak = dm(g.ak); p.a_bs = g.a_bs; p.conjA = g.conjA;
  p.a_mfast = single_unit(g.am) ? 1 : 0;
  p.B = g.B; p.bk = dm(g.bk); p.bn = dm(g.bn); p.b_bs = g.b_bs; p.conjB = g.conjB;
  p.b_nfast = single_unit(g.bn) ? 1 : 0;
  p.C = g.C; p.cm = dm(g.cm); p.cn = dm(g.cn); p.c_bs = g.c_bs;
  p.ar = g.alpha_re; p.ai = g.alpha_im; p.br = g.beta_re; p.bi = g.beta_im;
  p.dg = g.diag_re;
  p.batch = (int)g.batch;
  p.bst = g.bat.bs;
  p.zcount = (int)(g.batch * g.bat.nb);
  p.P = nullptr;
  REQUIRE(g.bat.nb >= 1 && g.batch * g.bat.nb <= 65535, "GEMM batch too large");
  if (g.K == 0) {
    // C = beta * C (alpha term vanishes): run the reducer with zero splits.
    p.splits = 0;
    dim3 rg((unsigned)std::min<int64_t>((g.M * g.N * g.batch + 255) / 256, 65536), (unsigned)g.bat.nb);
    if (g.tc == F64)
      MPS_LAUNCH(splitk_reduce_kernel<double2>, rg, dim3(256), 0, s, p);
    else
      MPS_LAUNCH(splitk_reduce_kernel<float2>, rg, dim3(256), 0, s, p);
    return;
  }
  if (pl.splits > 1) {
    const size_t need = (size_t)pl.splits * g.batch * g.M * g.N * sizeof(double2);
    REQUIRE(ws != nullptr && ws_bytes >= need, "GEMM workspace too small");
    p.P = ws;
  }
  dim3 grid((unsigned)(pl.tiles_m * pl.tiles_n), (unsigned)pl.splits, (unsigned)(g.batch * g.bat.nb));
  if (g.max_wg > 0 && (int64_t)grid.x * grid.y * grid.z > g.max_wg) {
    // capped: x workgroups per split walk all tiles x zcount output blocks
    const int64_t blocks = (int64_t)grid.x * grid.z;
    grid.x = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks, g.max_wg / pl.splits));
    grid.z = 1;
    launch_gemm_capped(p, g.ta, g.tb, g.tc, cmp, pl.tm, grid, s);
  } else if (g.ta == F64 && g.tb == F64)
    launch_tc<double2, double2>(p, g.tc, cmp, pl.tm, grid, s);
  else if (g.ta == F64 && g.tb == F32)
    launch_tc<double2, float2>(p, g.tc, cmp, pl.tm, grid, s);
  else if (g.ta == F32 && g.tb == F64)
    launch_tc<float2, double2>(p, g.tc, cmp, pl.tm, grid, s);
  else
    launch_tc<float2, float2>(p, g.tc, cmp, pl.tm, grid, s);
  HIPCHK(hipGetLastError());
  if (pl.splits > 1) {
    dim3 rg((unsigned)std::min<int64_t>((g.M * g.N * g.batch + 255) / 256, 65536), (unsigned)g.bat.nb);
    if (g.tc == F64)
      MPS_LAUNCH(splitk_reduce_kernel<double2>, rg, dim3(256), 0, s, p);
    else
      MPS_LAUNCH(splitk_reduce_kernel<float2>, rg, dim3(256), 0, s, p);
    HIPCHK(hipGetLastError());
  }
}

}  // namespace mpship
