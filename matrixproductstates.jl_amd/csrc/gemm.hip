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
#include "common.h"

namespace mpship {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int BK = 16, NT = 256;

struct DevMap {
  int32_t n0;
  int64_t s0, s1;
};

struct GemmParams {
  int M, N, K;
  int tiles_m, tiles_n, splits, kchunk;
  const void* A; DevMap am, ak; int64_t a_bs; int conjA; int a_mfast;
  const void* B; DevMap bk, bn; int64_t b_bs; int conjB; int b_nfast;
  void* C; DevMap cm, cn; int64_t c_bs;
  double ar, ai, br, bi;
  double dg;      // added to the real part of every diagonal element C[i][i]
  void* P;        // split-K partials: [(split*batch + b)][n][m]
  int batch;
  int64_t bst;    // byte stride of the lockstep copies (blockIdx.z = copy * batch + b)
};

__device__ __forceinline__ int64_t moff(const DevMap& mp, int i) {
  if (mp.n0 == 0) return (int64_t)i * mp.s0;
  const int q = i / mp.n0;
  return (int64_t)(i - q * mp.n0) * mp.s0 + (int64_t)q * mp.s1;
}

// loader offsets: without two-level maps (TWO = false) every map is a single stride
template <bool TWO>
__device__ __forceinline__ int64_t loff(const DevMap& mp, int i) {
  if constexpr (TWO) return moff(mp, i);
  return (int64_t)i * mp.s0;
}

__device__ __forceinline__ void ldc(const double2* p, double& re, double& im) {
  const double2 v = *p;
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void ldc(const float2* p, double& re, double& im) {
  const float2 v = *p;
  re = v.x;
  im = v.y;
}
__device__ __forceinline__ void stc(double2* p, double re, double im) { *p = make_double2(re, im); }
__device__ __forceinline__ void stc(float2* p, double re, double im) {
  *p = make_float2((float)re, (float)im);
}

// ------------------------------------------------------------ tile loaders
// Each thread owns TM*BK/NT elements of an operand's K-step tile (TM rows of
// M (A) or N (B) by BK of K).  The row part of every element's address, its
// place in the tile and its row bound are fixed for the whole K loop and set
// up once (OpLoader::init); per K step, fetch issues the global loads into
// registers and put masks them (and conjugates) into LDS as Re/Im planes
// S[plane][k][m|n].  When K is a multiple of BK (KF) every step is full: the
// k part of each address is a per-step scalar added to a pointer set up once,
// with no clamp and no k mask; otherwise addresses are clamped and masked.
template <int TM>
struct Frag {
  static constexpr int N = (TM * BK) / NT;
  double re[N], im[N];
};

template <int TM, bool TWO, bool KF, typename T>
struct OpLoader {
  static constexpr int NF = (TM * BK) / NT;
  const T* row[NF];   // base + row part of the address (row clamped) [+ k part of kk when KF, single map]
  int kk[NF];         // k within the step
  int rr[NF];         // row within the tile
  DevMap km;

  __device__ __forceinline__ void init(const T* base, const DevMap& rmap, const DevMap& kmap, int r0, int R,
                                       int fast) {
    const int tid = threadIdx.x;
    km = kmap;
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const int r = fast ? (tid % TM) : (tid / BK + j * (NT / BK));
      kk[j] = fast ? (tid / TM + j * (NT / TM)) : (tid % BK);
      rr[j] = r;
      row[j] = base + loff<TWO>(rmap, min(r0 + r, R - 1));
      if (KF && !TWO) row[j] += (int64_t)kk[j] * km.s0;
    }
  }
  template <bool FULL>
  __device__ __forceinline__ void fetch(int k0, int kend, Frag<TM>& f) const {
    if constexpr (KF && !TWO) {
      const int64_t ko = (int64_t)k0 * km.s0;
#pragma unroll
      for (int j = 0; j < NF; ++j) ldc(row[j] + ko, f.re[j], f.im[j]);
      return;
    }
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const int gk = FULL ? k0 + kk[j] : min(k0 + kk[j], kend - 1);
      ldc(row[j] + loff<TWO>(km, gk), f.re[j], f.im[j]);
    }
  }
  // Rows past the operand's edge hold a clamped row's values: they feed only
  // output rows (of A) or columns (of B) past the edge, which are never
  // stored, so only k past the chunk needs a mask.
  template <bool FULL, typename S>
  __device__ __forceinline__ void put(int k0, int kend, int conj, const Frag<TM>& f, S (*Xs)[BK][TM]) const {
#pragma unroll
    for (int j = 0; j < NF; ++j) {
      const bool ok = FULL || k0 + kk[j] < kend;
      const int c = rr[j] ^ kk[j];   // swizzled row (see zgemm_body)
      Xs[0][kk[j]][c] = (S)(ok ? f.re[j] : 0.0);
      Xs[1][kk[j]][c] = (S)(ok ? (conj ? -f.im[j] : f.im[j]) : 0.0);
    }
  }
};

// --------------------------------------------------------------- epilogue
// MFMA 16x16 output layouts (measured on gfx950 with a lane probe):
//   v_mfma_f32_16x16x4_f32: lane l, element r -> row 4*(l/16) + r, col l%16
//   v_mfma_f64_16x16x4_f64: lane l, element r -> row (l/16) + 4*r, col l%16
// Operands in both: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k.
// One output element (row < M, col < N): a split-K partial, or
// alpha * acc + beta * C (+ dg on the diagonal) through the output maps.
template <typename TC>
__device__ __forceinline__ void store_elem(const GemmParams& p, int b, int64_t bo, int split, int row, int col,
                                           double cr, double ci) {
  if (p.splits > 1) {
    double2* P = boff((double2*)p.P, bo) + ((int64_t)split * p.batch + b) * (int64_t)p.M * p.N;
    P[(int64_t)col * p.M + row] = make_double2(cr, ci);
    return;
  }
  TC* C = boff((TC*)p.C, bo) + (int64_t)b * p.c_bs + moff(p.cm, row) + moff(p.cn, col);
  double re = p.ar * cr - p.ai * ci;
  double im = p.ar * ci + p.ai * cr;
  if (p.br != 0.0 || p.bi != 0.0) {
    double orr, oi;
    ldc(C, orr, oi);
    re += p.br * orr - p.bi * oi;
    im += p.br * oi + p.bi * orr;
  }
  re += (row == col) ? p.dg : 0.0;
  stc(C, re, im);
}

// ------------------------------------------------------------ MFMA kernels
// Complex products with three real MFMAs (3M): P1 = Re a Re b, P2 = Im a Im b,
// P3 = (Re a + Im a)(Re b + Im b); Re = P1 - P2, Im = P3 - P1 - P2.  Rounding
// stays normwise (about 3 eps |a||b| per product against 2 eps for four
// MFMAs), and the MFMA work drops by a quarter: MFMA throughput bounds these
// tiles.  LDS holds the Re and Im planes, double buffered (one barrier per K
// step); the sum planes are formed in registers after the LDS reads.  Step
// it: store the tile of step it+1 (fetched one step earlier) into the other
// buffer, issue the global loads of step it+2, run the MFMAs of step it.
template <bool F64>
struct MfmaT;
template <>
struct MfmaT<true> {
  using S = double;
  using V = v4d;
  static __device__ __forceinline__ V mfma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct MfmaT<false> {
  using S = float;
  using V = v4f;
  static __device__ __forceinline__ V mfma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <bool F64, int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__device__ __forceinline__ void zgemm_body(const GemmParams& p) {
  using S = typename MfmaT<F64>::S;
  using V = typename MfmaT<F64>::V;
  constexpr int WT = TM / 32;   // 16x16 MFMA tiles per wave and dimension
  constexpr int WB = TM / 2;    // a wave's output block is WB x WB
  // LDS: the double-buffered operand tiles, then (after the K loop) each
  // wave's output block, column major.  Rows are not padded but swizzled: element (k, r) of an operand tile
  // sits at column r ^ k (k < BK = 16 <= TM flips only the low four bits of
  // r), and element (row, col) of a wave's output block at row ^ (col & 7).
  // The k-major tile stores of the loaders and the row-major MFMA reads then
  // meet distinct banks, as the padded layout did, and the largest variant
  // (FP64, TM = 64) takes exactly 64 KiB of LDS.
  constexpr size_t TILE_BYTES = 2 * sizeof(S[2][2][BK][TM]);
  constexpr size_t OUT_BYTES = 4 * sizeof(double2) * WB * WB;
  __shared__ __align__(16) char smem[TILE_BYTES > OUT_BYTES ? TILE_BYTES : OUT_BYTES];
  auto As = reinterpret_cast<S(*)[2][BK][TM]>(smem);
  auto Bs = reinterpret_cast<S(*)[2][BK][TM]>(smem + TILE_BYTES / 2);
  const int tile = blockIdx.x, split = blockIdx.y, b = blockIdx.z % p.batch;
  const int64_t bo = (int64_t)(blockIdx.z / p.batch) * p.bst;   // lockstep copy
  const int tm = tile % p.tiles_m, tn = tile / p.tiles_m;
  const int m0 = tm * TM, n0 = tn * TM;
  const int kbeg = split * p.kchunk;
  const int kend = min(p.K, kbeg + p.kchunk);
  const int nit = (kend - kbeg + BK - 1) / BK;
  const TA* A = boff((const TA*)p.A, bo) + (int64_t)b * p.a_bs;
  const TB* B = boff((const TB*)p.B, bo) + (int64_t)b * p.b_bs;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave & 1) * (TM / 2), wn = (wave >> 1) * (TM / 2);
  const int lr = lane & 15, lk = lane >> 4;

  V p1[WT][WT], p2[WT][WT], p3[WT][WT];
#pragma unroll
  for (int i = 0; i < WT; ++i)
#pragma unroll
    for (int j = 0; j < WT; ++j) {
      p1[i][j] = V{0, 0, 0, 0};
      p2[i][j] = V{0, 0, 0, 0};
      p3[i][j] = V{0, 0, 0, 0};
    }

  OpLoader<TM, TWO, KF, TA> la;
  OpLoader<TM, TWO, KF, TB> lb;
  la.init(A, p.am, p.ak, m0, p.M, p.a_mfast);
  lb.init(B, p.bn, p.bk, n0, p.N, p.b_nfast);
  auto stage = [&](int k0, Frag<TM>& fa, Frag<TM>& fb) {
    la.template fetch<KF>(k0, kend, fa);
    lb.template fetch<KF>(k0, kend, fb);
  };
  auto store = [&](int k0, const Frag<TM>& fa, const Frag<TM>& fb, int buf) {
    la.template put<KF>(k0, kend, p.conjA, fa, As[buf]);
    lb.template put<KF>(k0, kend, p.conjB, fb, Bs[buf]);
  };
  Frag<TM> fa, fb;
  stage(kbeg, fa, fb);
  store(kbeg, fa, fb, 0);
  if (nit > 1) stage(kbeg + BK, fa, fb);
  __syncthreads();
  for (int it = 0; it < nit; ++it) {
    const int cur = it & 1;
    if (it + 1 < nit) {
      const int k1 = kbeg + (it + 1) * BK;
      store(k1, fa, fb, cur ^ 1);
      if (it + 2 < nit) stage(k1 + BK, fa, fb);
    }
#pragma unroll
    for (int kq = 0; kq < BK / 4; ++kq) {
      const int k = kq * 4 + lk;
      S a_r[WT], a_i[WT], a_s[WT], b_r[WT], b_i[WT], b_s[WT];
#pragma unroll
      for (int i = 0; i < WT; ++i) {
        const int ra = (wm + i * 16 + lr) ^ k, rb = (wn + i * 16 + lr) ^ k;
        a_r[i] = As[cur][0][k][ra];
        a_i[i] = As[cur][1][k][ra];
        b_r[i] = Bs[cur][0][k][rb];
        b_i[i] = Bs[cur][1][k][rb];
        a_s[i] = a_r[i] + a_i[i];
        b_s[i] = b_r[i] + b_i[i];
      }
#pragma unroll
      for (int i = 0; i < WT; ++i)
#pragma unroll
        for (int j = 0; j < WT; ++j) {
          p1[i][j] = MfmaT<F64>::mfma(a_r[i], b_r[j], p1[i][j]);
          p2[i][j] = MfmaT<F64>::mfma(a_i[i], b_i[j], p2[i][j]);
          p3[i][j] = MfmaT<F64>::mfma(a_s[i], b_s[j], p3[i][j]);
        }
    }
    __syncthreads();
  }
  // Epilogue through LDS (the loop's last barrier has retired the operand
  // tiles): MFMA layouts (measured on gfx950 with a lane probe)
  //   v_mfma_f64_16x16x4_f64: lane l, element r -> row (l/16) + 4r, col l%16
  //   v_mfma_f32_16x16x4_f32: lane l, element r -> row 4(l/16) + r, col l%16
  // go to the wave's block, then each store instruction writes whole column
  // segments of WB rows (64 / WB columns per instruction).
  double2* blk = reinterpret_cast<double2*>(smem) + wave * WB * WB;
#pragma unroll
  for (int i = 0; i < WT; ++i)
#pragma unroll
    for (int j = 0; j < WT; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const S x1 = p1[i][j][r], x2 = p2[i][j][r], x3 = p3[i][j][r];
        const int row = i * 16 + (F64 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
        const int col = j * 16 + (lane & 15);
        blk[col * WB + (row ^ (col & 7))] = make_double2((double)(x1 - x2), (double)(x3 - x1 - x2));
      }
  __syncthreads();
  constexpr int CPI = 64 / WB;   // columns per store instruction
#pragma unroll
  for (int q = 0; q < WB / CPI; ++q) {
    const int rl = lane % WB, cl = q * CPI + lane / WB;
    const int row = m0 + wm + rl, col = n0 + wn + cl;
    const double2 v = blk[cl * WB + (rl ^ (cl & 7))];
    if (row < p.M && col < p.N) store_elem<TC>(p, b, bo, split, row, col, v.x, v.y);
  }
}

// complex FP64 (v_mfma_f64_16x16x4_f64, FP64 accumulation)
template <int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__global__ __launch_bounds__(NT, 2) void zgemm_f64_kernel(GemmParams p) {
  zgemm_body<true, TM, TA, TB, TC, TWO, KF>(p);
}

// complex FP32 (v_mfma_f32_16x16x4_f32, FP32 accumulation)
template <int TM, typename TA, typename TB, typename TC, bool TWO, bool KF>
__global__ __launch_bounds__(NT, 2) void zgemm_f32_3m_kernel(GemmParams p) {
  zgemm_body<false, TM, TA, TB, TC, TWO, KF>(p);
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
  if (g.ta == F64 && g.tb == F64)
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
