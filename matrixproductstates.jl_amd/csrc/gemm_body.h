// Complex GEMM tile body shared by the GEMM translation units: kernel
// parameters, operand loaders, epilogue and the MFMA main loop of one
// output block (see gemm.hip for the scheme).
#pragma once
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
  int zcount;     // batch * lockstep copies: output blocks = tiles_m * tiles_n * zcount per split
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
__device__ __forceinline__ void zgemm_body(const GemmParams& p, int tile, int zi) {
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
  const int split = blockIdx.y, b = zi % p.batch;
  const int64_t bo = (int64_t)(zi / p.batch) * p.bst;   // lockstep copy
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

// capped-grid launch of the complex GEMM (gemm_capped.hip): grid.x
// workgroups per split walk every output block; the arithmetic of each
// block is that of the one-block-per-workgroup kernels
void launch_gemm_capped(const GemmParams& p, DType ta, DType tb, DType tc, Compute cmp, int tm, dim3 grid,
                        hipStream_t s);

}  // namespace mpship
