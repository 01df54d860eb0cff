// FP64 complex Householder QR with LAPACK conventions, Q formed in place.
//
// Replaces qrq! (orthonormalisation.jl:180-198), which on the reference's
// GPU path copies every panel to the host for LAPACK zgeqrf + zungqr.  The
// algorithm is zgeqr2 + zung2r:
//   zlarfg: beta = -sign(Re alpha) * ||[alpha; x]||, tau = (beta - alpha)/beta,
//           v = [1; x / (alpha - beta)]; tau = 0 (H = I) when x = 0 and
//           Im alpha = 0 -- or, with a rank cutoff tol > 0, when both are
//           <= tol * ||A||_F;
//   geqr2:  A(i:m, i+1:n) <- H(i)^H A(i:m, i+1:n)  (conj(tau));
//   ung2r:  Q = H(1) ... H(n) I(:, 1:n), built from the last reflector back.
// Reflectors with tau = 0 are skipped entirely (no trailing work), which is
// what makes zero-padded, low-rank panels cheap.
//
// The panel is addressed through (rs, cs) strides, so rightorth!'s transpose
// (orthonormalisation.jl:230-232) is a view, and Q lands directly in the
// site tensor's own layout.
//
// Work split: panel-blocked (below); a single-workgroup unblocked kernel
// covers panels taller than the blocked variants hold in registers.
#include "common.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <type_traits>

namespace mpship {

constexpr int QR_NT = 1024;
constexpr int QR_NW = QR_NT / 64;

__device__ __forceinline__ double wave_sum(double v) { return wave_sum_dpp(v); }

// Deterministic workgroup sum: every thread reads the wave partials in the same order.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < QR_NW; ++w) s += sh[w];
  return s;
}

__global__ __launch_bounds__(QR_NT) void qr_householder_kernel(double2* __restrict__ A, int m, int n,
                                                               int64_t rs, int64_t cs, double tol,
                                                               double2* __restrict__ tau,
                                                               double2* __restrict__ Rout, int64_t bs) {
  A = boff(A, blockIdx.y * bs);
  tau = boff(tau, blockIdx.y * bs);
  if (Rout != nullptr) Rout = boff(Rout, blockIdx.y * bs);
  __shared__ double sh[QR_NW];
  __shared__ double2 s_tau, s_scale;
  __shared__ double s_beta;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#define AT(r, c) A[(int64_t)(r) * rs + (int64_t)(c) * cs]

  double thr = 0.0;
  if (tol > 0.0) {
    double s = 0.0;
    const int64_t tot = (int64_t)m * n;
    for (int64_t e = tid; e < tot; e += QR_NT) {
      const int c = (int)(e / m), r = (int)(e - (int64_t)c * m);
      const double2 v = AT(r, c);
      s += v.x * v.x + v.y * v.y;
    }
    thr = tol * sqrt(block_sum(s, sh));
  }

  // ---------------------------------------------------------------- geqr2
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int r = i + 1 + tid; r < m; r += QR_NT) {
      const double2 v = AT(r, i);
      s += v.x * v.x + v.y * v.y;
    }
    const double xn2 = block_sum(s, sh);
    if (tid == 0) {
      const double2 alpha = AT(i, i);
      const double xnorm = sqrt(xn2);
      const bool ident = (thr > 0.0) ? (xnorm <= thr && fabs(alpha.y) <= thr)
                                     : (xn2 == 0.0 && alpha.y == 0.0);
      if (ident) {
        s_tau = make_double2(0.0, 0.0);
        s_beta = 0.0;
      } else {
        const double beta = -copysign(sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2), alpha.x);
        s_tau = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
        // scale = 1 / (alpha - beta)
        const double dr = alpha.x - beta, di = alpha.y;
        const double den = dr * dr + di * di;
        s_scale = make_double2(dr / den, -di / den);
        s_beta = beta;
      }
      tau[i] = s_tau;
    }
    __syncthreads();
    const double2 t = s_tau;
    if (t.x != 0.0 || t.y != 0.0) {
      const double2 sc = s_scale;
      for (int r = i + 1 + tid; r < m; r += QR_NT) {
        const double2 v = AT(r, i);
        AT(r, i) = make_double2(v.x * sc.x - v.y * sc.y, v.x * sc.y + v.y * sc.x);
      }
      __syncthreads();
      // trailing columns: a_j <- a_j - conj(tau) v (v^H a_j)
      for (int j = i + 1 + wave; j < n; j += QR_NW) {
        double wr = 0.0, wi = 0.0;
        for (int r = i + lane; r < m; r += 64) {
          const double2 v = (r == i) ? make_double2(1.0, 0.0) : AT(r, i);
          const double2 a = AT(r, j);
          wr += v.x * a.x + v.y * a.y;   // conj(v) * a
          wi += v.x * a.y - v.y * a.x;
        }
        wr = wave_sum(wr);
        wi = wave_sum(wi);
        // f = conj(tau) * w
        const double fr = t.x * wr + t.y * wi, fi = t.x * wi - t.y * wr;
        for (int r = i + lane; r < m; r += 64) {
          const double2 v = (r == i) ? make_double2(1.0, 0.0) : AT(r, i);
          double2 a = AT(r, j);
          a.x -= v.x * fr - v.y * fi;
          a.y -= v.x * fi + v.y * fr;
          AT(r, j) = a;
        }
      }
      __syncthreads();
      if (tid == 0) AT(i, i) = make_double2(s_beta, 0.0);
    }
    __syncthreads();
  }

  // R (upper triangle) for leftorth / rightorth (Julia qr)
  if (Rout != nullptr) {
    for (int e = tid; e < n * n; e += QR_NT) {
      const int c = e / n, r = e - c * n;
      Rout[e] = (r <= c) ? AT(r, c) : make_double2(0.0, 0.0);
    }
    __syncthreads();
  }

  // ---------------------------------------------------------------- ung2r
  for (int i = n - 1; i >= 0; --i) {
    const double2 t = tau[i];
    const bool nz = (t.x != 0.0 || t.y != 0.0);
    if (i < n - 1 && nz) {
      for (int j = i + 1 + wave; j < n; j += QR_NW) {
        double wr = 0.0, wi = 0.0;
        for (int r = i + lane; r < m; r += 64) {
          const double2 v = (r == i) ? make_double2(1.0, 0.0) : AT(r, i);
          const double2 a = AT(r, j);
          wr += v.x * a.x + v.y * a.y;
          wi += v.x * a.y - v.y * a.x;
        }
        wr = wave_sum(wr);
        wi = wave_sum(wi);
        // f = tau * w
        const double fr = t.x * wr - t.y * wi, fi = t.x * wi + t.y * wr;
        for (int r = i + lane; r < m; r += 64) {
          const double2 v = (r == i) ? make_double2(1.0, 0.0) : AT(r, i);
          double2 a = AT(r, j);
          a.x -= v.x * fr - v.y * fi;
          a.y -= v.x * fi + v.y * fr;
          AT(r, j) = a;
        }
      }
      __syncthreads();
    }
    for (int r = i + 1 + tid; r < m; r += QR_NT) {
      const double2 v = AT(r, i);
      // -tau * v
      AT(r, i) = make_double2(-(t.x * v.x - t.y * v.y), -(t.x * v.y + t.y * v.x));
    }
    for (int r = tid; r < i; r += QR_NT) AT(r, i) = make_double2(0.0, 0.0);
    if (tid == 0) AT(i, i) = make_double2(1.0 - t.x, -t.y);
    __syncthreads();
  }
#undef AT
}

// ----------------------------------------------------------------------------
// Panel-blocked form (the default for m <= 2560).
//
// Same reflectors as zgeqr2 / zung2r, applied through the compact-WY form
// H(c0)...H(c0+nb-1) = I - V T V^H of each panel:
//   qr_colnorm  one wave per column: ||a_c||^2 (rank cutoff only)
//   qr_panel    one workgroup factors an NB-column panel held in registers,
//               rows distributed over threads
//   qr_apply    trailing columns <- (I - V T V^H)^H columns, one per workgroup
//   Q formation, n <= 256: G = V^H V (GEMM), W from T^{-1} W = V[0:n,:]^H
//               (qr_wsolve), Q = I(:, 0:n) - V W (GEMM), over the panel
//   qr_formq    n > 256: Q = (I - V_0 T_0 V_0^H) ... (I - V_p T_p V_p^H) I(:, 0:n),
//               4 columns per workgroup (reflectors read from a separate
//               buffer, so Q can overwrite the panel in place)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a - v f for complex v, f = (fr, fi), with the rounding pinned by explicit fmas
// (the panel's update and its look-ahead of the next diagonal must agree bit
// for bit)
__device__ __forceinline__ double2 cupd(double2 a, double2 v, double fr, double fi) {
  return make_double2(fma(v.y, fi, fma(-v.x, fr, a.x)), fma(-v.y, fr, fma(-v.x, fi, a.y)));
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {   // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// ||A(:, c)||^2 per column: four loads per thread in flight at a time, then
// a fixed-order sum (per thread, wave butterfly, then the four waves)
constexpr int CN_NT = 256;
// (norms = false: none).  For the chain kernel (prog != nullptr) workgroup c
// also zeroes column c of the reflector store Vs (m rows; identity
// reflectors are not stored there), and workgroup 0 zeroes the progress
// words prog[0 .. nprog).
__global__ __launch_bounds__(CN_NT) void qr_colnorm_kernel(const double2* __restrict__ A, int m, int64_t rs,
                                                           int64_t cs, double* __restrict__ colnrm,
                                                           unsigned* __restrict__ prog, int nprog,
                                                           double2* __restrict__ Vs, double2* __restrict__ tau,
                                                           bool norms, int64_t bs) {
  A = boff(A, blockIdx.y * bs);
  colnrm = boff(colnrm, blockIdx.y * bs);
  const int c = blockIdx.x, tid = threadIdx.x;
  if (prog != nullptr) {
    // the chain factorization's inputs: progress words zero, V zero (its
    // identity columns are never stored), tau NaN (every factored column
    // stores its own, so a panel that no workgroup took shows as NaN in Q)
    prog = boff(prog, blockIdx.y * bs);
    Vs = boff(Vs, blockIdx.y * bs);
    tau = boff(tau, blockIdx.y * bs);
    if (c == 0)
      for (int i = tid; i < nprog; i += CN_NT) prog[i] = 0u;
    if (tid == 0) tau[c] = make_double2(__builtin_nan(""), __builtin_nan(""));
    for (int r = tid; r < m; r += CN_NT) Vs[(int64_t)c * m + r] = make_double2(0.0, 0.0);
  }
  if (!norms) return;
  __shared__ double sh[CN_NT / 64];
  double s = 0.0;
  for (int r0 = 0; r0 < m; r0 += 4 * CN_NT) {
    double2 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = r0 + tid + CN_NT * i;
      v[i] = r < m ? A[(int64_t)r * rs + (int64_t)c * cs] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) s += v[i].x * v[i].x + v[i].y * v[i].y;
  }
  s = wave_sum(s);
  if ((tid & 63) == 0) sh[tid >> 6] = s;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < CN_NT / 64; ++w) t += sh[w];
    colnrm[c] = t;
  }
}

// Cross-half sums with the gfx950 lane swaps (VALU, no LDS round trip):
//   add32(x, y): lanes 0-31 get x[l] + x[l+32], lanes 32-63 get y[l-32] + y[l]
//   add16(x, y): within each 32-lane half, the 16-lane row with (l & 16) == 0
//                gets x[l] + x[l^16], the other row y[l^16] + y[l]
__device__ __forceinline__ double swap_add32(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(x), __double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(x), __double2hiint(y), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
__device__ __forceinline__ double swap_add16(double x, double y) {
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
  return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
// value of lane l ^ 4: row_shl:4 into banks 0 and 2, row_shr:4 into banks 1 and 3
__device__ __forceinline__ double xor4_d(double v) {
  int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x104, 0xf, 0x5, false);
  int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x104, 0xf, 0x5, false);
  lo = __builtin_amdgcn_update_dpp(lo, __double2loint(v), 0x114, 0xf, 0xa, false);
  hi = __builtin_amdgcn_update_dpp(hi, __double2hiint(v), 0x114, 0xf, 0xa, false);
  return __hiloint2double(hi, lo);
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// (j, l) of the p-th pair j < l, ordered by l then j
__host__ __device__ constexpr int pair_l(int p) {
  int l = 1;
  while (p >= l) { p -= l; ++l; }
  return l;
}
__host__ __device__ constexpr int pair_j(int p) {
  int l = 1;
  while (p >= l) { p -= l; ++l; }
  return p;
}

// Butterfly reduce-scatter of 8 doubles over a wave: afterwards lane l holds
// the wave total of v[l >> 3].  The exchanges across 32 and 16 lanes halve
// the values a lane carries (lane swaps), the one across 8 lanes picks one
// of the remaining two (DPP row_ror:8), and xor 4, 2, 1 add within the group
// of 8; those final adds commute, so the 8 lanes hold identical bits.
__device__ __forceinline__ double wave_rs8(const double (&v)[8]) {
  const int lane = threadIdx.x & 63;
  double w4[4], w2[2];
#pragma unroll
  for (int q = 0; q < 4; ++q) w4[q] = swap_add32(v[q], v[q + 4]);
#pragma unroll
  for (int q = 0; q < 2; ++q) w2[q] = swap_add16(w4[q], w4[q + 2]);
  const bool h3 = (lane & 8) != 0;
  const double keep = h3 ? w2[1] : w2[0];
  const double send = h3 ? w2[0] : w2[1];
  double w = keep + dpp_d<0x128>(send);   // row_ror:8 pairs lane i with i ^ 8
  w += xor4_d(w);
  w += dpp_d<0x4E>(w);                     // quad_perm(2,3,0,1): i ^ 2
  w += dpp_d<0xB1>(w);                     // quad_perm(1,0,3,2): i ^ 1
  return w;
}

// 1/x from the hardware estimate and two Newton steps (x finite, nonzero)
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// 1/sqrt(x) from the hardware estimate and two Newton steps (x > 0 normal):
// r <- r (3 - x r^2) / 2, the second step in the residual form h = 1/2 - (x r) r / 2
__device__ __forceinline__ double rsq_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  r = r * fma(-hx * r, r, 1.5);
  const double h = fma(-hx * r, r, 0.5);
  return fma(r, h, r);
}

// One workgroup factors panel columns c0 .. c0+nbp-1 (nbp <= NB), rows
// c0..m-1.  Thread t holds rows t, t+NT, ..., t+(R-1)NT of the panel in
// registers (rows past the panel are zero and stay zero through every step,
// so only the first row slot, the one that can meet the diagonal, is
// masked), so every dot product of a step is a local partial sum; only the
// zlarfg scalars and the dot products v_k^H a_j cross lanes (butterfly in
// the wave, LDS across waves).  The column loop is not unrolled (the
// unrolled body overflows the instruction cache): the panel shifts left by
// one column per step, so the active column always sits in a[.][0].  Each
// finished column goes to Vs as the explicit reflector and, on rows <= k,
// to A as a column of R (the reflector tail below the diagonal is not
// stored in A: Q is formed from Vs and overwrites the panel).
// PT: also form the compact-WY factor T (H(0)...H(nbp-1) = I - V T V^H,
// zlarft) from the Gram entries G[j][l] = v_j^H v_l; otherwise the apply
// kernel works from the Gram entries itself and T is formed after the
// factorization from the Gram GEMM.
// Exchange between the chain kernel's workgroups, which all run on one XCD:
// plain stores (the L1 writes through; a store is complete once that XCD's
// L2 holds it) and raw-buffer loads with sc1 set (device scope, straight into
// LDS), which miss the reading CU's L1 and read the L2 the stores went to.
constexpr int AUX_SC1 = 16;                 // cache-policy operand: sc1
constexpr uint32_t OFF_NONE = 0xFFFFFFF0u;  // past every buffer: the hardware drops the access
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ void st_buf(__amdgpu_buffer_rsrc_t r, uint32_t byte_off, double2 x) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  const v4i v = {__double2loint(x.x), __double2hiint(x.x), __double2loint(x.y), __double2hiint(x.y)};
  __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)byte_off, 0, 0);
}
// where the chain kernel's panel_factor writes: R entries to A, the
// reflectors to Vs and tau; byte offsets of this thread's slot-0 row of
// column c0 and the steps per column and per row slot
struct ChainIO {
  __amdgpu_buffer_rsrc_t rA, rV, rT;
  uint32_t abase, acol, vbase, vcol, vslot;
  unsigned* prog;
};
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0F70);
}

// LDS reduction slots of the panel factorization
template <int NB, int NT>
struct PanelShape {
  static constexpr int NW = NT / 64;
  static constexpr int GSTEP = (2 * (NB - 1) + 7) / 8 + 1;   // 8-value groups per column step (+1 look-ahead)
  static_assert(GSTEP <= 8, "panel too wide");
  static constexpr int GMAX = GSTEP > 4 ? GSTEP : 4;         // the up-front reduction uses 4
  static constexpr int RED = NW * GMAX * 64;
};

// Factors the NB-column panel held in a[R][NB] (columns c0 .. c0+nbp-1).
// Thread t's slot i holds row base + t + NT i, where base = c0 for the panel
// kernel (rows from the panel's diagonal down) and 0 for the chain kernel
// (all m rows); tk = t + base - c0 is slot 0's row relative to the panel's
// first diagonal element, so the diagonal of column k is tk == k, rows with
// tk < k are R entries and slots past M rows are zero.  Ak / Vk point at this
// thread's slot-0 row of column c0 in A and in the reflector store Vs.
// CHAIN: the column stores go through io (see ChainIO), and the reflectors
// are published through io->prog, which holds the number c of complete
// reflectors in bits 0-7 and, in bits 8-15, which of them are identities
// (tau = 0).  A column's stores are issued after its step's barrier and
// published, once complete, at the next barrier.
template <int R, int NB, int NT, bool CHAIN>
__device__ __forceinline__ void panel_factor(double2 (&a)[R][NB], const int tk, const int M, const int m,
                                             const int64_t cs, const int c0, const int nbp, const double tol,
                                             double tsum, double2* const Ak, double2* const Vk,
                                             double2* tau, double2* s_tau,
                                             double (&red)[2][PanelShape<NB, NT>::RED], const ChainIO* io) {
  constexpr int NW = PanelShape<NB, NT>::NW;
  constexpr int GMAX = PanelShape<NB, NT>::GMAX;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // sum over waves of the butterfly results of group g; lane l gets value l >> 3
  auto wave_sum_slots = [&](const double* r, int g) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += r[(w * GMAX + g) * 64 + lane];
    return s;
  };
  auto bcast = [&](double v, int srclane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), srclane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), srclane);
    return __hiloint2double(hi, lo);
  };
  int ph = 0;
  unsigned idmask = 0;   // CHAIN: identity reflectors of this panel so far (bit k)
  int stored = 0;        // CHAIN: columns whose stores have been issued
  // CHAIN: column k's stores: R entries (rows <= k) to A, tau from thread 0,
  // and the reflector to Vs unless it is an identity (Vs was zeroed before
  // the launch; an identity's column of V never reaches Q, as its row of W
  // is zero)
  auto chain_store = [&](int k, double2 rval, const double2* vv, double2 tval) {
    if constexpr (CHAIN) {
      st_buf(io->rA, tk <= k ? io->abase + (uint32_t)k * io->acol : OFF_NONE, rval);
      if (vv != nullptr) {
#pragma unroll
        for (int i = 0; i < R; ++i)
          st_buf(io->rV,
                    tid + NT * i < M ? io->vbase + (uint32_t)k * io->vcol + (uint32_t)i * io->vslot : OFF_NONE,
                    vv[i]);
      }
      st_buf(io->rT, tid == 0 ? (uint32_t)(c0 + k) * 16u : OFF_NONE, tval);
      stored = k + 1;
    }
  };
  // look-ahead: the zlarfg inputs of the next column, downdated from sums
  // reduced together with the current column's dot products (uniform)
  bool have_next = false;
  double nx_xs = 0.0, nx_ar = 0.0, nx_ai = 0.0;
  // Identity prefix (8-column panels): the zlarfg inputs of every column,
  // reduced together up front.  While the panel's reflectors are identities
  // no column changes, so column k's inputs are these values, with the same
  // bits as its own reduction would give (the same per-thread sums, and the
  // butterfly pairs lanes alike for every value); a panel of identity
  // reflectors then costs one reduction instead of one per column.
  bool up_valid = false;
  double up_xs = 0.0, up_ar = 0.0, up_ai = 0.0;   // lanes 8k..8k+7 hold column k's value
  if constexpr (NB == 8) {
    double vx[8], vr[8], vi[8];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const double2 x = a[i][k];
        const double n2 = fma(x.x, x.x, x.y * x.y);
        s += (i > 0 || tk > k) ? n2 : 0.0;
      }
      vx[k] = s;
      vr[k] = (tk == k) ? a[0][k].x : 0.0;
      vi[k] = (tk == k) ? a[0][k].y : 0.0;
    }
    red[ph][(wave * GMAX + 0) * 64 + lane] = wave_rs8(vx);
    red[ph][(wave * GMAX + 1) * 64 + lane] = wave_rs8(vr);
    red[ph][(wave * GMAX + 2) * 64 + lane] = wave_rs8(vi);
    if (tol > 0.0) {
      const double vt[8] = {tsum, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      red[ph][(wave * GMAX + 3) * 64 + lane] = wave_rs8(vt);
    }
    __syncthreads();
    up_xs = wave_sum_slots(red[ph], 0);
    up_ar = wave_sum_slots(red[ph], 1);
    up_ai = wave_sum_slots(red[ph], 2);
    if (tol > 0.0) tsum = bcast(wave_sum_slots(red[ph], 3), 0);
    ph ^= 1;
    up_valid = true;
  } else if (tol > 0.0) {
    // red[1] is first written by the second column step, after the first
    // step's barrier, so every thread has read these partials by then
    const double w = wave_sum(tsum);
    if (lane == 0) red[1][wave] = w;
    __syncthreads();
    tsum = 0.0;
#pragma unroll
    for (int q = 0; q < NW; ++q) tsum += red[1][q];
  }
  const double thr = tol > 0.0 ? tol * sqrt(tsum) : 0.0;
  const double thr2 = thr * thr;
  if constexpr (CHAIN && NB == 8) {
    // a panel of identity reflectors only (each column's zlarfg test on the
    // up-front values, the same values its step uses): nothing of it is read
    // by the other workgroups, so it is published at once
    const bool idl = (thr > 0.0) ? (up_xs <= thr2 && fabs(up_ai) <= thr) : (up_xs == 0.0 && up_ai == 0.0);
    const uint64_t ib = __ballot((lane & 7) == 0 && (lane >> 3) < nbp && idl);
    const uint64_t all = [&] {
      uint64_t b = 0;
      for (int k = 0; k < nbp; ++k) b |= 1ull << (8 * k);
      return b;
    }();
    if (ib == all && tid == 0)
      __hip_atomic_store(io->prog, (unsigned)nbp | (((1u << nbp) - 1u) << 8), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  // One column step: column k sits in slot O of a[.][], the columns trailing
  // it in slots O+1 .. W-1 (zero past the panel) and are updated in place.
  // Steps go in pairs (O = 0, then 1) and the panel then shifts left by two
  // slots, so a register shift is paid once per two columns; once at most
  // W-1 columns trail column k, the step shrinks to width W.
  auto step = [&](auto Wc, auto Oc, int k) {
    constexpr int W = decltype(Wc)::value;
    constexpr int O = decltype(Oc)::value;
    constexpr int T = W - 1 - O;                 // trailing slots
    static_assert(T >= 0, "slot O must be within the register panel");
    constexpr int GW = (2 * T + 7) / 8;
    if (k >= NB) __builtin_unreachable();
    // zlarfg on column k (= a[.][O]): ||x||^2 below row k and alpha = a(k, k)
    double xs, ar, ai;
    if (have_next) {
      xs = nx_xs;
      ar = nx_ar;
      ai = nx_ai;
    } else if (up_valid) {
      xs = bcast(up_xs, 8 * k);
      ar = bcast(up_ar, 8 * k);
      ai = bcast(up_ai, 8 * k);
    } else {
      double v8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const double2 x = a[i][O];
        const double n2 = fma(x.x, x.x, x.y * x.y);
        v8[0] += (i > 0 || tk > k) ? n2 : 0.0;
      }
      v8[1] = (tk == k) ? a[0][O].x : 0.0;
      v8[2] = (tk == k) ? a[0][O].y : 0.0;
      red[ph][wave * GMAX * 64 + lane] = wave_rs8(v8);
      __syncthreads();
      const double s0 = wave_sum_slots(red[ph], 0);
      ph ^= 1;
      xs = bcast(s0, 0);
      ar = bcast(s0, 8);
      ai = bcast(s0, 16);
    }
    have_next = false;
    const bool ident = (thr > 0.0) ? (xs <= thr2 && fabs(ai) <= thr) : (xs == 0.0 && ai == 0.0);
    if (ident) {
      idmask |= 1u << k;
      if (tid == 0) s_tau[k] = make_double2(0.0, 0.0);
      // column k out unchanged (R part to A; Vs: 1 at row k, zero elsewhere);
      // the trailing columns stay as they are
      if constexpr (CHAIN) {
        chain_store(k, a[0][O], nullptr, make_double2(0.0, 0.0));
      } else {
        if (tid == 0) tau[c0 + k] = make_double2(0.0, 0.0);
        if (tk <= k) Ak[(int64_t)k * cs] = a[0][O];
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const int rr = tid + NT * i;
          if (rr < M) Vk[(int64_t)k * m + NT * i] = make_double2((i == 0 && tk == k) ? 1.0 : 0.0, 0.0);
        }
      }
    } else {
      up_valid = false;   // this reflector changes the columns after k
      // beta = -sign(Re alpha) ||(alpha, x)||, tau = (beta - alpha) / beta,
      // v = x / (alpha - beta); alpha - beta != 0 here.  One reciprocal
      // square root gives both beta and 1/beta (the norm is far from the
      // subnormal range whenever the column is not an identity reflector).
      const double ns = fma(ar, ar, fma(ai, ai, xs));
      const double rn = rsq_nr(ns);
      const double beta = -copysign(ns * rn, ar);
      const double ib = -copysign(rn, ar);
      const double dr = ar - beta;
      const double iden = rcp_nr(fma(dr, dr, ai * ai));
      const double2 t = make_double2((beta - ar) * ib, -ai * ib);
      const double2 sc = make_double2(dr * iden, -ai * iden);
      if (tid == 0) s_tau[k] = t;
      if constexpr (!CHAIN) {
        if (tid == 0) tau[c0 + k] = t;
      }
      // vm = v_k on rows k..M-1 and 0 elsewhere (masks the dots and the update)
      double2 vm[R];
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const double2 y = cmul(a[i][O], sc);
        if (i == 0)
          vm[0] = (tk == k) ? make_double2(1.0, 0.0) : ((tk > k) ? y : make_double2(0.0, 0.0));
        else
          vm[i] = y;
      }
      // column k out: R part to A, the reflector to Vs (CHAIN: after the
      // barrier below)
      if constexpr (!CHAIN) {
        if (tk <= k) Ak[(int64_t)k * cs] = (tk == k) ? make_double2(beta, 0.0) : a[0][O];
#pragma unroll
        for (int i = 0; i < R; ++i)
          if (tid + NT * i < M) Vk[(int64_t)k * m + NT * i] = vm[i];
      }
      if constexpr (T == 0) {
        // the panel's last column: nothing trails it, so no dots, no update
        // and no barrier (CHAIN: the end of the panel publishes it)
        if constexpr (CHAIN) chain_store(k, (tk == k) ? make_double2(beta, 0.0) : a[0][O], vm, t);
        return;
      }
      // s_j = v_k^H a_j for the T trailing slots, four columns per butterfly
#pragma unroll
      for (int g = 0; g < GW; ++g) {
        double v8[8];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int j = O + 1 + 4 * g + p;
          double sr = 0.0, si = 0.0;
          if (j < W) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
              const double2 x = a[i][j];
              sr = fma(vm[i].x, x.x, fma(vm[i].y, x.y, sr));
              si = fma(vm[i].x, x.y, fma(-vm[i].y, x.x, si));
            }
          }
          v8[2 * p] = sr;
          v8[2 * p + 1] = si;
        }
        red[ph][(wave * GMAX + g) * 64 + lane] = wave_rs8(v8);
      }
      // look-ahead sums for column k+1 (a[.][O+1] before its update), rows > k+1:
      // ||a||^2, v_k^H a, ||v_k||^2, and a(k+1), v_k(k+1) from their owner
      if constexpr (T >= 1) {
        double v8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < R; ++i) {
          const double2 x = a[i][O + 1], v = vm[i];
          const double2 xb = (i > 0 || tk > k + 1) ? x : make_double2(0.0, 0.0);
          const double2 vb = (i > 0 || tk > k + 1) ? v : make_double2(0.0, 0.0);
          v8[0] = fma(xb.x, xb.x, fma(xb.y, xb.y, v8[0]));
          v8[1] = fma(vb.x, xb.x, fma(vb.y, xb.y, v8[1]));
          v8[2] = fma(vb.x, xb.y, fma(-vb.y, xb.x, v8[2]));
          v8[3] = fma(vb.x, vb.x, fma(vb.y, vb.y, v8[3]));
        }
        const bool at = (tk == k + 1);
        v8[4] = at ? a[0][O + 1].x : 0.0;
        v8[5] = at ? a[0][O + 1].y : 0.0;
        v8[6] = at ? vm[0].x : 0.0;
        v8[7] = at ? vm[0].y : 0.0;
        red[ph][(wave * GMAX + GW) * 64 + lane] = wave_rs8(v8);
      }
      // CHAIN: every column stored so far (issued a step ago) is complete in
      // every thread past the barrier: publish them, then store column k
      if constexpr (CHAIN) wait_vm<0>();
      __syncthreads();
      if constexpr (CHAIN) {
        if (tid == 0 && stored > 0)
          __hip_atomic_store(io->prog, (unsigned)stored | (idmask << 8), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WORKGROUP);
        chain_store(k, (tk == k) ? make_double2(beta, 0.0) : a[0][O], vm, t);
      }
      double sg[GW > 0 ? GW : 1];   // (GW = 0 only in the last-column step, which returned above)
#pragma unroll
      for (int g = 0; g < GW; ++g) sg[g] = wave_sum_slots(red[ph], g);
      const double la = wave_sum_slots(red[ph], GW);
      ph ^= 1;
      // a_j <- a_j - conj(tau) v_k s_j, in place
#pragma unroll
      for (int j = O + 1; j < W; ++j) {
        const int f = 2 * (j - O - 1), g = f / 8, q = f % 8;
        const double sr = bcast(sg[g], 8 * q), si = bcast(sg[g], 8 * q + 8);
        const double fr = t.x * sr + t.y * si, fi = t.x * si - t.y * sr;
        if (j == O + 1 && k + 1 < nbp) {
          // column k+1 after the update: alpha' = a(k+1) - c v_k(k+1) (the
          // owner's update, same operations) and ||x'||^2 =
          // ||a||^2 - 2 Re(conj(c) v_k^H a) + |c|^2 ||v_k||^2, c = conj(tau) s;
          // used where the cancellation costs at most ~1e3 ulps
          // (scale >= the sum of the terms' magnitudes)
          const double u1 = bcast(la, 0), u2r = bcast(la, 8), u2i = bcast(la, 16), u3 = bcast(la, 24);
          const double2 ap = make_double2(bcast(la, 32), bcast(la, 40));
          const double2 v1 = make_double2(bcast(la, 48), bcast(la, 56));
          const double2 al = cupd(ap, v1, fr, fi);
          const double c2 = fr * fr + fi * fi;
          const double xn = u1 - 2.0 * (fr * u2r + fi * u2i) + c2 * u3;
          const double scale = 2.0 * (u1 + c2 * u3);
          if (xn >= 1e-3 * scale) {
            have_next = true;
            nx_xs = xn;
            nx_ar = al.x;
            nx_ai = al.y;
          }
        }
#pragma unroll
        for (int i = 0; i < R; ++i) a[i][j] = cupd(a[i][j], vm[i], fr, fi);
      }
    }
  };
  auto segment = [&](auto Wc, int kb, int ke) {
    constexpr int W = decltype(Wc)::value;
#pragma unroll 1
    for (int k = kb; k < ke; k += 2) {
      step(Wc, std::integral_constant<int, 0>{}, k);
      if (k + 1 < ke) step(Wc, std::integral_constant<int, 1>{}, k + 1);
      // shift the panel: column k+2 becomes a[.][0]
#pragma unroll
      for (int i = 0; i < R; ++i) {
#pragma unroll
        for (int j = 0; j + 2 < W; ++j) a[i][j] = a[i][j + 2];
        a[i][W - 2] = make_double2(0.0, 0.0);
        a[i][W - 1] = make_double2(0.0, 0.0);
      }
    }
  };
  if constexpr (NB == 16) {
    segment(std::integral_constant<int, 16>{}, 0, min(nbp, 4));
    segment(std::integral_constant<int, 12>{}, 4, min(nbp, 8));
    segment(std::integral_constant<int, 8>{}, 8, min(nbp, 12));
    segment(std::integral_constant<int, 4>{}, 12, nbp);
  } else {
    static_assert(NB == 8, "panel width 8 or 16");
    // columns 6 and 7 in a 2-slot panel: column 6's step works on column 7
    // alone, and column 7's step has nothing trailing (no reduction, no barrier)
    segment(std::integral_constant<int, 8>{}, 0, min(nbp, 4));
    segment(std::integral_constant<int, 4>{}, 4, min(nbp, 6));
    segment(std::integral_constant<int, 2>{}, 6, nbp);
  }
  if constexpr (CHAIN) {
    // the rest, once complete
    wait_vm<0>();
    __syncthreads();
    if (tid == 0)
      __hip_atomic_store(io->prog, (unsigned)nbp | (idmask << 8), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

template <int R, int NB, int NT, bool PT>
__global__ __launch_bounds__(NT) void qr_panel_kernel(double2* __restrict__ A, int m, int n, int64_t rs,
                                                      int64_t cs, int c0, int nbp, double tol,
                                                      const double* __restrict__ colnrm,
                                                      double2* __restrict__ Vs, double2* __restrict__ tau,
                                                      double2* __restrict__ Tp, int64_t bs) {
  static_assert(NT > NB, "the first row slot must cover the panel's diagonal");
  {
    const int64_t bo = blockIdx.y * bs;   // lockstep copy
    A = boff(A, bo);
    colnrm = boff(colnrm, bo);
    Vs = boff(Vs, bo);
    tau = boff(tau, bo);
    Tp = boff(Tp, bo);
  }
  constexpr int NW = NT / 64;
  __shared__ double red[2][PanelShape<NB, NT>::RED];
  __shared__ double2 s_tau[NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int M = m - c0;
  // rank cutoff tol * ||A||_F from the column norms, summed in a fixed order
  // (per thread, then the wave butterfly, then over waves); the loads are
  // issued with the panel's own
  double tsum = 0.0;
  if (tol > 0.0)
    for (int c = tid; c < n; c += NT) tsum += colnrm[c];
  double2 a[R][NB];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int rr = tid + NT * i;
    const int64_t ro = (int64_t)(c0 + min(rr, M - 1)) * rs;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double2 x = A[ro + (int64_t)(c0 + min(j, nbp - 1)) * cs];
      a[i][j] = (rr < M && j < nbp) ? x : make_double2(0.0, 0.0);
    }
  }
  double2* const Ak = A + (int64_t)(c0 + tid) * rs + (int64_t)c0 * cs;   // row c0 + tid (slot 0), column c0
  double2* const Vk = Vs + (int64_t)c0 * m + c0 + tid;                    // column c0, row c0 + tid
  panel_factor<R, NB, NT, false>(a, tid, M, m, cs, c0, nbp, tol, tsum, Ak, Vk, tau, s_tau, red,
                                 (const ChainIO*)nullptr);
  // rows above the panel of its reflectors are zero (the GEMMs of the Q
  // formation run over all m rows)
  for (int r = tid; r < c0; r += NT)
    for (int j = 0; j < nbp; ++j) Vs[(int64_t)(c0 + j) * m + r] = make_double2(0.0, 0.0);
  if constexpr (PT) {
    constexpr int NP = NB * (NB - 1) / 2;                 // Gram pairs
    constexpr int NPB = (NP + 31) / 32;                   // blocks of 8 groups of 4 pairs
    __shared__ double gram[NW * NPB * 64];
    __shared__ double2 s_G[NB][NB];
    __shared__ double2 s_T[NB][NB + 1];
    __syncthreads();
    // Gram entries G[j][l] = v_j^H v_l (j < l) from this thread's own rows of
    // Vs (explicit zeros above each reflector's diagonal, so the sums run over
    // rows >= l by themselves): 4 pairs per butterfly, 8 butterflies per
    // stored block, lane t keeping group (t & 7)'s value (t >> 3)
    double2 b[R][NB];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int rr = tid + NT * i;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const double2 x = Vs[(int64_t)(c0 + min(j, nbp - 1)) * m + c0 + min(rr, M - 1)];
        b[i][j] = (rr < M && j < nbp) ? x : make_double2(0.0, 0.0);
      }
    }
    static_for<0, NPB>([&](auto bc) {
      constexpr int blk = decltype(bc)::value;
      double keep = 0.0;
      static_for<0, 8>([&](auto gc) {
        constexpr int gg = decltype(gc)::value;
        constexpr int g = 8 * blk + gg;
        if constexpr (4 * g < NP) {
          double v8[8];
          static_for<0, 4>([&](auto pc) {
            constexpr int pq = decltype(pc)::value;
            constexpr int p = 4 * g + pq;
            double gr = 0.0, gi = 0.0;
            if constexpr (p < NP) {
              constexpr int l = pair_l(p), j = pair_j(p);
#pragma unroll
              for (int i = 0; i < R; ++i) {
                gr += b[i][j].x * b[i][l].x + b[i][j].y * b[i][l].y;
                gi += b[i][j].x * b[i][l].y - b[i][j].y * b[i][l].x;
              }
            }
            v8[2 * pq] = gr;
            v8[2 * pq + 1] = gi;
          });
          const double t = wave_rs8(v8);
          keep = ((lane & 7) == gg) ? t : keep;
        }
      });
      gram[(wave * NPB + blk) * 64 + lane] = keep;
    });
    __syncthreads();
    for (int e = tid; e < NB * NB; e += NT) {
      const int j = e % NB, l = e / NB;
      double2 gv = make_double2(0.0, 0.0);
      if (j < l && l < nbp) {
        // pair p = l(l-1)/2 + j: group p/4 (block g/8, lane slot g%8), values 2(p%4), 2(p%4)+1
        const int p = l * (l - 1) / 2 + j, g = p / 4, f = 2 * (p % 4);
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          gv.x += gram[(w * NPB + g / 8) * 64 + 8 * f + (g % 8)];
          gv.y += gram[(w * NPB + g / 8) * 64 + 8 * (f + 1) + (g % 8)];
        }
      }
      s_G[j][l] = gv;
    }
    __syncthreads();
    // T: T[k][k] = tau_k, T[0:k, k] = -tau_k T[0:k, 0:k] G[0:k, k]; lane j
    // keeps row j of T in registers, G columns are broadcast reads
    if (tid < 64) {
      double2 trow[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) trow[c] = make_double2(0.0, 0.0);
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        if (k < nbp) {
          const double2 tk = s_tau[k];
          double2 acc = make_double2(0.0, 0.0);
#pragma unroll
          for (int l = 0; l < k; ++l) {
            const double2 p = cmul(trow[l], s_G[l][k]);   // trow[l] = 0 for l < lane
            acc.x += p.x;
            acc.y += p.y;
          }
          const double2 v = cmul(tk, acc);
          trow[k] = (lane < k) ? make_double2(-v.x, -v.y) : ((lane == k) ? tk : make_double2(0.0, 0.0));
        }
      }
      if (lane < NB) {
#pragma unroll
        for (int c = 0; c < NB; ++c) s_T[lane][c] = (lane < nbp && c < nbp) ? trow[c] : make_double2(0.0, 0.0);
      }
    }
    __syncthreads();
    for (int e = tid; e < NB * NB; e += NT) Tp[e] = s_T[e % NB][e / NB];
  }
}

// T factors of all panels from the Gram matrix G = V^H V (n x n, column
// major; its diagonal NB x NB blocks are the panels' Gram entries): one
// wave per panel, lane j keeps row j of T (recurrence as zlarft, forward
// columnwise).  Tall[p] is NB x NB column major.
template <int NB>
__global__ __launch_bounds__(64) void qr_tfactor_kernel(const double2* __restrict__ G, int n,
                                                        const double2* __restrict__ tau,
                                                        double2* __restrict__ Tall, int64_t bs) {
  G = boff(G, blockIdx.y * bs);
  tau = boff(tau, blockIdx.y * bs);
  Tall = boff(Tall, blockIdx.y * bs);
  __shared__ double2 sG[NB][NB + 1];
  const int p = blockIdx.x, c0 = p * NB, nbp = min(NB, n - c0), lane = threadIdx.x;
  for (int e = lane; e < NB * NB; e += 64) {
    const int j = e % NB, l = e / NB;
    sG[j][l] = (j < nbp && l < nbp) ? G[(int64_t)(c0 + l) * n + c0 + j] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  double2 trow[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) trow[c] = make_double2(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    if (k < nbp) {
      const double2 tk = tau[c0 + k];
      double2 acc = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < k; ++l) {
        const double2 q = cmul(trow[l], sG[l][k]);
        acc.x += q.x;
        acc.y += q.y;
      }
      const double2 v = cmul(tk, acc);
      trow[k] = (lane < k) ? make_double2(-v.x, -v.y) : ((lane == k) ? tk : make_double2(0.0, 0.0));
    }
  }
  if (lane < NB) {
#pragma unroll
    for (int c = 0; c < NB; ++c)
      Tall[(int64_t)p * NB * NB + lane + c * NB] = (lane < nbp && c < nbp) ? trow[c] : make_double2(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// Chain QR: the whole factorization (zgeqr2's reflectors) in one launch.
// Each of the P = ceil(n/8) panels has a workgroup that holds its columns
// 8p .. 8p+7 over all m rows in registers (thread t, slot i: row t + NT i).
// It applies the reflectors of panels 0 .. p-1 one at a time, as their owners
// publish them,
//   a_j <- a_j - conj(tau_c) v_c (v_c^H a_j)        (H(c)^H, as zgeqr2),
// then factors its own panel with panel_factor, which publishes each
// reflector (its Vs column and tau) by advancing prog[p].  The owner of the
// next panel applies the current panel's reflectors while they are being
// produced, so no panel waits for a separate trailing-update launch, and the
// factorization is one dispatch instead of a panel and an update launch per
// 8 columns.
// One XCD per factorization: the grid has 8 workgroups per panel, dealt to
// the 8 XCDs round robin (qr_chain_probe checks that once per device), and
// only those on one XCD (read from the XCC_ID hardware register) work, so
// every exchange goes through that XCD's L2.  They take panel tickets from
// prog[P] in arrival order (an atomic in that same L2); the others exit at
// once.  Lockstep copies: block g = [8 P g, 8 P (g + 1)) of a one-dimensional
// grid of ceil(nb / 8) blocks serves copies 8 g .. 8 g + 7, copy 8 g + x on
// XCD x (round robin puts exactly P of the block's workgroups on each XCD,
// whichever XCD the launch starts on).
// prog[0 .. P+1) are zero at launch (qr_colnorm_kernel clears them); bits 0-7
// of prog[k] count the complete reflectors of panel k, bits 8-15 mark which
// of them are identities (those are skipped without loading them), and
// PROG_ABORT says that panel's owner gave up.  A workgroup waits only on
// panels with lower tickets, taken by workgroups already running.  Every wait
// is bounded by CHAIN_WAIT_TICKS of s_memrealtime (100 MHz): past it the
// workgroup marks its own panel aborted and stores NaN as its tau, so a
// failure reaches Q as NaN instead of hanging the queue.
constexpr unsigned PROG_ABORT = 0x80000000u;
constexpr uint64_t CHAIN_WAIT_TICKS = 100000000ull;   // 1 s
constexpr int CH_NB = 8;
// reflectors staged per batch: 5 x 20 KiB of LDS at m <= 1280 (256 threads),
// 2 x 40 KiB at m <= 2560 (512 threads)
template <int R, int NT>
constexpr int chain_slots() { return (R * NT * 16) <= 20480 ? 5 : 2; }
constexpr int CH_WG_PER_PANEL = 8;   // one per XCD
__device__ __forceinline__ unsigned xcc_id() { return __builtin_amdgcn_s_getreg(6164) & 15u; }   // hwreg(HW_REG_XCC_ID, 0, 4)

template <int R, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) void qr_chain_kernel(
    double2* __restrict__ A, int m, int n, int64_t rs, int64_t cs, double tol, const double* __restrict__ colnrm,
    double2* Vs, double2* tau, unsigned* prog, unsigned nb, int64_t bs) {
  // Vs, tau and prog are written by other workgroups while this one reads
  // them: no __restrict__ (with it the compiler takes a polled word as
  // invariant and hoists its load out of the wait loop)
  constexpr int NB = CH_NB;
  static_assert(NT >= 256, "the first row slot must cover every diagonal row (n <= NT)");
  const int P = (n + NB - 1) / NB;
  const unsigned copy = (blockIdx.x / (unsigned)(CH_WG_PER_PANEL * P)) * CH_WG_PER_PANEL + xcc_id();
  if (copy >= nb) return;
  {
    const int64_t bo = (int64_t)copy * bs;   // lockstep copy
    A = boff(A, bo);
    colnrm = boff(colnrm, bo);
    Vs = boff(Vs, bo);
    tau = boff(tau, bo);
    prog = boff(prog, bo);
  }
  constexpr int NW = NT / 64;
  constexpr int GMAX = PanelShape<NB, NT>::GMAX;
  __shared__ double red[2][PanelShape<NB, NT>::RED];
  __shared__ double2 s_tau[NB];
  __shared__ unsigned s_poll[2];
  __shared__ int s_ticket;
  // published reflectors staged in LDS (rows t + NT i, as the columns) with
  // their tau (one copy per wave), CH_SLOTS of them per batch
  constexpr int CH_SLOTS = chain_slots<R, NT>();
  __shared__ double2 ring[CH_SLOTS][R * NT];
  __shared__ double2 ring_tau[CH_SLOTS][NT / 64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_ticket = (int)__hip_atomic_fetch_add(prog + P, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int p = s_ticket;
  if (p >= P) return;   // every panel taken (never under round robin)
  const int c0 = NB * p, nbp = min(NB, n - c0);
  ChainIO io;
  io.rA = buf_rsrc(A, (uint32_t)(((int64_t)(m - 1) * rs + (int64_t)(n - 1) * cs + 1) * sizeof(double2)));
  io.rV = buf_rsrc(Vs, (uint32_t)((size_t)m * n * sizeof(double2)));
  io.rT = buf_rsrc(tau, (uint32_t)(n * sizeof(double2)));
  io.abase = (uint32_t)(((int64_t)tid * rs + (int64_t)c0 * cs) * sizeof(double2));
  io.acol = (uint32_t)(cs * sizeof(double2));
  io.vbase = (uint32_t)(((int64_t)c0 * m + tid) * sizeof(double2));
  io.vcol = (uint32_t)(m * sizeof(double2));
  io.vslot = (uint32_t)(NT * sizeof(double2));
  io.prog = prog + p;
  double tsum = 0.0;
  if (tol > 0.0)
    for (int c = tid; c < n; c += NT) tsum += colnrm[c];
  double2 a[R][NB];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    const int rr = tid + NT * i;
    const int64_t ro = (int64_t)min(rr, m - 1) * rs;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double2 x = A[ro + (int64_t)(c0 + min(j, nbp - 1)) * cs];
      a[i][j] = (rr < m && j < nbp) ? x : make_double2(0.0, 0.0);
    }
  }
  auto bcast = [&](double v, int srclane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), srclane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), srclane);
    return __hiloint2double(hi, lo);
  };
  // reflector c into ring slot sl: each wave's rows straight into LDS (16-byte
  // device-scope buffer loads with the LDS flag; rows past m read as zero),
  // and tau by the wave's lane 0 into its own copy
  auto stage_refl = [&](int c, int sl) {
    typedef __attribute__((address_space(3))) void* lds_ptr;
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int rr = tid + NT * i;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(io.rV, (lds_ptr)&ring[sl][wave * 64 + NT * i], 16,
                                               rr < m ? (int)(((size_t)c * m + rr) * sizeof(double2)) : (int)OFF_NONE,
                                               0, 0, AUX_SC1);
    }
    if (lane == 0)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(io.rT, (lds_ptr)&ring_tau[sl][wave], 16, (int)(c * sizeof(double2)), 0, 0,
                                               AUX_SC1);
  };
  int ph = 0, pp = 0;
  bool fail = false;
  // applies reflector v (tau t, the same in every thread) to the 8 columns
  auto apply = [&](const double2 (&v)[R], double2 t) {
    // s_j = v^H a_j, columns 0-3 and 4-7 in one butterfly each
    double v8[2][8];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const double2 x = a[i][j];
        sr = fma(v[i].x, x.x, fma(v[i].y, x.y, sr));
        si = fma(v[i].x, x.y, fma(-v[i].y, x.x, si));
      }
      v8[j / 4][2 * (j % 4)] = sr;
      v8[j / 4][2 * (j % 4) + 1] = si;
    }
    red[ph][(wave * GMAX + 0) * 64 + lane] = wave_rs8(v8[0]);
    red[ph][(wave * GMAX + 1) * 64 + lane] = wave_rs8(v8[1]);
    __syncthreads();
    double sg[2] = {0.0, 0.0};
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      sg[0] += red[ph][(w * GMAX + 0) * 64 + lane];
      sg[1] += red[ph][(w * GMAX + 1) * 64 + lane];
    }
    ph ^= 1;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int q = j % 4;
      const double sr = bcast(sg[j / 4], 16 * q), si = bcast(sg[j / 4], 16 * q + 8);
      const double fr = t.x * sr + t.y * si, fi = t.x * si - t.y * sr;
#pragma unroll
      for (int i = 0; i < R; ++i) a[i][j] = cupd(a[i][j], v[i], fr, fi);
    }
  };
  for (int k = 0; k < p && !fail; ++k) {
    unsigned done = 0;
    while (done < (unsigned)NB) {
      if (tid == 0) {
        unsigned v = __hip_atomic_load(prog + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((v & 0xFFu) <= done) {
          const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
          do {
            __builtin_amdgcn_s_sleep(1);
            v = __hip_atomic_load(prog + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          } while ((v & 0xFFu) <= done && __builtin_amdgcn_s_memrealtime() - t0 < CHAIN_WAIT_TICKS);
        }
        s_poll[pp] = v;
      }
      __syncthreads();
      const unsigned pv = s_poll[pp];
      pp ^= 1;
      const unsigned avail = pv & 0xFFu;
      if ((pv & PROG_ABORT) != 0u || avail <= done || avail > (unsigned)NB) {
        fail = true;
        break;
      }
      // the non-identity reflectors among done .. avail-1 (at most CH_SLOTS
      // per batch): all staged at once, one wait, then applied in order from
      // LDS.  (Staging the next batch while this one is applied measured no
      // faster: when the owner publishes one reflector at a time, each still
      // costs its discovery, load and apply in turn.)
      unsigned rem = ((1u << avail) - (1u << done)) & ~(pv >> 8);
      int ns = 0;
      unsigned upto = avail;
      while (rem != 0u && ns < CH_SLOTS) {
        const int q = __builtin_ctz(rem);
        rem &= rem - 1u;
        stage_refl(NB * k + q, ns++);
      }
      if (rem != 0u) upto = (unsigned)__builtin_ctz(rem);   // the rest in the next batch
      if (ns > 0) {
        wait_vm<0>();
        __syncthreads();
        for (int sl = 0; sl < ns; ++sl) {
          double2 v[R];
#pragma unroll
          for (int i = 0; i < R; ++i) v[i] = ring[sl][tid + NT * i];
          apply(v, ring_tau[sl][wave]);
        }
      }
      done = upto;
    }
  }
  if (fail) {
    if (tid == 0) __hip_atomic_store(prog + p, PROG_ABORT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (tid < nbp) tau[c0 + tid] = make_double2(__builtin_nan(""), __builtin_nan(""));
    return;
  }
  __syncthreads();   // the last apply's reads of red[] before panel_factor writes it
  panel_factor<R, NB, NT, true>(a, tid - c0, m, m, cs, c0, nbp, tol, tsum, A + (int64_t)tid * rs + (int64_t)c0 * cs,
                                Vs + (int64_t)c0 * m + tid, tau, s_tau, red, &io);
}

// ---------------------------------------------------------------------------
// Compact-WY application of one panel (V: rows r0.., NB columns, leading dim
// m; T: NB x NB column-major) to the 4 register columns of a workgroup (one
// per wave; row r0 + lane + 64 i in x[i]).  V is staged through LDS in
// 64*CR-row chunks (CR rows per lane), double buffered, with the next chunk's
// global loads issued before the current chunk is consumed.
//   CONJT: x <- x - V T^H (V^H x)   (geqrf trailing update, H^H)
//  !CONJT: x <- x - V T   (V^H x)   (Q formation, H)
constexpr int WY_NT = 256;
constexpr int WY_CR = 2;
constexpr int WY_CH = 64 * WY_CR;

template <int NB>
struct WyStage {
  double2 buf[2][WY_CH][NB + 1];
  double2 T[NB][NB];
};

template <int NB>
__device__ __forceinline__ void wy_fetch(const double2* __restrict__ V, int64_t ldv, int rows, int nbp, int ch,
                                         double2 (&r)[WY_CH * NB / WY_NT]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < WY_CH * NB / WY_NT; ++q) {
    const int e = tid + WY_NT * q;
    const int row = e % WY_CH, col = e / WY_CH;
    const int gr = min(ch * WY_CH + row, rows - 1);
    r[q] = V[(int64_t)min(col, nbp - 1) * ldv + gr];
  }
}

template <int NB>
__device__ __forceinline__ void wy_put(int rows, int nbp, int ch, const double2 (&r)[WY_CH * NB / WY_NT],
                                       double2 (*b)[NB + 1]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < WY_CH * NB / WY_NT; ++q) {
    const int e = tid + WY_NT * q;
    const int row = e % WY_CH, col = e / WY_CH;
    const bool ok = (ch * WY_CH + row < rows) && (col < nbp);
    b[row][col] = ok ? r[q] : make_double2(0.0, 0.0);
  }
}

template <int MAXR, int NB, bool CONJT>
__device__ __forceinline__ void wy_apply(double2 (&x)[MAXR], const double2* __restrict__ V, int64_t ldv,
                                         int rows, int nbp, const double2* __restrict__ Tg, WyStage<NB>& S) {
  constexpr int NCH = (MAXR + WY_CR - 1) / WY_CR;
  const int lane = threadIdx.x & 63;
  const int nch = (rows + WY_CH - 1) / WY_CH;
  for (int e = threadIdx.x; e < NB * NB; e += WY_NT) S.T[e % NB][e / NB] = Tg[e];
  double2 r[WY_CH * NB / WY_NT];
  double2 w[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) w[k] = make_double2(0.0, 0.0);
  // pass 1: w = V^H x
  wy_fetch<NB>(V, ldv, rows, nbp, 0, r);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if (ch < nch) {
      wy_put<NB>(rows, nbp, ch, r, S.buf[ch & 1]);
      __syncthreads();
      if (ch + 1 < nch) wy_fetch<NB>(V, ldv, rows, nbp, ch + 1, r);
#pragma unroll
      for (int c = 0; c < WY_CR; ++c) {
        const int i = ch * WY_CR + c;
        if (i < MAXR) {
#pragma unroll
          for (int k = 0; k < NB; ++k) {
            const double2 v = S.buf[ch & 1][lane + 64 * c][k];
            w[k].x += v.x * x[i].x + v.y * x[i].y;
            w[k].y += v.x * x[i].y - v.y * x[i].x;
          }
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    w[k].x = wave_sum(w[k].x);
    w[k].y = wave_sum(w[k].y);
  }
  double2 u[NB];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    double ur = 0.0, ui = 0.0;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const double2 tv = CONJT ? S.T[i][k] : S.T[k][i];
      // CONJT: conj(T[i][k]) w_i ; else T[k][i] w_i
      if (CONJT) {
        ur += tv.x * w[i].x + tv.y * w[i].y;
        ui += tv.x * w[i].y - tv.y * w[i].x;
      } else {
        ur += tv.x * w[i].x - tv.y * w[i].y;
        ui += tv.x * w[i].y + tv.y * w[i].x;
      }
    }
    u[k] = make_double2(ur, ui);
  }
  // pass 2: x <- x - V u
  wy_fetch<NB>(V, ldv, rows, nbp, 0, r);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    if (ch < nch) {
      wy_put<NB>(rows, nbp, ch, r, S.buf[ch & 1]);
      __syncthreads();
      if (ch + 1 < nch) wy_fetch<NB>(V, ldv, rows, nbp, ch + 1, r);
#pragma unroll
      for (int c = 0; c < WY_CR; ++c) {
        const int i = ch * WY_CR + c;
        if (i < MAXR) {
          double dr = 0.0, di = 0.0;
#pragma unroll
          for (int k = 0; k < NB; ++k) {
            const double2 v = S.buf[ch & 1][lane + 64 * c][k];
            dr += v.x * u[k].x - v.y * u[k].y;
            di += v.x * u[k].y + v.y * u[k].x;
          }
          x[i].x -= dr;
          x[i].y -= di;
        }
      }
    }
  }
  __syncthreads();
}

// geqrf trailing update: 4 trailing columns per workgroup, rows c0..m-1
// geqrf trailing update, one workgroup per trailing column j:
//   a_j <- (I - V T V^H)^H a_j = a_j - V u,  u = T^H (V^H a_j)   on rows c0..m-1.
// Thread t holds rows t, t+NT, ... (RA of them) of a_j and of the panel's NB
// reflectors in registers for both passes; V^H a_j is one butterfly
// reduction per wave plus an LDS sum over waves.
// OWN_G: T is not read; the same reduction also yields the Gram entries
// G[l][q] = v_l^H v_q (l < q), and since T^{-1} = diag(1/tau) + striu(G),
// u solves the lower-triangular system
//   u_q = conj(tau_q) (w_q - sum_{l<q} conj(G[l][q]) u_l)
// (lane q of the first wave holds row q; u_l is broadcast as it completes).
template <int RA, int NB, int NT, bool OWN_G>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) void qr_apply_col_kernel(double2* __restrict__ A, int m, int64_t rs, int64_t cs,
                                                          int c0, int nbp, const double2* __restrict__ Vs,
                                                          const double2* __restrict__ Tp,
                                                          const double2* __restrict__ tau, int64_t bs) {
  {
    const int64_t bo = blockIdx.y * bs;   // lockstep copy
    A = boff(A, bo);
    Vs = boff(Vs, bo);
    Tp = boff(Tp, bo);
    tau = boff(tau, bo);
  }
  constexpr int NW = NT / 64;
  constexpr int NPAIR = OWN_G ? NB * (NB - 1) / 2 : 0;
  constexpr int NVAL = 2 * NB + 2 * NPAIR;               // w (NB complex), then the Gram pairs
  constexpr int NG = (NVAL + 7) / 8;
  static_assert(NVAL <= NT, "one thread per reduced value");
  __shared__ double red[NW * NG * 64];
  __shared__ double s_val[NG * 8];
  __shared__ double2 s_u[NB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // Workgroups are dealt to the XCDs round-robin (workgroup x on XCD x % 8)
  // and the next panel's factorisation, a single workgroup, runs on XCD 0:
  // where the grid allows, the next panel's columns go to workgroups 0, 8,
  // 16, ... so they are written through that XCD's L2 (a bijection of
  // [0, trail) otherwise in order).
  const int trail = (int)gridDim.x;
  const int P = min(NB, trail);
  int col = blockIdx.x;
  if (8 * (P - 1) < trail) {
    const int x = blockIdx.x;
    col = (x % 8 == 0 && x / 8 < P) ? x / 8 : P + x - min(P, (x + 7) / 8);
  }
  const int j = c0 + nbp + col;
  const int M = m - c0;
  const double2* V = Vs + (int64_t)c0 * m + c0;   // reflector q (panel-local), row c0 + r: V[r + q m]
  // lane q < nbp of every wave reads tau_q (loaded with the operands)
  const int lane_q = (threadIdx.x & 63) < nbp ? (int)(threadIdx.x & 63) : nbp - 1;
  const double2 tq = tau[c0 + lane_q];
  double2 x[RA], v[RA][NB];
#pragma unroll
  for (int i = 0; i < RA; ++i) {
    const int rr = tid + NT * i;
    const bool ok = rr < M;
    const int rc = min(rr, M - 1);
    const double2 xa = A[(int64_t)(c0 + rc) * rs + (int64_t)j * cs];
    x[i] = ok ? xa : make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      const double2 va = V[rc + (int64_t)min(q, nbp - 1) * m];
      v[i][q] = (ok && q < nbp) ? va : make_double2(0.0, 0.0);
    }
  }
  // a panel of identity reflectors leaves every trailing column as it is
  // (the ballot is the same in every wave: the whole workgroup returns)
  if (__ballot((threadIdx.x & 63) < nbp && (tq.x != 0.0 || tq.y != 0.0)) == 0) return;
  const double2 ct = (OWN_G && tid < nbp) ? tq : make_double2(0.0, 0.0);
  // value e: e < 2 NB: Re/Im of w_{e/2} = v_{e/2}^H x; else Re/Im of pair
  // p = (e - 2 NB) / 2, G[pair_j(p)][pair_l(p)].  Rows outermost: the 8
  // sums of a group advance together, so consecutive fmas are independent
  // (a sum finished before the next starts is a chain of dependent FP64
  // latencies at one wave per SIMD).
  static_for<0, NG>([&](auto gc) {
    constexpr int g = decltype(gc)::value;
    double v8[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < RA; ++i) {
      static_for<0, 4>([&](auto hc) {
        constexpr int h = decltype(hc)::value;
        constexpr int e = 8 * g + 2 * h;
        if constexpr (e < 2 * NB) {
          constexpr int q = e / 2;
          const double2 a = v[i][q], b = x[i];
          v8[2 * h] = fma(a.x, b.x, fma(a.y, b.y, v8[2 * h]));
          v8[2 * h + 1] = fma(a.x, b.y, fma(-a.y, b.x, v8[2 * h + 1]));
        } else if constexpr (e < NVAL) {
          constexpr int p = (e - 2 * NB) / 2;
          constexpr int l = pair_l(p), jj = pair_j(p);
          const double2 a = v[i][jj], b = v[i][l];
          v8[2 * h] = fma(a.x, b.x, fma(a.y, b.y, v8[2 * h]));
          v8[2 * h + 1] = fma(a.x, b.y, fma(-a.y, b.x, v8[2 * h + 1]));
        }
      });
    }
    red[(wave * NG + g) * 64 + lane] = wave_rs8(v8);
  });
  __syncthreads();
  if (tid < NVAL) {
    const int g = tid / 8, f = tid % 8;
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += red[(w * NG + g) * 64 + 8 * f];
    s_val[tid] = s;
  }
  __syncthreads();
  if constexpr (OWN_G) {
    if (tid < 64) {
      // lane q: acc = w_q, then the substitution over l = 0 .. NB-1
      const int q = min(lane, NB - 1);
      double2 acc = make_double2(s_val[2 * q], s_val[2 * q + 1]);
      double2 u = make_double2(0.0, 0.0);
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        if (lane == l) u = make_double2(ct.x * acc.x + ct.y * acc.y, ct.x * acc.y - ct.y * acc.x);
        const int lo_r = __builtin_amdgcn_readlane(__double2loint(u.x), l);
        const int hi_r = __builtin_amdgcn_readlane(__double2hiint(u.x), l);
        const int lo_i = __builtin_amdgcn_readlane(__double2loint(u.y), l);
        const int hi_i = __builtin_amdgcn_readlane(__double2hiint(u.y), l);
        const double2 ul = make_double2(__hiloint2double(hi_r, lo_r), __hiloint2double(hi_i, lo_i));
        if (l + 1 < NB && lane > l && lane < NB) {
          // conj(G[l][q]), pair index p = q (q - 1) / 2 + l
          const int p = q * (q - 1) / 2 + l;
          const double gr = s_val[2 * NB + 2 * p], gi = s_val[2 * NB + 2 * p + 1];
          acc.x -= gr * ul.x + gi * ul.y;
          acc.y -= gr * ul.y - gi * ul.x;
        }
      }
      if (lane < NB) s_u[lane] = u;
    }
  } else {
    // u = T^H w
    if (tid < NB) {
      double ur = 0.0, ui = 0.0;
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const double2 t = Tp[i + tid * NB];
        const double wr = s_val[2 * i], wi = s_val[2 * i + 1];
        ur += t.x * wr + t.y * wi;
        ui += t.x * wi - t.y * wr;
      }
      s_u[tid] = make_double2(ur, ui);
    }
  }
  __syncthreads();
  double2 u[NB];
#pragma unroll
  for (int q = 0; q < NB; ++q) u[q] = s_u[q];
#pragma unroll
  for (int i = 0; i < RA; ++i) {
    const int rr = tid + NT * i;
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
      dr += v[i][q].x * u[q].x - v[i][q].y * u[q].y;
      di += v[i][q].x * u[q].y + v[i][q].y * u[q].x;
    }
    if (rr < M) A[(int64_t)(c0 + rr) * rs + (int64_t)j * cs] = make_double2(x[i].x - dr, x[i].y - di);
  }
}

// Q = H(0) ... H(n-1) I(:, 0:n): 4 columns per workgroup (same panel), panel
// blocks from the columns' own panel down to 0, over all m rows (V is zero
// above each panel).
template <int MAXR, int NB>
__global__ __launch_bounds__(WY_NT) void qr_formq_kernel(double2* __restrict__ A, int m, int64_t rs, int64_t cs,
                                                         const double2* __restrict__ Vs,
                                                         const double2* __restrict__ Tall,
                                                         double2* __restrict__ Rout, int n, int64_t bs) {
  A = boff(A, blockIdx.y * bs);
  Vs = boff(Vs, blockIdx.y * bs);
  Tall = boff(Tall, blockIdx.y * bs);
  if (Rout != nullptr) Rout = boff(Rout, blockIdx.y * bs);
  __shared__ WyStage<NB> S;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = min(4 * (int)blockIdx.x + wave, n - 1);
  const bool live = 4 * (int)blockIdx.x + wave < n;
  if (Rout != nullptr && live) {
    for (int r = lane; r < n; r += 64)
      Rout[(int64_t)c * n + r] = (r <= c) ? A[(int64_t)r * rs + (int64_t)c * cs] : make_double2(0.0, 0.0);
  }
  double2 q[MAXR];
#pragma unroll
  for (int i = 0; i < MAXR; ++i) q[i] = make_double2((lane + 64 * i) == c ? 1.0 : 0.0, 0.0);
  const int pmax = (4 * (int)blockIdx.x) / NB;
  for (int p = pmax; p >= 0; --p) {
    const int c0 = p * NB;
    const int nbp = min(NB, n - c0);
    wy_apply<MAXR, NB, false>(q, Vs + (int64_t)c0 * m, m, m, nbp, Tall + (int64_t)p * NB * NB, S);
  }
  if (live) {
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
      const int rr = lane + 64 * i;
      if (rr < m) A[(int64_t)rr * rs + (int64_t)c * cs] = q[i];
    }
  }
}

// ---------------------------------------------------------------------------
// Q from the reflectors with two MFMA GEMMs and one substitution.  The
// product of all n reflectors is H(0)...H(n-1) = I - V T V^H with
//   T^{-1} = diag(1/tau) + striu(V^H V)
// (T upper triangular; the UT form of the compact-WY factor), so
//   Q = H(0)...H(n-1) I(:, 0:n) = I(:, 0:n) - V W,   W = T V[0:n, :]^H.
// The GEMM G = V^H V comes first; qr_wsolve_kernel solves T^{-1} W = V[0:n,:]^H
// by backward substitution, w_i = tau_i (b_i - sum_{k>i} G[i][k] w_k)
// (tau_i = 0 gives w_i = 0), and copies R out; then one GEMM forms
// Q = I(:, 0:n) - V W over the panel, the identity added in its epilogue.
constexpr int FW_NMAX = 256;       // n limit of the substitution kernel

// One wave per column c of W (4 per workgroup).  Lane l keeps rows
// l, l+64, ... of that column's running right-hand side; at step i the lane
// holding row i finishes w_i (and keeps it in place: rows above i are never
// touched again), a lane read broadcasts it, and every lane subtracts
// G[k][i] w_i from its rows k < i (column i of G's strict upper triangle;
// the staged zeros leave the finished rows k >= i alone).  G is staged in LDS
// CH columns at a time, the next chunk's loads in flight during the current
// chunk's steps; each chunk's G entries and tau values move to registers,
// and the steps are unrolled per row slot, so a step is a short VALU chain.
template <int NS>
__global__ __launch_bounds__(256) void qr_wsolve_kernel(const double2* __restrict__ A, int m, int n, int64_t rs,
                                                        int64_t cs, const double2* __restrict__ Vs,
                                                        const double2* __restrict__ G,
                                                        const double2* __restrict__ tau, double2* __restrict__ W,
                                                        double2* __restrict__ Rout, int64_t bs) {
  {
    const int64_t bo = blockIdx.y * bs;   // lockstep copy
    A = boff(A, bo);
    Vs = boff(Vs, bo);
    G = boff(G, bo);
    tau = boff(tau, bo);
    W = boff(W, bo);
    if (Rout != nullptr) Rout = boff(Rout, bo);
  }
  constexpr int NR = 64 * NS;                  // rows covered (n <= NR)
  constexpr int CH = (NS <= 2) ? 16 : 8;       // G columns per chunk (divides 64)
  constexpr int LPT = (CH * NR) / 256;         // staged elements per thread per chunk
  __shared__ double2 sG[CH][NR];
  __shared__ double2 sTau[NR];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int c = 4 * (int)blockIdx.x + wave;
  const int cl = min(c, n - 1);
  // R (upper triangle of the packed panel) out before Q overwrites it
  if (Rout != nullptr) {
    for (int64_t e = (int64_t)blockIdx.x * 256 + tid; e < (int64_t)n * n; e += (int64_t)gridDim.x * 256) {
      const int r = (int)(e % n), q = (int)(e / n);
      Rout[e] = (r <= q) ? A[(int64_t)r * rs + (int64_t)q * cs] : make_double2(0.0, 0.0);
    }
  }
  for (int r = tid; r < NR; r += 256) sTau[r] = (r < n) ? tau[r] : make_double2(0.0, 0.0);
  // b = V[c, 0:n]^H (row c of V, conjugated): b_k = conj(v_k[c])
  double2 acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int r = lane + 64 * s;
    const double2 x = Vs[(int64_t)min(r, n - 1) * m + cl];
    acc[s] = (r < n) ? make_double2(x.x, -x.y) : make_double2(0.0, 0.0);
  }
  // chunk q holds G[0:NR, q*CH + cc] (cc < CH), zero on and below the diagonal
  double2 reg[LPT];
  auto fetch = [&](int q) {
#pragma unroll
    for (int t = 0; t < LPT; ++t) {
      const int e = tid + 256 * t, col = q * CH + e / NR, r = e % NR;
      const double2 x = G[(int64_t)min(col, n - 1) * n + min(r, n - 1)];
      reg[t] = (col < n && r < col) ? x : make_double2(0.0, 0.0);
    }
  };
  auto put = [&]() {
#pragma unroll
    for (int t = 0; t < LPT; ++t) {
      const int e = tid + 256 * t;
      sG[e / NR][e % NR] = reg[t];
    }
  };
  // W is upper triangular (w_r = 0 in column c for r > c: b_r = conj(v_r[c])
  // is zero there, and a zero w_r changes no other row), so the chunks above
  // this workgroup's last column are skipped; the result is the same
  const int qtop = min(4 * (int)blockIdx.x + 3, n - 1) / CH;
  fetch(qtop);
  put();
  __syncthreads();
  for (int q = qtop; q >= 0; --q) {
    if (q > 0) fetch(q - 1);
    double2 gv[CH][NS], tv[CH];
#pragma unroll
    for (int cc = 0; cc < CH; ++cc) {
      tv[cc] = sTau[q * CH + cc];
#pragma unroll
      for (int s = 0; s < NS; ++s) gv[cc][s] = sG[cc][lane + 64 * s];
    }
    const int sl = (q * CH) >> 6, ob = (q * CH) & 63;   // row slot and first owner lane of the chunk
    static_for<0, NS>([&](auto sc) {
      constexpr int S = decltype(sc)::value;
      if (sl == S) {
#pragma unroll
        for (int cc = CH - 1; cc >= 0; --cc) {
          // an identity reflector (tau = 0) has w_i = 0 and changes no other
          // row: its step is skipped (rank-deficient panels skip most steps)
          if (q * CH + cc < n && (tv[cc].x != 0.0 || tv[cc].y != 0.0)) {
            const int owner = ob + cc;
            const double2 wl = cmul(tv[cc], acc[S]);
            const int lo_r = __builtin_amdgcn_readlane(__double2loint(wl.x), owner);
            const int hi_r = __builtin_amdgcn_readlane(__double2hiint(wl.x), owner);
            const int lo_i = __builtin_amdgcn_readlane(__double2loint(wl.y), owner);
            const int hi_i = __builtin_amdgcn_readlane(__double2hiint(wl.y), owner);
            const double2 wi = make_double2(__hiloint2double(hi_r, lo_r), __hiloint2double(hi_i, lo_i));
#pragma unroll
            for (int s = 0; s < NS; ++s) {
              const double2 g = gv[cc][s];
              const double2 u = make_double2(fma(g.y, wi.y, fma(-g.x, wi.x, acc[s].x)),
                                             fma(-g.y, wi.x, fma(-g.x, wi.y, acc[s].y)));
              if (s == S)
                acc[s] = (lane == owner) ? wl : u;
              else
                acc[s] = u;
            }
          }
        }
      }
    });
    if (q > 0) {
      __syncthreads();
      put();
      __syncthreads();
    }
  }
  if (c < n) {
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int r = lane + 64 * s;
      if (r < n) {
        const double2 tr = sTau[r];   // rows of skipped steps: w_r = 0
        W[(int64_t)c * n + r] = (tr.x != 0.0 || tr.y != 0.0) ? acc[s] : make_double2(0.0, 0.0);
      }
    }
  }
}

struct QrGemms {
  Gemm gram, q;
};
static QrGemms qr_gemms(double2* A, int64_t m, int64_t n, int64_t rs, int64_t cs, const double2* Vs, double2* G,
                        double2* W) {
  QrGemms g;
  // G = V^H V
  g.gram.M = n; g.gram.N = n; g.gram.K = m;
  g.gram.A = Vs; g.gram.am = im1(m); g.gram.ak = im1(1); g.gram.conjA = true;
  g.gram.B = Vs; g.gram.bk = im1(1); g.gram.bn = im1(m);
  g.gram.C = G; g.gram.cm = im1(1); g.gram.cn = im1(n);
  // Q = I(:, 0:n) - V W, in place over the reset panel
  g.q.M = m; g.q.N = n; g.q.K = n;
  g.q.A = Vs; g.q.am = im1(1); g.q.ak = im1(m);
  g.q.B = W; g.q.bk = im1(1); g.q.bn = im1(n);
  g.q.C = A; g.q.cm = im1(rs); g.q.cn = im1(cs);
  g.q.alpha_re = -1.0; g.q.diag_re = 1.0;
  return g;
}

// R rows per thread and panel width NB in the panel kernel (at most NTMAX
// threads); MAXR rows per lane in the WY kernels.
struct QrWork {
  double2 *Vs, *tau, *Tall, *G, *W;
  double* colnrm;
  unsigned* prog;
  void* gws;
  size_t gws_bytes;
  int num_cus;
};

// Q = I(:, 0:n) - V W over the panel from the reflectors in q.Vs / q.tau (n <= FW_NMAX)
static void qr_form_q(double2* A, int m, int n, int64_t rs, int64_t cs, const QrWork& q, double2* R_out,
                      hipStream_t s, Bat bat) {
  const unsigned nb = (unsigned)bat.nb;
  const int64_t bs = bat.bs;
  QrGemms g = qr_gemms(A, m, n, rs, cs, q.Vs, q.G, q.W);
  g.gram.bat = bat;
  g.q.bat = bat;
  launch_gemm(g.gram, CMP_F64, q.gws, q.gws_bytes, q.num_cus, s);
  const dim3 grid((unsigned)((n + 3) / 4), nb);
  if (n <= 64)
    MPS_LAUNCH((qr_wsolve_kernel<1>), grid, dim3(256), 0, s, A, m, n, rs, cs, q.Vs, q.G, q.tau, q.W, R_out, bs);
  else if (n <= 128)
    MPS_LAUNCH((qr_wsolve_kernel<2>), grid, dim3(256), 0, s, A, m, n, rs, cs, q.Vs, q.G, q.tau, q.W, R_out, bs);
  else
    MPS_LAUNCH((qr_wsolve_kernel<4>), grid, dim3(256), 0, s, A, m, n, rs, cs, q.Vs, q.G, q.tau, q.W, R_out, bs);
  launch_gemm(g.q, CMP_F64, q.gws, q.gws_bytes, q.num_cus, s);
}

// The chain factorization (qr_chain_kernel) then the Q formation; needs
// n <= min(NT, FW_NMAX) and m within R * NT rows.
template <int R, int NT>
static void qr_chain(double2* A, int m, int n, int64_t rs, int64_t cs, double tol, const QrWork& q, double2* R_out,
                     hipStream_t s, Bat bat) {
  const unsigned nb = (unsigned)bat.nb;
  const int64_t bs = bat.bs;
  const int P = (n + CH_NB - 1) / CH_NB;
  MPS_LAUNCH(qr_colnorm_kernel, dim3((unsigned)n, nb), dim3(CN_NT), 0, s, A, m, rs, cs, q.colnrm, q.prog, P + 1,
             q.Vs, q.tau, tol > 0.0, bs);
  const unsigned blocks = (nb + CH_WG_PER_PANEL - 1) / CH_WG_PER_PANEL;
  MPS_LAUNCH((qr_chain_kernel<R, NT>), dim3((unsigned)(CH_WG_PER_PANEL * P) * blocks), dim3(NT), 0, s, A, m, n, rs,
             cs, tol, q.colnrm, q.Vs, q.tau, q.prog, nb, bs);
  qr_form_q(A, m, n, rs, cs, q, R_out, s, bat);
}

// MPS_QR_CHAIN=0 in the environment selects the launch-per-panel factorization
static bool chain_enabled() {
  static const bool on = [] {
    const char* e = getenv("MPS_QR_CHAIN");
    return e == nullptr || atoi(e) != 0;
  }();
  return on;
}

// The chain kernel needs the workgroups of a launch dealt to the 8 XCDs round
// robin (so each block of 8 P workgroups puts exactly P on every XCD, one
// factorization's panels per XCD).  qr_chain_probe runs once per device, from
// context creation (never under graph capture), and records whether 64
// workgroups land on XCD x % 8; the chain path is used only where they did.
// (Were a launch still to miss an XCD's share, the panels no workgroup took
// keep the NaN tau the column-norm kernel wrote, and Q shows NaN.)
constexpr int PROBE_WG = 64;
__global__ void xcc_probe_kernel(unsigned* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}
static int g_chain_ok[64] = {};   // per device: 0 not probed, 1 round robin from XCD 0, -1 not
static std::mutex g_chain_mu;
void qr_chain_probe(int device) {
  if (device < 0 || device >= 64) return;
  std::lock_guard<std::mutex> lk(g_chain_mu);
  if (g_chain_ok[device] != 0) return;
  unsigned* d = nullptr;
  unsigned h[PROBE_WG];
  HIPCHK(hipMalloc((void**)&d, PROBE_WG * sizeof(unsigned)));
  MPS_LAUNCH(xcc_probe_kernel, dim3(PROBE_WG), dim3(64), 0, (hipStream_t)0, d);
  const hipError_t e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  HIPCHK(hipFree(d));
  HIPCHK(e);
  bool ok = true;
  for (int i = 0; i < PROBE_WG; ++i) ok = ok && h[i] == (unsigned)(i % CH_WG_PER_PANEL);
  g_chain_ok[device] = ok ? 1 : -1;
}
static bool chain_fits(int64_t n, int nt, const Bat& bat, int chain) {
  if (!(chain < 0 ? chain_enabled() : chain != 0) || n > nt || n > FW_NMAX) return false;
  if ((int64_t)(bat.nb + CH_WG_PER_PANEL - 1) * ((n + CH_NB - 1) / CH_NB) > 0x7FFFFFFF) return false;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  return g_chain_ok[dev] == 1;
}

// PT: the panel kernel forms each panel's T (16-column panels, where the
// apply kernel's own Gram reduction would cost more than the panel's);
// otherwise the apply kernel reduces the Gram entries with V^H a_j, and the T
// factors for the Q formation come from the Gram GEMM's diagonal blocks.
template <int R, int NB, int NT, int MAXR, int RA, int NTA>
static void qr_blocked(double2* A, int m, int n, int64_t rs, int64_t cs, double tol, const QrWork& q,
                       double2* R_out, hipStream_t s, Bat bat) {
  constexpr bool PT = (NB == 16);
  double2 *Vs = q.Vs, *tau = q.tau, *Tall = q.Tall;
  double* colnrm = q.colnrm;
  const unsigned nb = (unsigned)bat.nb;
  const int64_t bs = bat.bs;
  if (tol > 0.0)
    MPS_LAUNCH(qr_colnorm_kernel, dim3((unsigned)n, nb), dim3(CN_NT), 0, s, A, m, rs, cs, colnrm,
               (unsigned*)nullptr, 0, (double2*)nullptr, (double2*)nullptr, true, bs);
  for (int c0 = 0; c0 < n; c0 += NB) {
    const int nbp = std::min(NB, n - c0);
    double2* Tp = Tall + (int64_t)(c0 / NB) * NB * NB;
    MPS_LAUNCH((qr_panel_kernel<R, NB, NT, PT>), dim3(1, nb), dim3(NT), 0, s, A, m, n, rs, cs, c0, nbp,
                       tol, colnrm, Vs, tau, Tp, bs);
    const int trail = n - c0 - nbp;
    if (trail > 0)
      MPS_LAUNCH((qr_apply_col_kernel<RA, NB, NTA, !PT>), dim3((unsigned)trail, nb), dim3(NTA), 0, s, A,
                         m, rs, cs, c0, nbp, Vs, Tp, tau, bs);
  }
  if (n <= FW_NMAX) {
    qr_form_q(A, m, n, rs, cs, q, R_out, s, bat);
  } else {
    QrGemms g = qr_gemms(A, m, n, rs, cs, Vs, q.G, q.W);
    g.gram.bat = bat;
    if (!PT) {
      launch_gemm(g.gram, CMP_F64, q.gws, q.gws_bytes, q.num_cus, s);
      MPS_LAUNCH((qr_tfactor_kernel<NB>), dim3((unsigned)((n + NB - 1) / NB), nb), dim3(64), 0, s, q.G,
                         n, tau, Tall, bs);
    }
    MPS_LAUNCH((qr_formq_kernel<MAXR, NB>), dim3((unsigned)((n + 3) / 4), nb), dim3(WY_NT), 0, s, A, m,
                       rs, cs, Vs, Tall, R_out, n, bs);
  }
}

// V (m x n) | tau (n) | column norms (n) | T blocks (ceil(n/NB) x NB x NB,
// NB <= 16) | G (n x n) | W (n x n) | split-K partials of the two GEMMs
struct QrLayout {
  size_t vs, tau, cn, prog, tall, g, w, gws, total, gws_bytes;
};
static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }
static QrLayout qr_layout(int64_t m, int64_t n, int num_cus) {
  QrLayout L;
  size_t o = 0;
  L.vs = o; o = al256(o + (size_t)m * n * sizeof(double2));
  L.tau = o; o = al256(o + (size_t)n * sizeof(double2));
  L.cn = o; o = al256(o + (size_t)n * sizeof(double));
  L.prog = o; o = al256(o + (size_t)((n + CH_NB - 1) / CH_NB + 1) * sizeof(unsigned));
  L.tall = o; o = al256(o + (size_t)(n + 16) * 16 * sizeof(double2));
  L.g = o; o = al256(o + (size_t)n * n * sizeof(double2));
  L.w = o; o = al256(o + (size_t)n * n * sizeof(double2));
  // split-K partials: a bound that grows with (m, n) (at most 32 splits of
  // the n x n Gram, ceil(n/32) of the m x n product: chunks are >= 32 deep),
  // so a workspace sized for the largest panel of a state serves every smaller one
  const QrGemms g = qr_gemms(nullptr, m, n, 1, m, nullptr, nullptr, nullptr);
  const size_t need =
      std::max(gemm_workspace_bytes(g.gram, CMP_F64, num_cus), gemm_workspace_bytes(g.q, CMP_F64, num_cus));
  const size_t bound = std::max((size_t)32 * n * n, (size_t)((n + 31) / 32) * m * n) * sizeof(double2);
  L.gws_bytes = std::max(need, bound);
  L.gws = o; o = al256(o + L.gws_bytes);
  L.total = o;
  return L;
}

size_t qr_workspace_bytes(int64_t m, int64_t n, int num_cus) { return qr_layout(m, n, num_cus).total; }

// (plain contexts: the RK4 schedule lifts its third stream's GEMM cap where
// the bulk sites' QR is this single launch on one XCD)
bool qr_single_launch_applies(int64_t m, int64_t n, const Bat& bat, int chain) {
  return bat.nb == 1 && ((m > 768 && m <= 1280 && chain_fits(n, 256, bat, chain)) ||
                         (m > 1280 && m <= 2560 && chain_fits(n, 512, bat, chain)));
}

void launch_qr(void* base, int64_t m, int64_t n, int64_t rs, int64_t cs, double tol, void* ws, int num_cus,
               void* R_out, hipStream_t s, Bat bat, int chain) {
  REQUIRE(m >= n && n >= 1, "qrq!: requires m >= n >= 1");
  REQUIRE(bat.nb >= 1 && bat.nb <= 65535, "qrq!: too many lockstep copies");
  REQUIRE(m < (1LL << 31) && m * n < (1LL << 40), "qrq!: panel too large");
  char* w = (char*)ws;
  const QrLayout L = qr_layout(m, n, num_cus);
  QrWork q;
  q.Vs = (double2*)(w + L.vs);
  q.tau = (double2*)(w + L.tau);
  q.colnrm = (double*)(w + L.cn);
  q.prog = (unsigned*)(w + L.prog);
  q.Tall = (double2*)(w + L.tall);
  q.G = (double2*)(w + L.g);
  q.W = (double2*)(w + L.w);
  q.gws = L.gws_bytes ? (void*)(w + L.gws) : nullptr;
  q.gws_bytes = L.gws_bytes;
  q.num_cus = num_cus;
  double2* A = (double2*)base;
  double2* R = (double2*)R_out;
  // <panel rows per thread, panel width, panel threads, WY rows per lane
  //  (Q formation for n > 256), apply rows per thread, apply threads>; each
  // variant needs m within rows per thread * threads and 64 * WY rows
  // one wave per SIMD (256 threads) where the panel fits: the column loop is
  // VALU-bound, and a SIMD with two waves carries twice the work; 8-column
  // panels keep the rows of a thread in registers up to m = 1280 (measured:
  // m = 1152 panel 3.6 us/column vs 5.7 with 16 columns on 6 waves; at m = 256
  // the 16-column panel wins, at m = 2304 two waves per SIMD beat AGPR storage)
  // (m <= 256: the single launch for pools of up to 64 copies; above that the
  // per-panel launches' WY updates over every copy's columns take less time,
  // measured: 256 copies of 64 x 32 in 205 us against 278.  The pool size,
  // not the copies of this launch, decides, so a ragged pool's copy factors
  // its panels the same way at every site)
  if (m <= 256 && n > 16 && (bat.plan_nb > 0 ? bat.plan_nb : bat.nb) <= 64 && chain_fits(n, 256, bat, chain))
    qr_chain<1, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 256)
    qr_blocked<1, 16, 256, 4, 1, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 512 && chain_fits(n, 256, bat, chain))
    qr_chain<2, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 512)
    qr_blocked<2, 8, 256, 8, 2, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 768 && chain_fits(n, 256, bat, chain))
    qr_chain<3, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 768)
    qr_blocked<3, 8, 256, 12, 3, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 1280 && chain_fits(n, 256, bat, chain))
    qr_chain<5, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 1280)
    qr_blocked<5, 8, 256, 20, 5, 256>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 2560 && chain_fits(n, 512, bat, chain))
    qr_chain<5, 512>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else if (m <= 2560)
    qr_blocked<5, 8, 512, 40, 5, 512>(A, (int)m, (int)n, rs, cs, tol, q, R, s, bat);
  else
    MPS_LAUNCH(qr_householder_kernel, dim3(1, (unsigned)bat.nb), dim3(QR_NT), 0, s, A, (int)m, (int)n, rs,
                       cs, tol, q.tau, R, bat.bs);
  HIPCHK(hipGetLastError());
}

}  // namespace mpship
