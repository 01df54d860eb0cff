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
// Work split: one workgroup of 1024 threads per QR; 16 waves each own a
// strided subset of the trailing columns (wavefront-shuffle dot products),
// reflector norms use a workgroup reduction.
#include "common.h"

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
                                                               double2* __restrict__ Rout) {
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
// Panel-blocked form (the default for m <= QB_MAXR * 64).
//
// Same per-column operation order as zgeqr2 / zung2r: trailing column j
// receives reflectors 0, 1, ..., j-1 in order; Q column c receives H(c),
// H(c-1), ..., H(0) in order.  Only the schedule differs:
//   qr_colnorm  one wave per column: ||a_c||^2 (rank cutoff only)
//   qr_panel    one workgroup factors a 32-column panel held in registers
//               (wave w owns panel columns w and w+16; one LDS barrier per
//               reflector, double-buffered reflector in LDS)
//   qr_apply    one wave per trailing column applies the panel's reflectors
//   qr_formq    one wave per column of Q (reflectors read from a separate
//               buffer, so Q can overwrite the panel in place)
constexpr int QB_NT = 1024;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconjmul(double2 a, double2 b) {   // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

__global__ __launch_bounds__(64) void qr_colnorm_kernel(const double2* __restrict__ A, int m, int64_t rs,
                                                        int64_t cs, double* __restrict__ colnrm) {
  const int c = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int r = lane; r < m; r += 64) {
    const double2 v = A[(int64_t)r * rs + (int64_t)c * cs];
    s += v.x * v.x + v.y * v.y;
  }
  s = wave_sum(s);
  if (lane == 0) colnrm[c] = s;
}

// zlarfg on the owner wave's register column; the reflector goes to LDS.
template <int MAXR>
__device__ __forceinline__ void panel_reflector(double2 (&a)[MAXR], int lane, int rk, int M, double thr,
                                                double2* __restrict__ vb, double2& t_out) {
  double xs = 0.0, ar = 0.0, ai = 0.0;
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    const double n2 = a[i].x * a[i].x + a[i].y * a[i].y;
    xs += (rr > rk && rr < M) ? n2 : 0.0;
    ar += (rr == rk) ? a[i].x : 0.0;
    ai += (rr == rk) ? a[i].y : 0.0;
  }
  xs = wave_sum(xs);
  ar = wave_sum(ar);
  ai = wave_sum(ai);
  const double xnorm = sqrt(xs);
  const bool ident = (thr > 0.0) ? (xnorm <= thr && fabs(ai) <= thr) : (xs == 0.0 && ai == 0.0);
  double2 t = make_double2(0.0, 0.0), sc = make_double2(1.0, 0.0);
  double beta = ar;
  if (!ident) {
    beta = -copysign(sqrt(ar * ar + ai * ai + xs), ar);
    t = make_double2((beta - ar) / beta, -ai / beta);
    const double dr = ar - beta, di = ai, den = dr * dr + di * di;
    sc = make_double2(dr / den, -di / den);
  }
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    if (rr == rk) {
      if (!ident) a[i] = make_double2(beta, 0.0);
      vb[rr] = make_double2(1.0, 0.0);
    } else if (rr > rk) {
      if (!ident) a[i] = cmul(a[i], sc);
      vb[rr] = (rr < M) ? a[i] : make_double2(0.0, 0.0);
    }
  }
  t_out = t;
}

// a <- a - conj(tau) v (v^H a), v from LDS (rows >= rk), unconditional reads, masked use
template <int MAXR>
__device__ __forceinline__ void apply_reflector_lds(double2 (&a)[MAXR], int lane, int rk, int M,
                                                    const double2* __restrict__ v, double2 t) {
  double wr = 0.0, wi = 0.0;
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    const double2 x = v[rr];
    const bool ok = (rr >= rk && rr < M);
    wr += ok ? (x.x * a[i].x + x.y * a[i].y) : 0.0;
    wi += ok ? (x.x * a[i].y - x.y * a[i].x) : 0.0;
  }
  wr = wave_sum(wr);
  wi = wave_sum(wi);
  const double fr = t.x * wr + t.y * wi, fi = t.x * wi - t.y * wr;   // conj(tau) w
  // re-read v from LDS instead of keeping MAXR values live across the reduction
  asm volatile("" ::: "memory");
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    const double2 x = v[rr];
    if (rr >= rk && rr < M) {
      a[i].x -= x.x * fr - x.y * fi;
      a[i].y -= x.x * fi + x.y * fr;
    }
  }
}

// One workgroup factors panel columns c0 .. c0+nbp-1 (nbp <= NW), rows c0..m-1;
// wave w holds panel column w in registers.  Only LDS traffic inside the
// column loop.  Waves whose column is already factored accumulate the Gram
// entries G[w][k] = v_w^H v_k while the others apply reflector k, so the
// compact-WY factor T (H(0)...H(nbp-1) = I - V T V^H, zlarft) needs no
// further pass over V.
template <int MAXR, int NT>
__global__ __launch_bounds__(NT) void qr_panel_kernel(double2* __restrict__ A, int m, int n, int64_t rs,
                                                      int64_t cs, int c0, int nbp, double tol,
                                                      const double* __restrict__ colnrm,
                                                      double2* __restrict__ Vs, double2* __restrict__ tau,
                                                      double2* __restrict__ Tp) {
  constexpr int NW = NT / 64;
  __shared__ double2 vbuf[2][MAXR * 64];
  __shared__ double2 s_tau[NW];
  __shared__ double2 s_G[NW][NW];
  __shared__ double2 s_T[NW][NW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int M = m - c0;
  double thr = 0.0;
  if (tol > 0.0) {
    double t = 0.0;
    for (int c = 0; c < n; ++c) t += colnrm[c];
    thr = tol * sqrt(t);
  }
  const int j = wave;
  double2 a[MAXR];
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    const double2 x = A[(int64_t)(c0 + min(rr, M - 1)) * rs + (int64_t)(c0 + min(j, nbp - 1)) * cs];
    a[i] = (rr < M && j < nbp) ? x : make_double2(0.0, 0.0);
  }
  for (int k = 0; k < nbp; ++k) {
    const int buf = k & 1;
    if (wave == k) {
      double2 t;
      panel_reflector<MAXR>(a, lane, k, M, thr, vbuf[buf], t);
      if (lane == 0) s_tau[k] = t;
    }
    __syncthreads();
    const double2 t = s_tau[k];
    if (j > k && j < nbp) {
      if (t.x != 0.0 || t.y != 0.0) apply_reflector_lds<MAXR>(a, lane, k, M, vbuf[buf], t);
    } else if (j < k) {
      // G[j][k] = v_j^H v_k over rows >= k (v_k vanishes above k; row j of v_j is 1 < k)
      double gr = 0.0, gi = 0.0;
#pragma unroll
      for (int i = 0; i < MAXR; ++i) {
        const int rr = lane + 64 * i;
        const double2 x = vbuf[buf][rr];
        const bool ok = (rr >= k && rr < M);
        gr += ok ? (a[i].x * x.x + a[i].y * x.y) : 0.0;
        gi += ok ? (a[i].x * x.y - a[i].y * x.x) : 0.0;
      }
      gr = wave_sum(gr);
      gi = wave_sum(gi);
      if (lane == 0) s_G[j][k] = make_double2(gr, gi);
    }
  }
  __syncthreads();
  // T: T[k][k] = tau_k, T[0:k, k] = -tau_k T[0:k, 0:k] G[0:k, k]
  if (wave == 0) {
    for (int k = 0; k < nbp; ++k) {
      const double2 tk = s_tau[k];
      double2 acc = make_double2(0.0, 0.0);
      if (lane < k)
        for (int l = lane; l < k; ++l) {
          const double2 p = cmul(s_T[lane][l], s_G[l][k]);
          acc.x += p.x;
          acc.y += p.y;
        }
      __builtin_amdgcn_wave_barrier();
      if (lane < k) {
        const double2 v = cmul(tk, acc);
        s_T[lane][k] = make_double2(-v.x, -v.y);
      } else if (lane == k) {
        s_T[k][k] = tk;
      } else if (lane < NW) {
        s_T[lane][k] = make_double2(0.0, 0.0);
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
  // packed panel back to A; explicit reflectors to Vs (column c0+j: zero rows
  // 0..c0+j-1, 1 at row c0+j, the tail below; an identity reflector stores 0)
  if (j < nbp) {
    const bool id = (s_tau[j].x == 0.0 && s_tau[j].y == 0.0);
    double2* vcol = Vs + (int64_t)(c0 + j) * m;
    for (int r = lane; r < c0; r += 64) vcol[r] = make_double2(0.0, 0.0);
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
      const int rr = lane + 64 * i;
      if (rr < M) {
        A[(int64_t)(c0 + rr) * rs + (int64_t)(c0 + j) * cs] = a[i];
        vcol[c0 + rr] = (rr < j) ? make_double2(0.0, 0.0)
                                 : (rr == j ? make_double2(1.0, 0.0) : (id ? make_double2(0.0, 0.0) : a[i]));
      }
    }
  }
  if (tid < nbp) tau[c0 + tid] = s_tau[tid];
  __syncthreads();
  for (int e = tid; e < NW * NW; e += NT) {
    const int r = e % NW, c = e / NW;
    Tp[e] = (r < nbp && c < nbp) ? s_T[r][c] : make_double2(0.0, 0.0);
  }
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
template <int MAXR, int NB>
__global__ __launch_bounds__(WY_NT) void qr_apply_kernel(double2* __restrict__ A, int m, int64_t rs, int64_t cs,
                                                         int c0, int nbp, int ntrail,
                                                         const double2* __restrict__ Vs,
                                                         const double2* __restrict__ Tp) {
  __shared__ WyStage<NB> S;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int jj = 4 * blockIdx.x + wave;
  const int j = c0 + nbp + min(jj, ntrail - 1);
  const int M = m - c0;
  double2 x[MAXR];
#pragma unroll
  for (int i = 0; i < MAXR; ++i) {
    const int rr = lane + 64 * i;
    const double2 v = A[(int64_t)(c0 + min(rr, M - 1)) * rs + (int64_t)j * cs];
    x[i] = (rr < M) ? v : make_double2(0.0, 0.0);
  }
  wy_apply<MAXR, NB, true>(x, Vs + (int64_t)c0 * m + c0, m, M, nbp, Tp, S);
  if (jj < ntrail) {
#pragma unroll
    for (int i = 0; i < MAXR; ++i) {
      const int rr = lane + 64 * i;
      if (rr < M) A[(int64_t)(c0 + rr) * rs + (int64_t)j * cs] = x[i];
    }
  }
}

// Q = H(0) ... H(n-1) I(:, 0:n): 4 columns per workgroup (same panel), panel
// blocks from the columns' own panel down to 0, over all m rows (V is zero
// above each panel).
template <int MAXR, int NB>
__global__ __launch_bounds__(WY_NT) void qr_formq_kernel(double2* __restrict__ A, int m, int64_t rs, int64_t cs,
                                                         const double2* __restrict__ Vs,
                                                         const double2* __restrict__ Tall,
                                                         double2* __restrict__ Rout, int n) {
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

template <int MAXR, int NT>
static void qr_blocked(double2* A, int m, int n, int64_t rs, int64_t cs, double tol, double2* Vs,
                       double2* tau, double* colnrm, double2* Tall, double2* R_out, hipStream_t s) {
  constexpr int NB = NT / 64;
  if (tol > 0.0) hipLaunchKernelGGL(qr_colnorm_kernel, dim3((unsigned)n), dim3(64), 0, s, A, m, rs, cs, colnrm);
  for (int c0 = 0; c0 < n; c0 += NB) {
    const int nbp = std::min(NB, n - c0);
    double2* Tp = Tall + (int64_t)(c0 / NB) * NB * NB;
    hipLaunchKernelGGL((qr_panel_kernel<MAXR, NT>), dim3(1), dim3(NT), 0, s, A, m, n, rs, cs, c0, nbp, tol,
                       colnrm, Vs, tau, Tp);
    const int trail = n - c0 - nbp;
    if (trail > 0)
      hipLaunchKernelGGL((qr_apply_kernel<MAXR, NB>), dim3((unsigned)((trail + 3) / 4)), dim3(WY_NT), 0, s, A,
                         m, rs, cs, c0, nbp, trail, Vs, Tp);
  }
  hipLaunchKernelGGL((qr_formq_kernel<MAXR, NB>), dim3((unsigned)((n + 3) / 4)), dim3(WY_NT), 0, s, A, m, rs,
                     cs, Vs, Tall, R_out, n);
}

size_t qr_workspace_bytes(int64_t m, int64_t n) {
  // V (m x n) | tau (n) | column norms (n) | T blocks (ceil(n/NB) x NB x NB, NB <= 16)
  return (size_t)m * n * sizeof(double2) + (size_t)n * sizeof(double2) + (size_t)n * sizeof(double) +
         (size_t)(n + 16) * 16 * sizeof(double2) + 1024;
}

void launch_qr(void* base, int64_t m, int64_t n, int64_t rs, int64_t cs, double tol, void* ws,
               void* /*scal_ws*/, void* R_out, hipStream_t s) {
  REQUIRE(m >= n && n >= 1, "qrq!: requires m >= n >= 1");
  REQUIRE(m < (1LL << 31) && m * n < (1LL << 40), "qrq!: panel too large");
  char* w = (char*)ws;
  double2* Vs = (double2*)w;
  double2* tau = (double2*)(w + (size_t)m * n * sizeof(double2));
  double* colnrm = (double*)(w + (size_t)m * n * sizeof(double2) + (size_t)n * sizeof(double2));
  double2* Tall = (double2*)(w + (size_t)m * n * sizeof(double2) + (size_t)n * sizeof(double2) +
                             (((size_t)n * sizeof(double) + 15) & ~(size_t)15));
  double2* A = (double2*)base;
  double2* R = (double2*)R_out;
  // (rows per lane, panel workgroup size = 64 x panel width); a complex
  // double is 4 VGPRs and 16 waves of 1024 threads leave 128 VGPRs per lane
  if (m <= 640)
    qr_blocked<10, 1024>(A, (int)m, (int)n, rs, cs, tol, Vs, tau, colnrm, Tall, R, s);    // panel 16
  else if (m <= 1280)
    qr_blocked<20, 1024>(A, (int)m, (int)n, rs, cs, tol, Vs, tau, colnrm, Tall, R, s);    // panel 16
  else if (m <= 2560)
    qr_blocked<40, 512>(A, (int)m, (int)n, rs, cs, tol, Vs, tau, colnrm, Tall, R, s);     // panel 8
  else
    hipLaunchKernelGGL(qr_householder_kernel, dim3(1), dim3(QR_NT), 0, s, A, (int)m, (int)n, rs, cs, tol, tau,
                       R);
  HIPCHK(hipGetLastError());
}

}  // namespace mpship
