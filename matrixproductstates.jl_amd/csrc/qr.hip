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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

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

void launch_qr(void* base, int64_t m, int64_t n, int64_t rs, int64_t cs, double tol, void* tau_ws,
               void* /*scal_ws*/, void* R_out, hipStream_t s) {
  REQUIRE(m >= n && n >= 1, "qrq!: requires m >= n >= 1");
  REQUIRE(m < (1LL << 31) && m * n < (1LL << 40), "qrq!: panel too large");
  hipLaunchKernelGGL(qr_householder_kernel, dim3(1), dim3(QR_NT), 0, s, (double2*)base, (int)m,
                     (int)n, rs, cs, tol, (double2*)tau_ws, (double2*)R_out);
  HIPCHK(hipGetLastError());
}

}  // namespace mpship
