// One-sided (Hestenes) Jacobi SVD in FP64 for compressionSVD! (compression.jl:1280-1312).
//
// The reference calls LAPACK (Julia svd) on reshape(site * r, dl, dnew*d).  Here
// the columns of an m x n panel Y are made mutually orthogonal by plane
// rotations, Y J = W; then sigma_k = ||w_k|| and the singular vectors are
// W Sigma^-1 and J.  One-sided Jacobi keeps high relative accuracy for small
// singular values (nothing is squared).
//
// Schedule: a round-robin tournament pairs the n columns (n odd: one dummy),
// n' - 1 rounds per sweep, one workgroup per pair and one launch per round.
// Each workgroup forms alpha = ||y_p||^2, beta = ||y_q||^2, gamma = y_p^H y_q with
// a fixed-order reduction, skips the pair when |gamma| <= tol sqrt(alpha beta),
// and otherwise applies the rotation to columns p, q of Y and of J and sets
// the sweep's flag.  The host stops when a sweep sets no flag.
#include "common.h"

#include <algorithm>

namespace mpship {

constexpr int SVD_NT = 256;

// players at tournament positions: position 0 is fixed, the others rotate
__device__ __forceinline__ int svd_player(int pos, int round, int np) {
  return pos == 0 ? 0 : 1 + (pos - 1 + round) % (np - 1);
}

__global__ __launch_bounds__(SVD_NT) void svd_jacobi_round_kernel(double2* __restrict__ Y, int m, int n,
                                                                  double2* __restrict__ J, int round, int np,
                                                                  double tol, int* __restrict__ flag) {
  __shared__ double sh[4][SVD_NT / 64];
  int p = svd_player(blockIdx.x, round, np), q = svd_player(np - 1 - blockIdx.x, round, np);
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
  if (q >= n) return;   // paired with the dummy column
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double2* yp = Y + (int64_t)p * m;
  double2* yq = Y + (int64_t)q * m;
  double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
  for (int r = tid; r < m; r += SVD_NT) {
    const double2 x = yp[r], y = yq[r];
    a += x.x * x.x + x.y * x.y;
    b += y.x * y.x + y.y * y.y;
    gr += x.x * y.x + x.y * y.y;   // conj(x) y
    gi += x.x * y.y - x.y * y.x;
  }
  a = wave_sum_dpp(a);
  b = wave_sum_dpp(b);
  gr = wave_sum_dpp(gr);
  gi = wave_sum_dpp(gi);
  if (lane == 0) {
    sh[0][wave] = a;
    sh[1][wave] = b;
    sh[2][wave] = gr;
    sh[3][wave] = gi;
  }
  __syncthreads();
  a = b = gr = gi = 0.0;
#pragma unroll
  for (int w = 0; w < SVD_NT / 64; ++w) {
    a += sh[0][w];
    b += sh[1][w];
    gr += sh[2][w];
    gi += sh[3][w];
  }
  const double g = sqrt(gr * gr + gi * gi);
  if (!(g > tol * sqrt(a * b))) return;
  if (tid == 0) atomicOr(flag, 1);
  // gamma = g e^{i phi}; with y~_q = e^{-i phi} y_q the 2x2 Gram is real:
  // t solves t^2 + 2 zeta t - 1 = 0 (smaller root), zeta = (beta - alpha) / (2 g)
  const double zeta = (b - a) / (2.0 * g);
  const double t = copysign(1.0, zeta) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
  const double er = gr / g, ei = -gi / g;   // e^{-i phi}
  auto rot = [&](double2* cp, double2* cq, int len) {
    for (int r = tid; r < len; r += SVD_NT) {
      const double2 x = cp[r], y0 = cq[r];
      const double2 y = make_double2(er * y0.x - ei * y0.y, er * y0.y + ei * y0.x);
      cp[r] = make_double2(c * x.x - s * y.x, c * x.y - s * y.y);
      cq[r] = make_double2(s * x.x + c * y.x, s * x.y + c * y.y);
    }
  };
  rot(yp, yq, m);
  rot(J + (int64_t)p * n, J + (int64_t)q * n, n);
}

// sigma_k = ||Y[:, k]||, one workgroup per column, fixed reduction order
__global__ __launch_bounds__(SVD_NT) void svd_colnorm_kernel(const double2* __restrict__ Y, int m,
                                                             double* __restrict__ sigma) {
  __shared__ double sh[SVD_NT / 64];
  const double2* y = Y + (int64_t)blockIdx.x * m;
  double s = 0.0;
  for (int r = threadIdx.x; r < m; r += SVD_NT) s += y[r].x * y[r].x + y[r].y * y[r].y;
  s = wave_sum_dpp(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < SVD_NT / 64; ++w) t += sh[w];
    sigma[blockIdx.x] = sqrt(t);
  }
}

__global__ void svd_identity_kernel(double2* __restrict__ J, int n, int* __restrict__ flag) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)n * n;
       e += (int64_t)gridDim.x * blockDim.x)
    J[e] = make_double2((e % n) == (e / n) ? 1.0 : 0.0, 0.0);
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = 0;
}

__global__ void svd_clear_flag_kernel(int* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// Y (m x n) = X^H for X (n x m), or a plain copy (conj = false, X m x n)
__global__ void svd_load_kernel(double2* __restrict__ Y, const double2* __restrict__ X, int m, int n, int ct) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)m * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % m), c = (int)(e / m);
    if (ct) {
      const double2 x = X[c + (int64_t)n * r];
      Y[e] = make_double2(x.x, -x.y);
    } else {
      Y[e] = X[e];
    }
  }
}

// out[k + rows * col] = scale_k * (conj ? conj : id)(S[col + ld * perm[k]]), k < rows, col < cols
__global__ void svd_gather_kernel(double2* __restrict__ out, const double2* __restrict__ S, int64_t ld,
                                  const int* __restrict__ perm, const double* __restrict__ scale, int rows,
                                  int cols, int conj, int by_col) {
  // by_col: out[a + cols * k] = scale_k * S[a + ld * perm[k]]   (a < cols, k < rows)
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < (int64_t)rows * cols;
       e += (int64_t)gridDim.x * blockDim.x) {
    int k, col;
    if (by_col) {
      col = (int)(e % cols);
      k = (int)(e / cols);
    } else {
      k = (int)(e % rows);
      col = (int)(e / rows);
    }
    const double2 x = S[col + ld * perm[k]];
    const double f = scale ? scale[k] : 1.0;
    out[e] = make_double2(f * x.x, f * (conj ? -x.y : x.y));
  }
}

static unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 4096); }

void svd_load(void* Y, const void* X, int64_t m, int64_t n, bool conj_transpose, hipStream_t s) {
  MPS_LAUNCH(svd_load_kernel, dim3(grid_for(m * n)), dim3(256), 0, s, (double2*)Y, (const double2*)X,
                     (int)m, (int)n, conj_transpose ? 1 : 0);
  HIPCHK(hipGetLastError());
}

int svd_jacobi(void* Y, int64_t m, int64_t n, void* J, double tol, int* flag_dev, int* flag_host,
               int max_sweeps, hipStream_t s) {
  MPS_LAUNCH(svd_identity_kernel, dim3(grid_for(n * n)), dim3(256), 0, s, (double2*)J, (int)n, flag_dev);
  HIPCHK(hipGetLastError());
  if (n < 2) return 0;
  const int np = (int)(n + (n & 1));
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    if (sweep > 0) MPS_LAUNCH(svd_clear_flag_kernel, dim3(1), dim3(64), 0, s, flag_dev);
    for (int round = 0; round < np - 1; ++round)
      MPS_LAUNCH(svd_jacobi_round_kernel, dim3((unsigned)(np / 2)), dim3(SVD_NT), 0, s, (double2*)Y,
                         (int)m, (int)n, (double2*)J, round, np, tol, flag_dev);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(flag_host, flag_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (*flag_host == 0) return sweep + 1;
  }
  throw Error(-3, "compressionSVD!: Jacobi SVD did not converge");
}

void svd_colnorms(const void* Y, int64_t m, int64_t n, double* sigma, hipStream_t s) {
  MPS_LAUNCH(svd_colnorm_kernel, dim3((unsigned)n), dim3(SVD_NT), 0, s, (const double2*)Y, (int)m, sigma);
  HIPCHK(hipGetLastError());
}

void svd_gather(void* out, const void* S, int64_t ld, const int* perm, const double* scale, int64_t rows,
                int64_t cols, bool conj, bool by_col, hipStream_t s) {
  MPS_LAUNCH(svd_gather_kernel, dim3(grid_for(rows * cols)), dim3(256), 0, s, (double2*)out,
                     (const double2*)S, ld, perm, scale, (int)rows, (int)cols, conj ? 1 : 0, by_col ? 1 : 0);
  HIPCHK(hipGetLastError());
}

}  // namespace mpship
