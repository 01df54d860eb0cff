// Small device kernels: boundary-coefficient writes, norms, rescaling, copies.
#include "common.h"

#include <atomic>
#include <mutex>
#include <unordered_map>

namespace mpship {

namespace {
struct LaunchLimits {
  int max_threads;
  size_t static_lds;
};
std::mutex g_launch_mu;
std::unordered_map<const void*, LaunchLimits> g_launch_cache;
std::atomic<uint64_t> g_launches{0};
size_t g_device_lds = 0;
}  // namespace

void check_launch(const void* kernel, const char* name, dim3 grid, dim3 block, size_t dyn_lds) {
  auto fail = [&](const std::string& why) {
    throw Error(-4, std::string("invalid launch of ") + name + ": " + why);
  };
  if (grid.x < 1 || grid.y < 1 || grid.z < 1) fail("empty grid");
  if (block.x < 1 || block.y < 1 || block.z < 1) fail("empty block");
  const uint64_t threads = (uint64_t)block.x * block.y * block.z;
  if ((uint64_t)grid.x * block.x > 0xffffffffull || (uint64_t)grid.y * block.y > 0xffffffffull ||
      (uint64_t)grid.z * block.z > 0xffffffffull)
    fail("grid size exceeds 32 bits");
  LaunchLimits lim;
  {
    std::lock_guard<std::mutex> lk(g_launch_mu);
    auto it = g_launch_cache.find(kernel);
    if (it == g_launch_cache.end()) {
      hipFuncAttributes a;
      HIPCHK(hipFuncGetAttributes(&a, kernel));
      if (g_device_lds == 0) {
        int dev = 0, v = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        g_device_lds = (size_t)v;
      }
      it = g_launch_cache.emplace(kernel, LaunchLimits{a.maxThreadsPerBlock, a.sharedSizeBytes}).first;
    }
    lim = it->second;
  }
  if (threads > (uint64_t)lim.max_threads)
    fail("block of " + std::to_string(threads) + " threads, kernel maximum " + std::to_string(lim.max_threads));
  if (lim.static_lds + dyn_lds > g_device_lds)
    fail("LDS " + std::to_string(lim.static_lds + dyn_lds) + " bytes, device limit " + std::to_string(g_device_lds));
  g_launches.fetch_add(1, std::memory_order_relaxed);
}

uint64_t launches_checked() { return g_launches.load(std::memory_order_relaxed); }


// Batched copies: blockIdx.y selects the copy (byte offset blockIdx.y * bs).
__global__ void set_scalar_kernel(double2* p, double re, double im, int64_t bs) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *boff(p, blockIdx.y * bs) = make_double2(re, im);
}

void launch_set_scalar(void* p, double re, double im, hipStream_t s, Bat bat) {
  MPS_LAUNCH(set_scalar_kernel, dim3(1, bat.nb), dim3(64), 0, s, (double2*)p, re, im, bat.bs);
  HIPCHK(hipGetLastError());
}

__global__ void fill_zero_kernel(double2* p, int64_t n, int64_t bs) {
  p = boff(p, blockIdx.y * bs);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

void launch_fill_zero(void* p, int64_t n, hipStream_t s, Bat bat) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  MPS_LAUNCH(fill_zero_kernel, dim3(g, bat.nb), dim3(256), 0, s, (double2*)p, n, bat.bs);
  HIPCHK(hipGetLastError());
}

// sum |p_i|^2 with a fixed reduction order (one workgroup per copy) -> out[0]
__global__ __launch_bounds__(1024) void sum_abs2_kernel(const double2* p, int64_t n, double* out, int64_t bs) {
  p = boff(p, blockIdx.y * bs);
  out = boff(out, blockIdx.y * bs);
  __shared__ double sh[16];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    const double2 v = p[i];
    s += v.x * v.x + v.y * v.y;
  }
  s = wave_sum_dpp(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < 16; ++w) t += sh[w];
    out[0] = t;
  }
}

void launch_sum_abs2(const void* p, int64_t n, double* out, hipStream_t s, Bat bat) {
  MPS_LAUNCH(sum_abs2_kernel, dim3(1, bat.nb), dim3(1024), 0, s, (const double2*)p, n, out, bat.bs);
  HIPCHK(hipGetLastError());
}

// p <- p / sqrt(norm2[0])   (normalize_lo!, orthonormalisation.jl:97-103)
__global__ void scale_rsqrt_kernel(double2* p, int64_t n, const double* norm2, int64_t bs) {
  p = boff(p, blockIdx.y * bs);
  norm2 = boff(norm2, blockIdx.y * bs);
  const double f = 1.0 / sqrt(norm2[0]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 v = p[i];
    p[i] = make_double2(v.x * f, v.y * f);
  }
}

void launch_scale_rsqrt(void* p, int64_t n, const double* norm2, hipStream_t s, Bat bat) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 4096);
  MPS_LAUNCH(scale_rsqrt_kernel, dim3(g, bat.nb), dim3(256), 0, s, (double2*)p, n, norm2, bat.bs);
  HIPCHK(hipGetLastError());
}

__global__ void copy_c_kernel(double2* d, const double2* s, int64_t n, int64_t bs) {
  d = boff(d, blockIdx.y * bs);
  s = boff(s, blockIdx.y * bs);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

void launch_copy_c(void* dst, const void* src, int64_t n, hipStream_t s, Bat bat) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  MPS_LAUNCH(copy_c_kernel, dim3(g, bat.nb), dim3(256), 0, s, (double2*)dst, (const double2*)src, n,
                     bat.bs);
  HIPCHK(hipGetLastError());
}

__global__ void scale_c_kernel(double2* p, int64_t n, double f) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_double2(f * p[i].x, f * p[i].y);
}

void launch_scale_c(void* p, int64_t n, double f, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  MPS_LAUNCH(scale_c_kernel, dim3(g), dim3(256), 0, s, (double2*)p, n, f);
  HIPCHK(hipGetLastError());
}

__global__ void conj_c_kernel(double2* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i].y = -p[i].y;
}

void launch_conj_c(void* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  MPS_LAUNCH(conj_c_kernel, dim3(g), dim3(256), 0, s, (double2*)p, n);
  HIPCHK(hipGetLastError());
}

// d (cols x rows, column major) = the transpose of s (rows x cols, column
// major), through 32 x 32 LDS tiles so both sides are read and written in
// contiguous column segments
__global__ __launch_bounds__(256) void transpose_c_kernel(double2* __restrict__ d, const double2* __restrict__ s,
                                                          int64_t rows, int64_t cols, int64_t bs) {
  d = boff(d, blockIdx.z * bs);
  s = boff(s, blockIdx.z * bs);
  __shared__ double2 t[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)blockIdx.x * 32, c0 = (int64_t)blockIdx.y * 32;
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + tx, c = c0 + k;
    if (r < rows && c < cols) t[k][tx] = s[r + rows * c];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int64_t c = c0 + tx, r = r0 + k;
    if (r < rows && c < cols) d[c + cols * r] = t[tx][k];
  }
}

void launch_transpose_c(void* dst, const void* src, int64_t rows, int64_t cols, hipStream_t s, Bat bat) {
  if (rows <= 0 || cols <= 0) return;
  REQUIRE((rows + 31) / 32 < (1LL << 31) && (cols + 31) / 32 <= 65535, "transpose too large");
  const dim3 grid((unsigned)((rows + 31) / 32), (unsigned)((cols + 31) / 32), (unsigned)bat.nb);
  MPS_LAUNCH(transpose_c_kernel, grid, dim3(256), 0, s, (double2*)dst, (const double2*)src, rows, cols,
                     bat.bs);
  HIPCHK(hipGetLastError());
}

template <typename TD, typename TS>
__global__ void convert_kernel(TD* d, const TS* s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    d[i].x = s[i].x;
    d[i].y = s[i].y;
  }
}

void launch_convert(void* dst, DType td, const void* src, DType ts, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
  if (td == F32 && ts == F64)
    MPS_LAUNCH((convert_kernel<float2, double2>), dim3(g), dim3(256), 0, s, (float2*)dst,
                       (const double2*)src, n);
  else if (td == F64 && ts == F32)
    MPS_LAUNCH((convert_kernel<double2, float2>), dim3(g), dim3(256), 0, s, (double2*)dst,
                       (const float2*)src, n);
  else if (td == F64)
    MPS_LAUNCH((convert_kernel<double2, double2>), dim3(g), dim3(256), 0, s, (double2*)dst,
                       (const double2*)src, n);
  else
    MPS_LAUNCH((convert_kernel<float2, float2>), dim3(g), dim3(256), 0, s, (float2*)dst,
                       (const float2*)src, n);
  HIPCHK(hipGetLastError());
}

// out = trace(M) for a d x d column-major complex matrix
__global__ void trace_kernel(const double2* M, int64_t d, double2* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double r = 0.0, i = 0.0;
    for (int64_t k = 0; k < d; ++k) {
      r += M[k * d + k].x;
      i += M[k * d + k].y;
    }
    *out = make_double2(r, i);
  }
}

void launch_trace_small(const void* M, int64_t d, void* out, hipStream_t s) {
  MPS_LAUNCH(trace_kernel, dim3(1), dim3(64), 0, s, (const double2*)M, d, (double2*)out);
  HIPCHK(hipGetLastError());
}

}  // namespace mpship
