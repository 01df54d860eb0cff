// Internal declarations shared by the HIP kernels and the host-side sweep code.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <stdexcept>
#include <vector>

namespace mpship {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess)                                                            \
      throw ::mpship::Error(-2, std::string(#x) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define REQUIRE(cond, msg)                                               \
  do {                                                                   \
    if (!(cond)) throw ::mpship::Error(-1, std::string("DimensionMismatch: ") + (msg)); \
  } while (0)

// ------------------------------------------------------- launch checks
// Every kernel launch goes through MPS_LAUNCH, which checks the dispatch
// before it is queued: each grid and block dimension at least 1, the block
// within the kernel's maximum threads (hipFuncGetAttributes, cached per
// kernel), grid x block within 32 bits per dimension (the AQL packet's grid
// size is in work-items), and static plus dynamic LDS within the device's
// per-workgroup limit.  A failed check throws (MPS_E_STATE) instead of
// queueing a malformed packet.
void check_launch(const void* kernel, const char* name, dim3 grid, dim3 block, size_t dyn_lds);
uint64_t launches_checked();

#define MPS_LAUNCH(K, G, B, L, S, ...)                                                          \
  do {                                                                                          \
    ::mpship::check_launch(reinterpret_cast<const void*>(&(K)), #K, dim3(G), dim3(B), (size_t)(L)); \
    hipLaunchKernelGGL(K, G, B, L, S, __VA_ARGS__);                                            \
  } while (0)

// ------------------------------------------------------------- index maps
// A tensor index i (a flattened group of reference indices) maps to the
// element offset (i % n0) * s0 + (i / n0) * s1.  n0 = 0 means a single
// stride (offset i * s0).  Every reshape + permutedims of the reference
// becomes a choice of maps; nothing is materialised.
struct IdxMap {
  int32_t n0;
  int64_t s0, s1;
};
inline IdxMap im1(int64_t s) { return IdxMap{0, s, 0}; }
inline IdxMap im2(int64_t n0, int64_t s0, int64_t s1) { return IdxMap{(int32_t)n0, s0, s1}; }

enum DType { F64 = 0, F32 = 1 };   // storage of complex elements: double2 / float2

// A lockstep batch: nb copies of every buffer of a context, copy b at byte
// offset b * bs from copy 0 (all of a batched context's buffers live in one
// slab with that stride).  Launches through a batched context run every
// kernel for all copies at once (an extra grid dimension), with each copy's
// pointers offset by its stride.  The default is one copy.
struct Bat {
  int nb = 1;
  int64_t bs = 0;
  // copies the GEMM plan (tile size, split-K depth) is made for; 0: nb.  A
  // ragged pool launches a prefix of its copies per site but plans for all of
  // them, so a copy's summation order never depends on the other copies.
  int plan_nb = 0;
};
template <typename T>
__host__ __device__ inline T* boff(T* p, int64_t bytes) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + bytes);
}
template <typename T>
__host__ __device__ inline const T* boff(const T* p, int64_t bytes) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + bytes);
}

// C[m,n] = alpha * sum_k opA(A)[m,k] opB(B)[k,n] + beta * C[m,n], batched.
struct Gemm {
  int64_t M = 0, N = 0, K = 0, batch = 1;
  const void* A = nullptr; IdxMap am{}, ak{}; int64_t a_bs = 0; bool conjA = false; DType ta = F64;
  const void* B = nullptr; IdxMap bk{}, bn{}; int64_t b_bs = 0; bool conjB = false; DType tb = F64;
  void* C = nullptr; IdxMap cm{}, cn{}; int64_t c_bs = 0; DType tc = F64;
  double alpha_re = 1, alpha_im = 0, beta_re = 0, beta_im = 0;
  double diag_re = 0;   // added to C[i][i] after alpha/beta (Q = I - V W)
  Bat bat;              // lockstep copies (A, B, C and the split-K partials all offset)
  int max_splits = 0;   // cap on the split-K depth (0: none; 1: no partials buffer needed)
  // Plan selection: a contraction of the sweeps (tunable) takes its tile and
  // split-K depth from the autotuned table when its shape is there, else from
  // the fixed rule; force_tm > 0 overrides both (the tuner's timing runs).
  bool tunable = false;
  int force_tm = 0, force_splits = 0;
  // cap on the workgroups of the launch (0: none): the output blocks are then
  // walked by fewer workgroups, leaving CUs to concurrent work on other streams
  int max_wg = 0;
};

enum Compute { CMP_F64 = 0, CMP_F32_3M = 1 };

// Workspace for split-K partial sums (provided by the caller, grown on demand).
struct GemmWorkspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
size_t gemm_workspace_bytes(const Gemm& g, Compute cmp, int num_cus);
// the tile and split-K depth launch_gemm uses for g; the lockstep-copy class of the tuned table
void gemm_plan_info(const Gemm& g, Compute cmp, int num_cus, int* tm, int* splits);
int nb_class(int64_t nb);
void launch_gemm(const Gemm& g, Compute cmp, void* ws, size_t ws_bytes, int num_cus, hipStream_t s);

// Householder QR (zgeqrf + zungqr semantics, optional rank cutoff); in place Q.
// Element (r, c) at base + r*rs + c*cs; m >= n.  ws: qr_workspace_bytes(m, n,
// num_cus).  R_out (optional): n x n column-major upper triangle.
size_t qr_workspace_bytes(int64_t m, int64_t n, int num_cus);
// chain: 1 single-launch factorization where it applies (m <= 2560, n <= 256,
// one copy), 0 launch per panel, -1 the process default (MPS_QR_CHAIN)
void launch_qr(void* base, int64_t m, int64_t n, int64_t rs, int64_t cs, double tol, void* ws, int num_cus,
               void* R_out, hipStream_t s, Bat bat = Bat(), int chain = -1);
// once per device, outside graph capture: checks the workgroup-to-XCD layout
// the single-launch (chain) QR relies on; launch_qr uses that path only after it passed
void qr_chain_probe(int device);
// whether launch_qr factors an m x n panel in one launch (chain as for launch_qr)
bool qr_single_launch_applies(int64_t m, int64_t n, const Bat& bat, int chain);

// ------------------------------------------------------- device helpers
// DPP move of a double (two 32-bit halves), bound_ctrl: out-of-row sources read 0
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
// Sum over the 64 lanes of a wave, result in every lane: quad_perm and row_ror
// DPP steps inside each 16-lane row, two cross-row exchanges.
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v += dpp_d<0xB1>(v);    // quad_perm(1,0,3,2)
  v += dpp_d<0x4E>(v);    // quad_perm(2,3,0,1)
  v += dpp_d<0x124>(v);   // row_ror:4
  v += dpp_d<0x128>(v);   // row_ror:8
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// small kernels (misc.hip)
void launch_set_scalar(void* p, double re, double im, hipStream_t s, Bat bat = Bat());
void launch_fill_zero(void* p, int64_t n_complex, hipStream_t s, Bat bat = Bat());
void launch_sum_abs2(const void* p, int64_t n, double* out, hipStream_t s, Bat bat = Bat());
void launch_scale_rsqrt(void* p, int64_t n, const double* norm2, hipStream_t s, Bat bat = Bat());
void launch_copy_c(void* dst, const void* src, int64_t n, hipStream_t s, Bat bat = Bat());
void launch_conj_c(void* p, int64_t n, hipStream_t s);   // p <- conj(p), in place
void launch_scale_c(void* p, int64_t n, double f, hipStream_t s);   // p <- f p, in place
// dst (cols x rows) = transpose of src (rows x cols), both column major, complex FP64
void launch_transpose_c(void* dst, const void* src, int64_t rows, int64_t cols, hipStream_t s, Bat bat = Bat());
void launch_convert(void* dst, DType td, const void* src, DType ts, int64_t n, hipStream_t s);
void launch_trace_small(const void* M, int64_t d, void* out, hipStream_t s);
// one-sided Jacobi SVD (svd.hip)
void svd_load(void* Y, const void* X, int64_t m, int64_t n, bool conj_transpose, hipStream_t s);
int svd_jacobi(void* Y, int64_t m, int64_t n, void* J, double tol, int* flag_dev, int* flag_host,
               int max_sweeps, hipStream_t s);
void svd_colnorms(const void* Y, int64_t m, int64_t n, double* sigma, hipStream_t s);
void svd_gather(void* out, const void* S, int64_t ld, const int* perm, const double* scale, int64_t rows,
                int64_t cols, bool conj, bool by_col, hipStream_t s);

}  // namespace mpship
