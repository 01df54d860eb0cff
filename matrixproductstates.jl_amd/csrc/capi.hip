// extern "C" boundary (include/mpship.h): handles, transfers, error codes,
// hipGraph caching of RK4 steps.
#include "../../include/mpship.h"
#include "contract.h"
#include "sweeps.h"
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <mutex>

using namespace mpship;
using cd = std::complex<double>;

struct mps_ctx_s {
  Ctx c;
  // cached RK4 graphs, keyed by the handles and parameters baked into them
  struct Key {
    std::vector<const void*> h;
    double dt;
    int prec;
    double tol;
    size_t ws_gen;
    int refine;
    bool operator<(const Key& o) const {
      if (h != o.h) return h < o.h;
      if (dt != o.dt) return dt < o.dt;
      if (prec != o.prec) return prec < o.prec;
      if (tol != o.tol) return tol < o.tol;
      if (refine != o.refine) return refine < o.refine;
      return ws_gen < o.ws_gen;
    }
  };
  std::map<Key, hipGraphExec_t> graphs;
  std::map<Key, int> seen;
  size_t ws_gen = 0;
  // second operator environment for the RK4 stages whose environments are
  // built on the aux stream (same slots as the caller's envop)
  std::unique_ptr<Env> shadow;
  // RK4 refinement states (mps_ctx_set_refine_sweeps): psi's copy for the
  // 10-argument form and the refit's intermediate state
  std::unique_ptr<State> rkeep, rtmp;
  // the next call's stage-1 environments, queued by the last RK4 call: valid
  // for a call with these handles and precision if no writing entry point
  // ran since (g_epoch == spec_epoch + 1 inside that call)
  bool spec = false;
  uint64_t spec_epoch = 0;
  const void *spec_H1 = nullptr, *spec_psi = nullptr, *spec_envop = nullptr;
  int spec_prec = 0;
  void clear_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
    seen.clear();
  }
  ~mps_ctx_s() { clear_graphs(); }
};
struct mps_state_s {
  State s;
  mps_ctx_s* owner;
};
struct mps_mpo_s {
  MPO o;
  mps_ctx_s* owner;
};
struct mps_env_s {
  Env e;
  mps_ctx_s* owner;
};

static thread_local std::string g_err;

// Every entry point that may write a state, MPO or environment advances
// this counter (read-only entry points use guard_ro); an RK4 call checks it
// to know that nothing changed since the previous call queued the next
// call's first environments.
static std::atomic<uint64_t> g_epoch{0};

template <typename F>
static int guard_ro(F&& f) {
  try {
    f();
    return MPS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MPS_E_STATE;
  }
}

template <typename F>
static int guard(F&& f) {
  g_epoch.fetch_add(1);
  try {
    f();
    return MPS_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return MPS_E_STATE;
  }
}

static void set_device(Ctx* c) { HIPCHK(hipSetDevice(c->device)); }

// Entry points that produce per-call host values or touch single buffers run
// on plain contexts only; a batched context's copies go through views.
static void no_batch(const Ctx* c) {
  if (c->bat.nb != 1)
    throw Error(MPS_E_STATE, "not available on a batched context (use a view of one copy)");
}

// zero `bytes` at p in every lockstep copy of the context
static void zero_copies(Ctx* c, void* p, size_t bytes) {
  if (c->bat.nb == 1)
    HIPCHK(hipMemsetAsync(p, 0, bytes, c->stream));
  else
    HIPCHK(hipMemset2DAsync(p, (size_t)c->bat.bs, 0, bytes, (size_t)c->bat.nb, c->stream));
}

// the same host data into every lockstep copy
static void upload_copies(Ctx* c, void* dst, const void* host, size_t bytes) {
  for (int b = 0; b < c->bat.nb; ++b)
    HIPCHK(hipMemcpyAsync(boff((char*)dst, (int64_t)b * c->bat.bs), host, bytes, hipMemcpyHostToDevice, c->stream));
}

// Any change of device buffers or dims invalidates captured graphs.
static void invalidate(mps_ctx_s* x) {
  if (x) x->clear_graphs();
}

static void track_ws(mps_ctx_s* x, size_t before) {
  if (x->c.ws_bytes != before) {
    x->ws_gen++;
    x->clear_graphs();
  }
}

// ----------------------------------------------------- the no-alloc buffer
// mps_scal_prod_no_conj_noalloc works in this fixed device array instead of
// the context's workspace: the two running environments, the per-site
// product and the result of one call.  Every GPU the library runs on holds
// its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for it with the device current
// (scratch_on(device)).  Calls through the entry point on one GPU are
// serialised by a per-device lock held from the first launch to the final
// synchronisation, so they cannot overlap on the buffer.
static constexpr size_t kNoAllocScratchBytes = (size_t)24 << 20;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_init_mu;
static std::mutex g_scratch_call_mu[kMaxDevices];
static char* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on `device`, resolved once per device with
// that device current (called at context creation, the library's set-up)
static char* scratch_on(int device) {
  REQUIRE(device >= 0 && device < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_init_mu);
  if (g_scratch_ptr[device] == nullptr) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(cur);
    if (e != hipSuccess || ptr == nullptr)
      throw Error(MPS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (char*)ptr;
  }
  return g_scratch_ptr[device];
}

// A context-owned state shaped like `like` (its bonds at most the capacity);
// reallocated, and the context's graphs dropped, when it does not fit.
static State* internal_like(mps_ctx_s* x, std::unique_ptr<State>& s, const State& like) {
  bool fits = s && s->N == like.N && s->d == like.d;
  for (int j = 0; fits && j < like.N; ++j) fits = like.dims[j] <= s->cap_l[j] && like.dims[j + 1] <= s->cap_r[j];
  if (!fits) {
    x->clear_graphs();
    s.reset();
    auto n = std::make_unique<State>();
    n->ctx = &x->c;
    n->N = like.N;
    n->d = like.d;
    n->dims = like.dims;
    n->cap_l.assign(like.dims.begin(), like.dims.end() - 1);
    n->cap_r.assign(like.dims.begin() + 1, like.dims.end());
    n->off.resize(like.N);
    size_t tot = 0;
    for (int j = 0; j < like.N; ++j) {
      n->off[j] = (int64_t)tot;
      tot += (size_t)like.dims[j] * like.d[j] * like.dims[j + 1];
    }
    n->total = tot;
    n->buf = (double2*)x->c.alloc(tot * sizeof(double2), n->own);
    s = std::move(n);
  }
  s->dims = like.dims;
  return s.get();
}

extern "C" {

const char* mps_last_error(void) { return g_err.c_str(); }
int mps_version(void) { return 1; }

int mps_ctx_create(int device, mps_ctx_t* out) {
  return guard([&] {
    REQUIRE(out != nullptr, "null output");
    auto* x = new mps_ctx_s();
    try {
      x->c.device = device;
      HIPCHK(hipSetDevice(device));
      (void)scratch_on(device);
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, device));
      x->c.num_cus = prop.multiProcessorCount;
      qr_chain_probe(device);
      HIPCHK(hipStreamCreateWithFlags(&x->c.stream, hipStreamNonBlocking));
      x->c.own_stream = true;
      HIPCHK(hipMalloc((void**)&x->c.dscal, 256 * sizeof(double)));
      HIPCHK(hipHostMalloc((void**)&x->c.hscal, 256 * sizeof(double), hipHostMallocDefault));
    } catch (...) {
      delete x;
      throw;
    }
    *out = x;
  });
}

int mps_ctx_create_batch(int device, int nb, int64_t copy_bytes, int64_t ws_bytes, mps_ctx_t* out) {
  return guard([&] {
    REQUIRE(out != nullptr && nb >= 1 && nb <= 65535, "invalid arguments");
    REQUIRE(copy_bytes > 0 && ws_bytes >= 0, "invalid arguments");
    auto* x = new mps_ctx_s();
    try {
      Ctx& c = x->c;
      c.device = device;
      HIPCHK(hipSetDevice(device));
      hipDeviceProp_t prop;
      HIPCHK(hipGetDeviceProperties(&prop, device));
      c.num_cus = prop.multiProcessorCount;
      qr_chain_probe(device);
      HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
      c.own_stream = true;
      // copy stride: scalars + workspace + the caller's containers, 256-byte aligned
      const size_t ws = ((size_t)ws_bytes + 255) & ~(size_t)255;
      const size_t stride = ((size_t)copy_bytes + ws + 4096 + 255) & ~(size_t)255;
      c.bat.nb = nb;
      c.bat.bs = (int64_t)stride;
      if (hipMalloc((void**)&c.slab, stride * (size_t)nb) != hipSuccess) {
        c.slab = nullptr;
        throw Error(MPS_E_ALLOC, "batched context: slab allocation failed");
      }
      bool own = false;
      c.dscal = (double*)c.alloc(256 * sizeof(double), own);
      c.ws = (char*)c.alloc(ws, own);
      c.ws_bytes = ws;
      HIPCHK(hipHostMalloc((void**)&c.hscal, std::max(256, nb) * sizeof(double), hipHostMallocDefault));
    } catch (...) {
      delete x;
      throw;
    }
    *out = x;
  });
}

int mps_batch_workspace_bytes(int device, int nb, int64_t Dmax, int64_t d, int64_t wmax, int64_t* out) {
  return guard([&] {
    REQUIRE(out && nb >= 1 && Dmax >= 1 && d >= 1 && wmax >= 1, "invalid arguments");
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    // the largest plan of the batched calls: the RK4 step's (four overlap products)
    *out = (int64_t)plan_bytes(Dmax, d, wmax, 4, 0, prop.multiProcessorCount, nb);
  });
}

int mps_ctx_create_companion(mps_ctx_t batched, mps_ctx_t* out) {
  return guard([&] {
    REQUIRE(batched != nullptr && out != nullptr, "invalid arguments");
    auto* x = new mps_ctx_s();
    try {
      Ctx& c = x->c;
      c.device = batched->c.device;
      HIPCHK(hipSetDevice(c.device));
      c.num_cus = batched->c.num_cus;
      c.stream = batched->c.stream;   // same stream: calls on views order with the batched calls
      c.own_stream = false;
      c.qr_tol = batched->c.qr_tol;
      HIPCHK(hipMalloc((void**)&c.dscal, 256 * sizeof(double)));
      HIPCHK(hipHostMalloc((void**)&c.hscal, 256 * sizeof(double), hipHostMallocDefault));
    } catch (...) {
      delete x;
      throw;
    }
    *out = x;
  });
}

int mps_ctx_set_ragged(mps_ctx_t ctx, int nmax, const int* n) {
  return guard([&] {
    REQUIRE(ctx && n, "invalid arguments");
    Ctx& c = ctx->c;
    const int nb = c.bat.nb;
    REQUIRE(nb > 1, "a ragged pool is a batched context");
    REQUIRE(!c.rag, "the pool's chain lengths are already set");
    REQUIRE(n[0] == nmax && nmax >= 2, "copy 0 holds the longest chain, of nmax sites");
    auto r = std::make_shared<Ctx::Ragged>();
    r->nmax = nmax;
    r->n.assign(n, n + nb);
    for (int b = 0; b < nb; ++b) {
      REQUIRE(n[b] >= 2 && n[b] <= nmax, "chain lengths in [2, nmax]");
      REQUIRE(b == 0 || n[b] <= n[b - 1], "copies sorted by chain length, longest first");
      r->h.push_back(n[b] / 2);
    }
    // copy b is present at slot q if q lies in its left part or its right part
    r->kact.assign(nmax, 0);
    for (int q = 0; q < nmax; ++q) {
      int k = 0;
      for (int b = 0; b < nb; ++b) {
        const bool on = q < r->h[b] || q >= nmax - (n[b] - r->h[b]);
        if (on) {
          REQUIRE(k == b, "ragged pool: copies present at a slot are not a prefix");
          ++k;
        }
      }
      r->kact[q] = k;
    }
    c.rag = r;
    c.bat.plan_nb = nb;
    invalidate(ctx);
  });
}

int mps_ctx_destroy(mps_ctx_t ctx) {
  return guard([&] {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
  });
}

int mps_ctx_set_stream(mps_ctx_t ctx, void* stream) {
  return guard([&] {
    set_device(&ctx->c);
    HIPCHK(hipStreamSynchronize(ctx->c.stream));
    if (ctx->c.own_stream) HIPCHK(hipStreamDestroy(ctx->c.stream));
    if (stream) {
      ctx->c.stream = (hipStream_t)stream;
      ctx->c.own_stream = false;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
      ctx->c.own_stream = true;
    }
    invalidate(ctx);
  });
}

int mps_ctx_synchronize(mps_ctx_t ctx) { return guard_ro([&] { ctx->c.sync(); }); }

int mps_ctx_set_qr_rank_tol(mps_ctx_t ctx, double tol) {
  return guard([&] {
    REQUIRE(tol >= 0.0 && tol < 1.0, "qr rank tolerance must be in [0, 1)");
    ctx->c.qr_tol = tol;
  });
}

int mps_ctx_set_qr_single_launch(mps_ctx_t ctx, int on) {
  return guard([&] {
    REQUIRE(on >= -1 && on <= 1, "qr form: 1 single launch, 0 launch per panel, -1 process default");
    ctx->c.qr_chain = on;
  });
}

int mps_ctx_set_precision(mps_ctx_t ctx, int prec) {
  return guard([&] {
    REQUIRE(prec == MPS_PREC_FP64 || prec == MPS_PREC_FP32, "unknown precision");
    ctx->c.prec = prec;
  });
}

int mps_ctx_set_refine_sweeps(mps_ctx_t ctx, int n) {
  return guard([&] {
    REQUIRE(n >= 0 && n <= 64 && n % 2 == 0, "refinement half-sweeps: an even number in [0, 64]");
    ctx->c.refine = n;
    invalidate(ctx);
  });
}

int mps_ctx_set_graphs(mps_ctx_t ctx, int enable) {
  return guard([&] {
    ctx->c.graphs = enable != 0;
    if (!enable) invalidate(ctx);
  });
}

int mps_ctx_set_profiling(mps_ctx_t ctx, int enable) {
  return guard([&] {
    ctx->c.sync();
    ctx->c.profiling = enable != 0;
    ctx->c.prof.clear();
    ctx->c.ev_used = 0;
  });
}

// ------------------------------------------------------------ GEMM tuning
// Distinct shapes of the recorded contractions, in order of first launch:
// key (M, N, K, batch, copy class, arithmetic).
static std::vector<size_t> distinct_recorded(const Ctx& c) {
  std::vector<size_t> first;
  std::vector<std::array<int64_t, 6>> keys;
  for (size_t i = 0; i < c.rec.size(); ++i) {
    const Gemm& g = c.rec[i].g;
    const int64_t nbq = g.bat.plan_nb > 0 ? g.bat.plan_nb : g.bat.nb;
    const std::array<int64_t, 6> k{g.M, g.N, g.K, g.batch, (int64_t)nb_class(nbq), (int64_t)c.rec[i].cmp};
    if (std::find(keys.begin(), keys.end(), k) == keys.end()) {
      keys.push_back(k);
      first.push_back(i);
    }
  }
  return first;
}

int mps_ctx_gemm_record(mps_ctx_t ctx, int enable) {
  return guard_ro([&] {
    REQUIRE(ctx, "invalid arguments");
    Ctx& c = ctx->c;
    if (enable) {
      c.rec.clear();
      c.gemm_rec = true;
      ctx->clear_graphs();   // recorded calls launch eagerly
    } else {
      c.gemm_rec = false;
    }
  });
}

int mps_ctx_gemm_recorded(mps_ctx_t ctx, int64_t* out, int max_rows, int* rows) {
  return guard_ro([&] {
    REQUIRE(ctx && rows, "invalid arguments");
    const Ctx& c = ctx->c;
    const std::vector<size_t> first = distinct_recorded(c);
    *rows = (int)first.size();
    for (int r = 0; r < (int)first.size() && r < max_rows && out; ++r) {
      const Ctx::RecGemm& x = c.rec[first[r]];
      const int64_t nbq = x.g.bat.plan_nb > 0 ? x.g.bat.plan_nb : x.g.bat.nb;
      int tm = 0, sp = 0;
      gemm_plan_info(x.g, x.cmp, c.num_cus, &tm, &sp);
      int64_t* o = out + 10 * r;
      o[0] = x.g.M; o[1] = x.g.N; o[2] = x.g.K; o[3] = x.g.batch; o[4] = nb_class(nbq); o[5] = (int64_t)x.cmp;
      o[6] = tm; o[7] = sp; o[8] = nbq; o[9] = (int64_t)x.ws_bytes;
    }
  });
}

int mps_ctx_gemm_time(mps_ctx_t ctx, int row, int tm, int splits, int iters, double* us) {
  return guard([&] {
    REQUIRE(ctx && us && iters >= 1, "invalid arguments");
    REQUIRE(tm == 32 || tm == 64, "tile 32 or 64");
    REQUIRE(splits >= 1 && splits <= 64, "split-K depth in [1, 64]");
    Ctx& c = ctx->c;
    REQUIRE(!c.capturing && !c.gemm_rec, "stop recording before timing");
    const std::vector<size_t> first = distinct_recorded(c);
    REQUIRE(row >= 0 && row < (int)first.size(), "no such recorded GEMM");
    set_device(&c);
    Ctx::RecGemm x = c.rec[first[row]];
    x.g.force_tm = tm;
    x.g.force_splits = splits;
    // refused before anything is queued when the partials do not fit the workspace
    REQUIRE(gemm_workspace_bytes(x.g, x.cmp, c.num_cus) <= (x.ws ? x.ws_bytes : 0),
            "split-K partials exceed the call's workspace");
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    launch_gemm(x.g, x.cmp, x.ws, x.ws_bytes, c.num_cus, x.s);   // warm
    HIPCHK(hipEventRecord(a, x.s));
    for (int i = 0; i < iters; ++i) launch_gemm(x.g, x.cmp, x.ws, x.ws_bytes, c.num_cus, x.s);
    HIPCHK(hipEventRecord(b, x.s));
    HIPCHK(hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *us = 1e3 * (double)ms / iters;
  });
}

int mps_ctx_kernel_stats(mps_ctx_t ctx, double* out) {
  return guard([&] {
    Ctx& c = ctx->c;
    c.sync();
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (auto& r : c.prof) {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
      if (r.kind == 0) {
        v[0] += ms;
        v[1] += r.flops;
        v[2] += r.bytes;
        v[3] += 1;
      } else {
        v[4] += ms;
        v[5] += r.flops;
        v[6] += 1;
      }
    }
    v[7] = (double)launches_checked();
    std::memcpy(out, v, sizeof(v));
    c.prof.clear();
    c.ev_used = 0;
  });
}

// ------------------------------------------------------------------ states
int mps_state_create(mps_ctx_t ctx, int N, const int64_t* d, const int64_t* dims, mps_state_t* out) {
  return guard([&] {
    REQUIRE(ctx && out && d && dims && N >= 1, "invalid arguments");
    set_device(&ctx->c);
    auto* x = new mps_state_s();
    x->owner = ctx;
    State& s = x->s;
    s.ctx = &ctx->c;
    s.N = N;
    s.d.assign(d, d + N);
    s.dims.assign(dims, dims + N + 1);
    s.cap_l.resize(N);
    s.cap_r.resize(N);
    s.off.resize(N);
    size_t tot = 0;
    for (int j = 0; j < N; ++j) {
      if (!(s.d[j] >= 1 && s.dims[j] >= 1 && s.dims[j + 1] >= 1)) {
        delete x;
        throw Error(MPS_E_ARG, "DimensionMismatch: dims must be >= 1");
      }
      s.cap_l[j] = s.dims[j];
      s.cap_r[j] = s.dims[j + 1];
      s.off[j] = (int64_t)tot;
      tot += (size_t)s.dims[j] * s.d[j] * s.dims[j + 1];
    }
    s.total = tot;
    try {
      s.buf = (double2*)ctx->c.alloc(tot * sizeof(double2), s.own);
    } catch (...) {
      s.buf = nullptr;
      delete x;
      throw Error(MPS_E_ALLOC, "state allocation failed");
    }
    zero_copies(&ctx->c, s.buf, tot * sizeof(double2));
    ctx->c.sync();
    *out = x;
  });
}

int mps_state_destroy(mps_state_t s) {
  return guard([&] {
    if (!s) return;
    s->owner->c.sync();
    invalidate(s->owner);
    delete s;
  });
}

int mps_state_dims(mps_state_t s, int64_t* out) {
  return guard_ro([&] { std::memcpy(out, s->s.dims.data(), (s->s.N + 1) * sizeof(int64_t)); });
}

int mps_state_upload_site(mps_state_t s, int site, const double* host, int64_t dl, int64_t d, int64_t dr) {
  return guard([&] {
    State& st = s->s;
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    REQUIRE(dl == st.Dl(site) && d == st.d[site] && dr == st.Dr(site), "site shape mismatch");
    set_device(st.ctx);
    upload_copies(st.ctx, st.site(site), host, dl * d * dr * sizeof(double2));
    st.ctx->sync();
  });
}

int mps_state_download_site(mps_state_t s, int site, double* host, int64_t dl, int64_t d, int64_t dr) {
  return guard_ro([&] {
    State& st = s->s;
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    REQUIRE(dl == st.Dl(site) && d == st.d[site] && dr == st.Dr(site), "site shape mismatch");
    set_device(st.ctx);
    HIPCHK(hipMemcpyAsync(host, st.site(site), dl * d * dr * sizeof(double2), hipMemcpyDeviceToHost,
                          st.ctx->stream));
    st.ctx->sync();
  });
}

// ComplexF32 host buffers (Julia Array{ComplexF32}, the TN = Complex{Float32}
// of the reference's GPU examples): staged in the context's workspace and
// widened / narrowed on the device; the library stores and computes complex
// FP64 (or FP32 in the precision form's operator terms).
static void* f32_stage(Ctx* c, int64_t n) {
  REQUIRE(c->bat.nb == 1, "ComplexF32 transfers run on plain contexts (use a view of one copy)");
  c->ensure_ws((size_t)n * sizeof(float2) + 256);
  return c->ws;
}

int mps_state_upload_site_f32(mps_state_t s, int site, const float* host, int64_t dl, int64_t d, int64_t dr) {
  return guard([&] {
    State& st = s->s;
    REQUIRE(host != nullptr, "null host buffer");
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    REQUIRE(dl == st.Dl(site) && d == st.d[site] && dr == st.Dr(site), "site shape mismatch");
    set_device(st.ctx);
    const int64_t n = dl * d * dr;
    void* stage = f32_stage(st.ctx, n);
    HIPCHK(hipMemcpyAsync(stage, host, n * sizeof(float2), hipMemcpyHostToDevice, st.ctx->stream));
    launch_convert(st.site(site), F64, stage, F32, n, st.ctx->stream);
    st.ctx->sync();
  });
}

int mps_state_download_site_f32(mps_state_t s, int site, float* host, int64_t dl, int64_t d, int64_t dr) {
  return guard_ro([&] {
    State& st = s->s;
    REQUIRE(host != nullptr, "null host buffer");
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    REQUIRE(dl == st.Dl(site) && d == st.d[site] && dr == st.Dr(site), "site shape mismatch");
    set_device(st.ctx);
    const int64_t n = dl * d * dr;
    void* stage = f32_stage(st.ctx, n);
    launch_convert(stage, F32, st.site(site), F64, n, st.ctx->stream);
    HIPCHK(hipMemcpyAsync(host, stage, n * sizeof(float2), hipMemcpyDeviceToHost, st.ctx->stream));
    st.ctx->sync();
  });
}

int mps_state_copy(mps_state_t dst, mps_state_t src) {
  return guard([&] {
    State &a = dst->s, &b = src->s;
    REQUIRE(a.N == b.N, "chain lengths differ");
    for (int j = 0; j < a.N; ++j) {
      REQUIRE(a.d[j] == b.d[j], "local dims differ");
      REQUIRE(b.dims[j] <= a.cap_l[j] && b.dims[j + 1] <= a.cap_r[j], "destination capacity too small");
    }
    set_device(a.ctx);
    REQUIRE(a.ctx->bat.nb == b.ctx->bat.nb && a.ctx->bat.bs == b.ctx->bat.bs,
            "state copy between a batched and a plain container");
    // the source may still be written by work queued on another context's stream
    if (b.ctx != a.ctx) b.ctx->sync();
    bool dims_changed = a.dims != b.dims;
    a.dims = b.dims;
    for (int j = 0; j < a.N; ++j) launch_copy_c(a.site(j), b.site(j), b.size(j), a.ctx->stream, a.ctx->bat);
    if (dims_changed) invalidate(dst->owner);
  });
}

int mps_state_product(mps_state_t s, const double* states) {
  return guard([&] {
    State& st = s->s;
    if (st.ctx->rag) throw Error(MPS_E_STATE, "not available on a ragged pool (set each copy through a view)");
    set_device(st.ctx);
    // raw zero-padded sites with [0, :, 0] = state (constructors.jl:19-20), then
    // the left QR sweep that mpsproductstate interleaves with their construction
    for (int j = 0; j < st.N; ++j) {
      std::vector<double2> h((size_t)st.size(j), make_double2(0.0, 0.0));
      const int64_t Dl = st.Dl(j);
      for (int64_t a = 0; a < st.d[j]; ++a) {
        int64_t off = 0;
        for (int k = 0; k < j; ++k) off += st.d[k];
        h[(size_t)(a * Dl)] = make_double2(states[2 * (off + a)], states[2 * (off + a) + 1]);
      }
      upload_copies(st.ctx, st.site(j), h.data(), h.size() * sizeof(double2));
      st.ctx->sync();
    }
    size_t before = st.ctx->ws_bytes;
    Plan p(st.ctx, {&st}, {}, 0);
    track_ws(s->owner, before);
    left_orthogonalize(p, &st, 0);
    st.ctx->sync();
  });
}

// -------------------------------------------------------------------- MPOs
int mps_mpo_create(mps_ctx_t ctx, int N, const int64_t* d, const int64_t* wdims, mps_mpo_t* out) {
  return guard([&] {
    REQUIRE(ctx && out && d && wdims && N >= 1, "invalid arguments");
    set_device(&ctx->c);
    auto* x = new mps_mpo_s();
    x->owner = ctx;
    MPO& o = x->o;
    o.ctx = &ctx->c;
    o.N = N;
    o.d.assign(d, d + N);
    o.w.assign(wdims, wdims + N + 1);
    o.off.resize(N);
    size_t tot = 0;
    for (int j = 0; j < N; ++j) {
      o.off[j] = (int64_t)tot;
      tot += (size_t)o.w[j] * o.d[j] * o.w[j + 1] * o.d[j];
    }
    o.total = tot;
    try {
      o.buf = (double2*)ctx->c.alloc(tot * sizeof(double2), o.own);
    } catch (...) {
      o.buf = nullptr;
      delete x;
      throw Error(MPS_E_ALLOC, "MPO allocation failed");
    }
    zero_copies(&ctx->c, o.buf, tot * sizeof(double2));
    ctx->c.sync();
    *out = x;
  });
}

int mps_mpo_destroy(mps_mpo_t o) {
  return guard([&] {
    if (!o) return;
    o->owner->c.sync();
    invalidate(o->owner);
    delete o;
  });
}

int mps_mpo_upload_site(mps_mpo_t o, int site, const double* host, int64_t wl, int64_t d, int64_t wr) {
  return guard([&] {
    MPO& m = o->o;
    REQUIRE(site >= 0 && site < m.N, "site out of range");
    REQUIRE(wl == m.wl(site) && d == m.d[site] && wr == m.wr(site), "MPO site shape mismatch");
    set_device(m.ctx);
    upload_copies(m.ctx, m.site(site), host, wl * d * wr * d * sizeof(double2));
    m.ctx->sync();
  });
}

int mps_mpo_upload_site_f32(mps_mpo_t o, int site, const float* host, int64_t wl, int64_t d, int64_t wr) {
  return guard([&] {
    MPO& m = o->o;
    REQUIRE(host != nullptr, "null host buffer");
    REQUIRE(site >= 0 && site < m.N, "site out of range");
    REQUIRE(wl == m.wl(site) && d == m.d[site] && wr == m.wr(site), "MPO site shape mismatch");
    set_device(m.ctx);
    const int64_t n = wl * d * wr * d;
    void* stage = f32_stage(m.ctx, n);
    HIPCHK(hipMemcpyAsync(stage, host, n * sizeof(float2), hipMemcpyHostToDevice, m.ctx->stream));
    launch_convert(m.site(site), F64, stage, F32, n, m.ctx->stream);
    m.ctx->sync();
  });
}

int mps_mpo_set_block(mps_mpo_t o, int site, int64_t wl_idx, int64_t wr_idx, const double* op) {
  return guard([&] {
    MPO& m = o->o;
    REQUIRE(site >= 0 && site < m.N, "site out of range");
    const int64_t wl = m.wl(site), wr = m.wr(site), d = m.d[site];
    REQUIRE(wl_idx >= 0 && wl_idx < wl && wr_idx >= 0 && wr_idx < wr, "MPO block index out of range");
    set_device(m.ctx);
    // O[wl_idx, s, wr_idx, t] at wl_idx + wl*s + wl*d*wr_idx + wl*d*wr*t
    for (int b = 0; b < m.ctx->bat.nb; ++b)
      for (int64_t t = 0; t < d; ++t)
        HIPCHK(hipMemcpy2DAsync(boff(m.site(site), (int64_t)b * m.ctx->bat.bs) + wl_idx + wl * d * wr_idx +
                                    wl * d * wr * t,
                                wl * sizeof(double2), op + 2 * d * t, sizeof(double2), sizeof(double2), d,
                                hipMemcpyHostToDevice, m.ctx->stream));
    m.ctx->sync();
  });
}

// ------------------------------------------------------------ environments
static int env_create(mps_ctx_t ctx, int N, const int64_t* dims1, const int64_t* dims2, const int64_t* w,
                      bool op, mps_env_t* out) {
  return guard([&] {
    REQUIRE(ctx && out && dims1 && dims2 && N >= 1, "invalid arguments");
    set_device(&ctx->c);
    auto* x = new mps_env_s();
    x->owner = ctx;
    Env& e = x->e;
    e.ctx = &ctx->c;
    e.N = N;
    e.op = op;
    e.d1.resize(N + 1);
    e.wd.resize(N + 1);
    e.d2.resize(N + 1);
    e.off.resize(N + 1);
    size_t tot = 0;
    for (int j = 0; j <= N; ++j) {
      const bool edge = (j == 0 || j == N);
      e.d1[j] = edge ? 1 : dims1[j];
      e.d2[j] = edge ? 1 : dims2[j];
      e.wd[j] = (edge || !op) ? 1 : w[j];
      e.off[j] = (int64_t)tot;
      // slots hold left or right environments of either bond orientation
      const int64_t big = std::max(e.d1[j], e.d2[j]);
      tot += (size_t)big * e.wd[j] * big;
    }
    e.total = tot;
    try {
      e.buf = (double2*)ctx->c.alloc(tot * sizeof(double2), e.own);
    } catch (...) {
      e.buf = nullptr;
      delete x;
      throw Error(MPS_E_ALLOC, "environment allocation failed");
    }
    zero_copies(&ctx->c, e.buf, tot * sizeof(double2));
    // build_env: ones at both boundaries
    launch_set_scalar(e.slot(0), 1.0, 0.0, ctx->c.stream, ctx->c.bat);
    launch_set_scalar(e.slot(N), 1.0, 0.0, ctx->c.stream, ctx->c.bat);
    ctx->c.sync();
    *out = x;
  });
}

int mps_env_create_op(mps_ctx_t ctx, int N, const int64_t* dims1, const int64_t* dims2,
                      const int64_t* wdims, mps_env_t* out) {
  if (!wdims) {
    g_err = "null wdims";
    return MPS_E_ARG;
  }
  return env_create(ctx, N, dims1, dims2, wdims, true, out);
}

int mps_env_create(mps_ctx_t ctx, int N, const int64_t* dims1, const int64_t* dims2, mps_env_t* out) {
  return env_create(ctx, N, dims1, dims2, nullptr, false, out);
}

int mps_env_destroy(mps_env_t e) {
  return guard([&] {
    if (!e) return;
    e->owner->c.sync();
    invalidate(e->owner);
    delete e;
  });
}

int mps_env_download_slot(mps_env_t e, int slot, double* host, int64_t n) {
  return guard_ro([&] {
    Env& v = e->e;
    REQUIRE(slot >= 0 && slot <= v.N, "slot out of range");
    const int64_t cap = (slot < v.N ? v.off[slot + 1] : (int64_t)v.total) - v.off[slot];
    REQUIRE(n >= 0 && n <= cap, "slot size mismatch");
    set_device(v.ctx);
    HIPCHK(hipMemcpyAsync(host, v.slot(slot), n * sizeof(double2), hipMemcpyDeviceToHost, v.ctx->stream));
    v.ctx->sync();
  });
}

// ------------------------------------------------- raw L1 contractions
#define RAW_PLAN(ctx, Dm, dm, wm)            \
  set_device(&ctx->c);                       \
  no_batch(&ctx->c);                         \
  size_t before_ = ctx->c.ws_bytes;          \
  Plan p(&ctx->c, (Dm), (dm), (wm), 0);      \
  track_ws(ctx, before_);

static i64 mx(std::initializer_list<i64> v) {
  i64 m = 1;
  for (i64 x : v) m = std::max(m, x);
  return m;
}

int mps_dev_prod_AOR(mps_ctx_t ctx, void* out, const void* A, int64_t Dl, int64_t d, int64_t Dr,
                     const void* O, int64_t wl, int64_t wr, const void* R, int64_t Dr0) {
  return guard([&] {
    RAW_PLAN(ctx, mx({Dl, Dr, Dr0}), d, mx({wl, wr}));
    prod_AOR(p.work, p.op, out, A, Dl, d, Dr, O, wl, wr, R, Dr0);
    ctx->c.sync();
  });
}

int mps_dev_prod_AOL(mps_ctx_t ctx, void* out, const void* A, int64_t Dl, int64_t d, int64_t Dr,
                     const void* O, int64_t wl, int64_t wr, const void* L, int64_t Dl0) {
  return guard([&] {
    RAW_PLAN(ctx, mx({Dl, Dr, Dl0}), d, mx({wl, wr}));
    prod_AOL(p.work, p.op, out, A, Dl, d, Dr, O, wl, wr, L, Dl0);
    ctx->c.sync();
  });
}

int mps_dev_prod_LR_34(mps_ctx_t ctx, void* out, const void* L, const void* R, int64_t r1, int64_t r2,
                       int64_t l3, int64_t r3, int64_t r4) {
  return guard([&] {
    RAW_PLAN(ctx, mx({r1, l3, r4}), r3, r2);
    prod_LR_34(p.work, p.op, out, L, R, r1, r2, l3, r3, r4, false);
    ctx->c.sync();
  });
}

int mps_dev_prod_LR_43(mps_ctx_t ctx, void* out, const void* L, const void* R, int64_t l1, int64_t l2,
                       int64_t l3, int64_t l4, int64_t r3) {
  return guard([&] {
    RAW_PLAN(ctx, mx({l1, l3, r3}), l2, l4);
    prod_LR_43(p.work, p.op, out, L, R, l1, l2, l3, l4, r3, false);
    ctx->c.sync();
  });
}

int mps_dev_prod_LR_23(mps_ctx_t ctx, void* out, const void* L, const void* R, int64_t l1, int64_t r1,
                       int64_t r2, int64_t r3, int transpose_L) {
  return guard([&] {
    RAW_PLAN(ctx, mx({l1, r1, r3}), r2, 1);
    prod_LR_23(p.work, out, L, R, l1, r1, r2, r3, transpose_L != 0, false);
    ctx->c.sync();
  });
}

int mps_dev_prod_LR_32(mps_ctx_t ctx, void* out, const void* L, const void* R, int64_t l1, int64_t l2,
                       int64_t l3, int64_t r2) {
  return guard([&] {
    RAW_PLAN(ctx, mx({l1, l3, r2}), l2, 1);
    prod_LR_32(p.work, out, L, R, l1, l2, l3, r2, false);
    ctx->c.sync();
  });
}

int mps_dev_update_re(mps_ctx_t ctx, void* out, const void* A0, const void* AOR, int64_t r1, int64_t r2,
                      int64_t r3, int64_t r4, int64_t a1) {
  return guard([&] {
    RAW_PLAN(ctx, mx({r1, r4, a1}), r3, r2);
    update_re(p.work, p.op, out, AOR, A0, r1, r2, r3, r4, a1);
    ctx->c.sync();
  });
}

int mps_dev_update_le(mps_ctx_t ctx, void* out, const void* A0, const void* AOL, int64_t l1, int64_t l2,
                      int64_t l3, int64_t l4, int64_t a3) {
  return guard([&] {
    RAW_PLAN(ctx, mx({l1, l3, a3}), l2, l4);
    update_le(p.work, p.op, out, AOL, A0, l1, l2, l3, l4, a3);
    ctx->c.sync();
  });
}

int mps_dev_update_renv(mps_ctx_t ctx, void* out, const void* A0, const void* AR, int64_t a1, int64_t d,
                        int64_t a2, int64_t ar1) {
  return guard([&] {
    RAW_PLAN(ctx, mx({a1, a2, ar1}), d, 1);
    update_renv(p.work, out, A0, AR, a1, d, a2, ar1);
    ctx->c.sync();
  });
}

int mps_dev_update_lenv(mps_ctx_t ctx, void* out, const void* A0, const void* AL, int64_t a1, int64_t d,
                        int64_t a3, int64_t al3, int mode) {
  return guard([&] {
    REQUIRE(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    RAW_PLAN(ctx, mx({a1, a3, al3}), d, 1);
    update_lenv(p.work, out, A0, AL, a1, d, a3, al3, mode);
    ctx->c.sync();
  });
}

// M[s, s'] = sum_{a,b} A[a,s,b] conj(A0[a,s',b])  (local_exp's Ap * A0p^H)
static void local_exp_matrix(Work& w, void* M, const void* A0, const void* A, i64 a1, i64 d, i64 a3) {
  Gemm g;
  g.M = d;
  g.N = d;
  g.K = a1 * a3;
  g.A = A; g.am = im1(a1); g.ak = im2(a1, 1, a1 * d);
  g.B = A0; g.bk = im2(a1, 1, a1 * d); g.bn = im1(a1); g.conjB = true;
  g.C = M; g.cm = im1(1); g.cn = im1(d);
  run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes);
}

int mps_dev_local_exp(mps_ctx_t ctx, double* out, const void* A0, const void* A, int64_t a1, int64_t d,
                      int64_t a3, const double* O_host) {
  return guard([&] {
    RAW_PLAN(ctx, mx({a1, a3}), d, 1);
    local_exp_matrix(p.work, p.tmp, A0, A, a1, d, a3);
    std::vector<double2> M((size_t)d * d);
    HIPCHK(hipMemcpyAsync(M.data(), p.tmp, d * d * sizeof(double2), hipMemcpyDeviceToHost, ctx->c.stream));
    ctx->c.sync();
    // trace(O * M) (host: d x d)
    cd t = 0.0;
    for (int64_t s = 0; s < d; ++s)
      for (int64_t q = 0; q < d; ++q)
        t += cd(O_host[2 * (s + d * q)], O_host[2 * (s + d * q) + 1]) * cd(M[q + d * s].x, M[q + d * s].y);
    out[0] = t.real();
    out[1] = t.imag();
  });
}

// ------------------------------------------------------------ gauge moves
int mps_dev_qrq(mps_ctx_t ctx, void* dev, int64_t m, int64_t n, int64_t rs, int64_t cs, void* R_out) {
  return guard([&] {
    REQUIRE(m >= n && n >= 1, "qrq!: m >= n >= 1 required");
    RAW_PLAN(ctx, mx({m, n}), 1, 1);
    Scratch qs(&ctx->c);
    (void)qs;
    REQUIRE(qr_workspace_bytes(m, n, ctx->c.num_cus) <= p.scr.used, "qr workspace");
    launch_qr(dev, m, n, rs, cs, ctx->c.qr_tol, p.tau, ctx->c.num_cus, R_out, ctx->c.stream, Bat(), ctx->c.qr_chain);
    ctx->c.sync();
  });
}

#define STATE_PLAN(st, owner, n_ovl)                      \
  set_device(st.ctx);                                     \
  size_t before_ = st.ctx->ws_bytes;                      \
  Plan p(st.ctx, {&st}, {}, n_ovl);                       \
  track_ws(owner, before_);

int mps_leftorth_site(mps_state_t s, int site) {
  return guard([&] {
    State& st = s->s;
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    STATE_PLAN(st, s->owner, 0);
    qr_left(p, &st, site, nullptr);
    st.ctx->sync();
  });
}

int mps_rightorth_site(mps_state_t s, int site) {
  return guard([&] {
    State& st = s->s;
    REQUIRE(site >= 0 && site < st.N, "site out of range");
    STATE_PLAN(st, s->owner, 0);
    qr_right(p, &st, site, nullptr);
    st.ctx->sync();
  });
}

int mps_leftorthogonalize_qr(mps_state_t s, double* norm2) {
  return guard([&] {
    State& st = s->s;
    STATE_PLAN(st, s->owner, 0);
    left_orthogonalize(p, &st, 0);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

int mps_leftorthonormalize_qr(mps_state_t s, double* norm2) {
  return guard([&] {
    State& st = s->s;
    STATE_PLAN(st, s->owner, 0);
    left_orthogonalize(p, &st, 0);
    launch_scale_rsqrt(st.site(st.N - 1), st.size(st.N - 1), st.ctx->dscal, st.ctx->stream, st.ctx->bat);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

int mps_rightorthogonalize_qr(mps_state_t s, double* norm2) {
  return guard([&] {
    State& st = s->s;
    STATE_PLAN(st, s->owner, 0);
    right_orthogonalize(p, &st, 0);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

int mps_rightorthonormalize_qr(mps_state_t s, double* norm2) {
  return guard([&] {
    State& st = s->s;
    STATE_PLAN(st, s->owner, 0);
    right_orthogonalize(p, &st, 0);
    launch_scale_rsqrt(st.site(0), st.size(0), st.ctx->dscal, st.ctx->stream, st.ctx->bat);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

int mps_normalize_lo(mps_state_t s, double* norm2) {
  return guard([&] {
    State& st = s->s;
    set_device(st.ctx);
    const int n = st.N - 1;
    launch_sum_abs2(st.site(n), st.size(n), st.ctx->dscal, st.ctx->stream, st.ctx->bat);
    launch_scale_rsqrt(st.site(n), st.size(n), st.ctx->dscal, st.ctx->stream, st.ctx->bat);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

int mps_norm_lo(mps_state_t s, double* norm2) {
  return guard_ro([&] {
    State& st = s->s;
    set_device(st.ctx);
    const int n = st.N - 1;
    launch_sum_abs2(st.site(n), st.size(n), st.ctx->dscal, st.ctx->stream, st.ctx->bat);
    const double no = read_scalar(st.ctx, 0);
    if (norm2) *norm2 = no;
  });
}

// ------------------------------------------------------------- lockstep views
// A view of copy b of a batched container, owned by a companion context on
// the same stream: every plain entry point works on it.  It shares the
// container's memory and must not outlive it.
static void check_view(const Ctx* batched, const Ctx* comp, int b) {
  REQUIRE(batched->bat.nb > 1 && comp->bat.nb == 1, "views are of batched containers, on a plain context");
  REQUIRE(comp->stream == batched->stream, "the view's context must share the batched context's stream");
  REQUIRE(b >= 0 && b < batched->bat.nb, "copy index out of range");
}

// In a ragged pool, copy b's view is its own chain of n[b] sites: site j and
// bond i at the slots the pool keeps them in.
static const Ctx::Ragged* ragged_of(const Ctx* batched, int b, int N) {
  const Ctx::Ragged* r = batched->rag.get();
  if (r) REQUIRE(N == r->nmax, "ragged pool: containers span the pool's nmax slots");
  (void)b;
  return r;
}

int mps_state_view(mps_state_t s, int b, mps_ctx_t comp, mps_state_t* out) {
  return guard([&] {
    REQUIRE(s && comp && out, "invalid arguments");
    check_view(s->s.ctx, &comp->c, b);
    auto* x = new mps_state_s();
    x->owner = comp;
    x->s = State();
    State& v = x->s;
    const State& o = s->s;
    v.ctx = &comp->c;
    v.N = o.N;
    v.d = o.d;
    v.dims = o.dims;
    v.cap_l = o.cap_l;
    v.cap_r = o.cap_r;
    v.off = o.off;
    if (const Ctx::Ragged* r = ragged_of(o.ctx, b, o.N)) {
      const int n = r->n[b];
      v.N = n;
      v.d.resize(n);
      v.cap_l.resize(n);
      v.cap_r.resize(n);
      v.off.resize(n);
      v.dims.resize(n + 1);
      for (int j = 0; j < n; ++j) {
        const int q = r->slot_of_site(b, j);
        v.d[j] = o.d[q];
        v.cap_l[j] = o.cap_l[q];
        v.cap_r[j] = o.cap_r[q];
        v.off[j] = o.off[q];
      }
      for (int i = 0; i <= n; ++i) v.dims[i] = o.dims[r->slot_of_bond(b, i)];
    }
    v.total = o.total;
    v.buf = boff(o.buf, (int64_t)b * o.ctx->bat.bs);
    v.own = false;
    *out = x;
  });
}

int mps_env_view(mps_env_t e, int b, mps_ctx_t comp, mps_env_t* out) {
  return guard([&] {
    REQUIRE(e && comp && out, "invalid arguments");
    check_view(e->e.ctx, &comp->c, b);
    auto* x = new mps_env_s();
    x->owner = comp;
    Env& v = x->e;
    const Env& o = e->e;
    v.ctx = &comp->c;
    v.N = o.N;
    v.op = o.op;
    v.d1 = o.d1;
    v.wd = o.wd;
    v.d2 = o.d2;
    v.off = o.off;
    if (const Ctx::Ragged* r = ragged_of(o.ctx, b, o.N)) {
      const int n = r->n[b];
      v.N = n;
      v.d1.resize(n + 1);
      v.wd.resize(n + 1);
      v.d2.resize(n + 1);
      v.off.resize(n + 1);
      for (int i = 0; i <= n; ++i) {
        const int q = r->slot_of_bond(b, i);
        v.d1[i] = o.d1[q];
        v.wd[i] = o.wd[q];
        v.d2[i] = o.d2[q];
        v.off[i] = o.off[q];
      }
    }
    v.total = o.total;
    v.buf = boff(o.buf, (int64_t)b * o.ctx->bat.bs);
    v.own = false;
    *out = x;
  });
}

int mps_mpo_view(mps_mpo_t m, int b, mps_ctx_t comp, mps_mpo_t* out) {
  return guard([&] {
    REQUIRE(m && comp && out, "invalid arguments");
    check_view(m->o.ctx, &comp->c, b);
    auto* x = new mps_mpo_s();
    x->owner = comp;
    MPO& v = x->o;
    const MPO& o = m->o;
    v.ctx = &comp->c;
    v.N = o.N;
    v.d = o.d;
    v.w = o.w;
    v.off = o.off;
    if (const Ctx::Ragged* r = ragged_of(o.ctx, b, o.N)) {
      const int n = r->n[b];
      v.N = n;
      v.d.resize(n);
      v.off.resize(n);
      v.w.resize(n + 1);
      for (int j = 0; j < n; ++j) {
        v.d[j] = o.d[r->slot_of_site(b, j)];
        v.off[j] = o.off[r->slot_of_site(b, j)];
      }
      for (int i = 0; i <= n; ++i) v.w[i] = o.w[r->slot_of_bond(b, i)];
    }
    v.total = o.total;
    v.buf = boff(o.buf, (int64_t)b * o.ctx->bat.bs);
    v.own = false;
    *out = x;
  });
}

// normalize_lo! of every copy of a batched state: norm2[b] for copy b
int mps_normalize_lo_batch(mps_state_t s, double* norm2) {
  return guard([&] {
    REQUIRE(s && norm2, "invalid arguments");
    State& st = s->s;
    Ctx* c = st.ctx;
    set_device(c);
    const int n = st.N - 1;
    launch_sum_abs2(st.site(n), st.size(n), c->dscal, c->stream, c->bat);
    launch_scale_rsqrt(st.site(n), st.size(n), c->dscal, c->stream, c->bat);
    HIPCHK(hipMemcpy2DAsync(c->hscal, sizeof(double), c->dscal, (size_t)c->bat.bs, sizeof(double),
                            (size_t)c->bat.nb, hipMemcpyDeviceToHost, c->stream));
    c->sync();
    for (int b = 0; b < c->bat.nb; ++b) norm2[b] = c->hscal[b];
  });
}

// ------------------------------------------------------------------ sweeps
static void check_states(std::initializer_list<const State*> v) {
  const State* f = *v.begin();
  for (const State* s : v) {
    REQUIRE(s->N == f->N, "chain lengths differ");
    REQUIRE(s->ctx == f->ctx, "states belong to different contexts");
    for (int j = 0; j < s->N; ++j) REQUIRE(s->d[j] == f->d[j], "local dims differ");
  }
}

static void check_op_env(const Env& e, const State& s) {
  REQUIRE(e.op, "operator environment expected (build_env with mpo_dims)");
  REQUIRE(e.N == s.N, "environment length differs from chain length");
}

static void check_ovl_env(const Env& e, const State& s) {
  REQUIRE(!e.op, "overlap environment expected (build_env without mpo_dims)");
  REQUIRE(e.N == s.N, "environment length differs from chain length");
}

static void check_mpo(const MPO& o, const State& s) {
  REQUIRE(o.N == s.N, "MPO length differs from chain length");
  for (int j = 0; j < s.N; ++j) REQUIRE(o.d[j] == s.d[j], "MPO local dims differ");
}

// The H-applying sweeps, RK4 and Arnoldi index every state of a call with one
// set of bonds, as the reference builds them all from one dims vector; states
// of other bonds are rejected before anything is queued.
static void check_same_bonds(std::initializer_list<const State*> v) {
  const State* f = *v.begin();
  for (const State* s : v) REQUIRE(s->dims == f->dims, "states differ in bond dimensions");
}

// Slot j of an environment receives the [a.dims[j], w[j], b.dims[j]] block
// (w = 1 for overlaps) that a sweep pairing a with b writes there; env_create
// gave it max(d1, d2)^2 wd elements.  A slot too small for the states' bonds
// is an error here rather than a write past the slot.
static void check_env_fits(const Env& e, const State& a, const State& b, const MPO* H) {
  REQUIRE(e.N == a.N && e.N == b.N, "environment length differs from chain length");
  for (int j = 0; j <= e.N; ++j) {
    const i64 big = std::max(e.d1[j], e.d2[j]);
    const i64 need = a.dims[j] * (H ? H->w[j] : 1) * b.dims[j];
    REQUIRE(need <= big * e.wd[j] * big,
            "environment slot " + std::to_string(j) + " is smaller than the states' bonds need");
  }
}

// The entry points whose sweeps launch per slot for the copies present there
// (RK4 and the H-applying half-sweeps) accept a ragged pool's containers once
// every copy's two halves meet consistently: the states' bonds, the MPOs'
// bonds and every environment slot across the copy's gap agree, so each
// copy is a chain of its own with the dims the slots give it.
struct RaggedCall {
  Ctx* c;
  RaggedCall(Ctx* ctx, const std::vector<const State*>& ss, const std::vector<const MPO*>& ms,
             const std::vector<const Env*>& es)
      : c(ctx) {
    if (!c->rag) return;
    const Ctx::Ragged& r = *c->rag;
    for (size_t b = 0; b < r.n.size(); ++b) {
      const int qa = r.h[b], qb = r.nmax - r.n[b] + r.h[b];
      for (const State* s : ss) {
        REQUIRE(s->N == r.nmax, "ragged pool: states span the pool's nmax slots");
        REQUIRE(s->dims[qa] == s->dims[qb], "ragged pool: a chain's two parts meet at bonds of different dimension");
      }
      for (const MPO* m : ms) {
        REQUIRE(m->N == r.nmax, "ragged pool: MPOs span the pool's nmax slots");
        REQUIRE(m->w[qa] == m->w[qb], "ragged pool: a chain's two parts meet at MPO bonds of different dimension");
      }
      for (const Env* e : es) {
        REQUIRE(e->N == r.nmax, "ragged pool: environments span the pool's nmax slots");
        for (int q = qa; q <= qb; ++q)
          REQUIRE(e->d1[q] == e->d1[qa] && e->wd[q] == e->wd[qa] && e->d2[q] == e->d2[qa],
                  "ragged pool: environment slots differ inside a chain's gap");
      }
    }
    c->rag_ok = true;
  }
  ~RaggedCall() { c->rag_ok = false; }
  RaggedCall(const RaggedCall&) = delete;
  RaggedCall& operator=(const RaggedCall&) = delete;
};

int mps_compress_var_apply_H_left(mps_state_t A0, mps_env_t envop, mps_state_t Ain, mps_state_t A,
                                  mps_mpo_t H) {
  return guard([&] {
    check_states({&A0->s, &Ain->s, &A->s});
    check_op_env(envop->e, A0->s);
    check_mpo(H->o, A0->s);
    check_same_bonds({&A0->s, &Ain->s, &A->s});
    check_env_fits(envop->e, A0->s, A0->s, &H->o);
    Ctx* c = A0->s.ctx;
    set_device(c);
    RaggedCall rc(c, {&A0->s, &Ain->s, &A->s}, {&H->o}, {&envop->e});
    size_t before = c->ws_bytes;
    Plan p(c, {&A0->s, &Ain->s, &A->s}, {&H->o}, 0);
    track_ws(A0->owner, before);
    compress_var_apply_H_left(p, &A0->s, &envop->e, &Ain->s, &A->s, &H->o);
    c->sync();
  });
}

static void sum_var_common(mps_state_t A0, const mps_env_t* env, mps_env_t envop, mps_state_t Ain, int numv,
                           const double* coef, const mps_state_t* A, mps_mpo_t H, bool right) {
  // the reference always forms the first overlap term (compression.jl:918, 1036): numv >= 2
  REQUIRE(numv >= 2 && A && coef, "numv >= 2 states and coefficients required (BoundsError)");
  std::vector<State*> As(numv);
  std::vector<Env*> Es(numv - 1);
  std::vector<cd> cf(numv);
  for (int l = 0; l < numv; ++l) {
    As[l] = &A[l]->s;
    cf[l] = cd(coef[2 * l], coef[2 * l + 1]);
    check_states({&A0->s, As[l]});
  }
  for (int l = 0; l + 1 < numv; ++l) {
    Es[l] = &env[l]->e;
    check_ovl_env(*Es[l], A0->s);
  }
  check_states({&A0->s, &Ain->s});
  check_op_env(envop->e, A0->s);
  check_mpo(H->o, A0->s);
  check_same_bonds({&A0->s, &Ain->s});
  for (int l = 0; l < numv; ++l) check_same_bonds({&A0->s, As[l]});
  for (int l = 0; l + 1 < numv; ++l) check_env_fits(*Es[l], A0->s, A0->s, nullptr);
  check_env_fits(envop->e, A0->s, A0->s, &H->o);
  Ctx* c = A0->s.ctx;
  set_device(c);
  std::vector<const State*> rs{&A0->s, &Ain->s};
  std::vector<const Env*> re{&envop->e};
  for (int l = 0; l < numv; ++l) rs.push_back(As[l]);
  for (int l = 0; l + 1 < numv; ++l) re.push_back(Es[l]);
  RaggedCall rc(c, rs, {&H->o}, re);
  i64 Dm = std::max(A0->s.max_dim(), Ain->s.max_dim()), dm = 1;
  for (int l = 0; l < numv; ++l) Dm = std::max(Dm, As[l]->max_dim());
  for (int j = 0; j < A0->s.N; ++j) dm = std::max(dm, A0->s.d[j]);
  size_t before = c->ws_bytes;
  Plan pp(c, Dm, dm, H->o.max_w(), numv);
  track_ws(A0->owner, before);
  if (right)
    compress_sum_var_apply_H_right(pp, &A0->s, Es.data(), &envop->e, &Ain->s, numv, cf.data(), As.data(),
                                   &H->o);
  else
    compress_sum_var_apply_H_left(pp, &A0->s, Es.data(), &envop->e, &Ain->s, numv, cf.data(), As.data(),
                                  &H->o);
  c->sync();
}

int mps_compress_sum_var_apply_H_left(mps_state_t A0, const mps_env_t* env, mps_env_t envop,
                                      mps_state_t Ain, int numv, const double* coef, const mps_state_t* A,
                                      mps_mpo_t H) {
  return guard([&] { sum_var_common(A0, env, envop, Ain, numv, coef, A, H, false); });
}

int mps_compress_sum_var_apply_H_right(mps_state_t A0, const mps_env_t* env, mps_env_t envop,
                                       mps_state_t Ain, int numv, const double* coef, const mps_state_t* A,
                                       mps_mpo_t H) {
  return guard([&] { sum_var_common(A0, env, envop, Ain, numv, coef, A, H, true); });
}

int mps_compress_var_apply_H(mps_state_t A0, mps_env_t envop, mps_state_t Ain, mps_state_t A, mps_mpo_t H,
                             int maxsweeps) {
  return guard([&] {
    check_states({&A0->s, &Ain->s, &A->s});
    check_op_env(envop->e, A0->s);
    check_mpo(H->o, A0->s);
    REQUIRE(maxsweeps >= 0, "maxsweeps >= 0");
    check_same_bonds({&A0->s, &Ain->s, &A->s});
    check_env_fits(envop->e, A0->s, A0->s, &H->o);
    Ctx* c = A0->s.ctx;
    set_device(c);
    size_t before = c->ws_bytes;
    Plan p(c, {&A0->s, &Ain->s, &A->s}, {&H->o}, 0);
    track_ws(A0->owner, before);
    compress_var_apply_H(p, &A0->s, &envop->e, &Ain->s, &A->s, &H->o, maxsweeps);
    c->sync();
  });
}

int mps_compress_var(mps_state_t A0, mps_state_t A, mps_env_t env, int maxsweeps) {
  return guard([&] {
    check_states({&A0->s, &A->s});
    no_batch(A0->s.ctx);
    check_ovl_env(env->e, A0->s);
    REQUIRE(maxsweeps >= 0, "maxsweeps >= 0");
    check_env_fits(env->e, A0->s, A->s, nullptr);
    Ctx* c = A0->s.ctx;
    set_device(c);
    size_t before = c->ws_bytes;
    Plan p(c, {&A0->s, &A->s}, {}, 0);
    track_ws(A0->owner, before);
    compress_var(p, &A0->s, &A->s, &env->e, maxsweeps);
    c->sync();
  });
}

// Temporary overlap environments for the allocating reference routines.
struct TmpEnvs {
  std::vector<mps_env_t> v;
  ~TmpEnvs() {
    for (auto e : v) mps_env_destroy(e);
  }
};

static void make_tmp_envs(mps_ctx_s* ctx, const State& s, int n, TmpEnvs& t, std::vector<Env*>& out) {
  std::vector<int64_t> big(s.N + 1);
  for (int j = 0; j <= s.N; ++j) big[j] = s.dims[j];
  for (int i = 0; i < n; ++i) {
    mps_env_t e = nullptr;
    int rc = mps_env_create(ctx, s.N, big.data(), big.data(), &e);
    if (rc != MPS_OK) throw Error(rc, g_err);
    t.v.push_back(e);
    out.push_back(&e->e);
  }
}

int mps_compress_sum_var(mps_state_t A0out, mps_state_t Ain, int numv, const double* coef,
                         const mps_state_t* A, int maxsweeps) {
  return guard([&] {
    REQUIRE(numv >= 1 && A && coef, "numv >= 1 states and coefficients required");
    no_batch(Ain->s.ctx);
    check_states({&A0out->s, &Ain->s});
    std::vector<State*> As(numv);
    std::vector<cd> cf(numv);
    bool alias = false;
    for (int l = 0; l < numv; ++l) {
      As[l] = &A[l]->s;
      cf[l] = cd(coef[2 * l], coef[2 * l + 1]);
      check_states({&A0out->s, As[l]});
      check_same_bonds({&A0out->s, &Ain->s, As[l]});
      if (A[l] == A0out) alias = true;
    }
    Ctx* c = A0out->s.ctx;
    set_device(c);
    TmpEnvs te;
    std::vector<Env*> FL;
    make_tmp_envs(A0out->owner, Ain->s, numv, te, FL);
    // the reference returns a fresh A0: write to a temporary when A0out is one of A
    mps_state_t out = A0out;
    mps_state_t tmp = nullptr;
    if (alias) {
      int rc = mps_state_create(A0out->owner, Ain->s.N, Ain->s.d.data(), Ain->s.dims.data(), &tmp);
      if (rc != MPS_OK) throw Error(rc, g_err);
      out = tmp;
    }
    try {
      size_t before = c->ws_bytes;
      Plan p(c, {&out->s, &Ain->s}, {}, numv);
      track_ws(A0out->owner, before);
      compress_sum_var(p, &out->s, &Ain->s, numv, cf.data(), As.data(), maxsweeps, FL);
      if (alias) {
        A0out->s.dims = out->s.dims;
        for (int j = 0; j < out->s.N; ++j) launch_copy_c(A0out->s.site(j), out->s.site(j), out->s.size(j), c->stream, c->bat);
      }
      c->sync();
    } catch (...) {
      if (tmp) mps_state_destroy(tmp);
      throw;
    }
    if (tmp) mps_state_destroy(tmp);
  });
}

int mps_compress_sum_var_apply_H(mps_state_t A0, mps_env_t envop, int numv, const double* coef,
                                 const mps_state_t* A, mps_mpo_t H, int maxsweeps) {
  return guard([&] {
    REQUIRE(numv >= 2 && A && coef, "numv >= 2 states and coefficients required (BoundsError)");
    std::vector<State*> As(numv);
    std::vector<cd> cf(numv);
    for (int l = 0; l < numv; ++l) {
      As[l] = &A[l]->s;
      cf[l] = cd(coef[2 * l], coef[2 * l + 1]);
      check_states({&A0->s, As[l]});
      check_same_bonds({&A0->s, As[l]});
      REQUIRE(A[l] != A0, "compress_sum_var_apply_H!: A0 must not be one of A");
    }
    check_op_env(envop->e, A0->s);
    check_mpo(H->o, A0->s);
    check_env_fits(envop->e, A0->s, A0->s, &H->o);
    Ctx* c = A0->s.ctx;
    set_device(c);
    TmpEnvs te;
    std::vector<Env*> env;
    make_tmp_envs(A0->owner, A0->s, numv - 1, te, env);
    size_t before = c->ws_bytes;
    Plan p(c, {&A0->s}, {&H->o}, numv);
    track_ws(A0->owner, before);
    compress_sum_var_apply_H(p, &A0->s, &envop->e, numv, cf.data(), As.data(), &H->o, maxsweeps, env);
    c->sync();
  });
}

// compressionSVD! (compression.jl:1280-1312), right to left: site i times the
// carried r, SVD of reshape(., dl, d*dnew), truncation by the cumulative
// weight rule, site i <- V[:, 1:dnew]^H, r <- U[:, 1:dnew] S.  The SVD is the
// one-sided Jacobi of svd.hip on whichever of X^H / X has fewer columns.
int mps_compression_svd(mps_state_t s, int64_t dmax, double tol, double* err_glob, double* err_trim,
                        double* ent_ent) {
  return guard([&] {
    State& st = s->s;
    no_batch(st.ctx);
    REQUIRE(dmax >= 1, "compressionSVD!: dmax >= 1 required");
    REQUIRE(st.dims[0] == 1 && st.dims[st.N] == 1, "compressionSVD!: open boundary bonds must be 1");
    // small device buffers: sigma | scale | perm | flag, in the plan's extra buffer
    // (the SVD of a site has at most max_dim * max d columns)
    i64 dm = 1;
    for (int j = 0; j < st.N; ++j) dm = std::max(dm, st.d[j]);
    const i64 nmax = st.max_dim() * dm + 64;
    set_device(st.ctx);
    size_t before_ = st.ctx->ws_bytes;
    Plan p(st.ctx, {&st}, {}, 0, (size_t)nmax * (2 * sizeof(double) + sizeof(int)) + 256);
    track_ws(s->owner, before_);
    Ctx* c = st.ctx;
    const std::vector<i64> old = st.dims;
    double* sig_d = (double*)p.extra;
    double* scl_d = sig_d + nmax;
    int* perm_d = (int*)(scl_d + nmax);
    int* flag_d = perm_d + nmax;
    int* flag_h = (int*)c->hscal;
    launch_set_scalar(p.AL, 1.0, 0.0, c->stream, c->bat);   // r = 1 (1 x 1)
    i64 dnew = 1;
    std::vector<double> sig;
    for (int i = st.N - 1; i >= 0; --i) {
      const i64 dl = old[i], d = st.d[i], dr = old[i + 1];
      // tmp (dl d x dnew) = reshape(site, dl d, dr) * r (dr x dnew)
      Gemm g;
      g.M = dl * d; g.N = dnew; g.K = dr;
      g.A = st.site(i); g.am = im1(1); g.ak = im1(dl * d);
      g.B = p.AL; g.bk = im1(1); g.bn = im1(dr);
      g.C = p.tmp; g.cm = im1(1); g.cn = im1(dl * d);
      run_gemm(c, g, CMP_F64, p.splitk, p.splitk_bytes);
      // X = reshape(tmp, dl, d dnew); Y = X^H (cols = dl) or X (cols = d dnew)
      const i64 mc = d * dnew;
      const bool ct = mc >= dl;
      const i64 m = ct ? mc : dl, n = ct ? dl : mc;
      REQUIRE(nmax > n, "compressionSVD!: bond dimension too large");
      svd_load(p.T1, p.tmp, m, n, ct, c->stream);
      // rotation threshold: rounding in y_p^H y_q over m rows is ~sqrt(m) eps sqrt(alpha beta)
      svd_jacobi(p.T1, m, n, p.T2, 4.0 * std::sqrt((double)m) * 2.220446049250313e-16, flag_d, flag_h, 60,
                 c->stream);
      svd_colnorms(p.T1, m, n, sig_d, c->stream);
      sig.assign(n, 0.0);
      HIPCHK(hipMemcpyAsync(sig.data(), sig_d, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      c->sync();
      // s = s[s .> 0], descending
      std::vector<int> order;
      for (int k = 0; k < n; ++k)
        if (sig[k] > 0.0) order.push_back(k);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
      REQUIRE(!order.empty(), "compressionSVD!: zero site");
      double stot = 0.0;
      for (int k : order) stot += sig[k] * sig[k];
      i64 cnt = 0;
      double cum = 0.0;
      for (int k : order) {
        cum += sig[k] * sig[k] / stot;
        if (cum < 1.0 - tol) ++cnt;
      }
      const i64 dn = std::min<i64>(cnt + 1, dmax);
      REQUIRE(dn <= (i64)order.size(), "compressionSVD!: truncation rank exceeds the number of singular values");
      double trim = 0.0, ent = 0.0;
      for (size_t k = 0; k < order.size(); ++k) {
        const double s2 = sig[order[k]] * sig[order[k]];
        if ((i64)k >= dn) trim += s2;
        ent += s2 * std::log(sig[order[k]]);
      }
      err_trim[i] = trim / stot;
      ent_ent[i] = -2.0 * ent;
      std::vector<double> scl(dn);
      std::vector<int> perm(order.begin(), order.begin() + dn);
      for (i64 k = 0; k < dn; ++k) scl[k] = ct ? 1.0 / sig[perm[k]] : sig[perm[k]];
      HIPCHK(hipMemcpyAsync(perm_d, perm.data(), dn * sizeof(int), hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipMemcpyAsync(scl_d, scl.data(), dn * sizeof(double), hipMemcpyHostToDevice, c->stream));
      // site i <- reshape(V[:, 1:dn]^H, dn, d, dnew): V = Y Sigma^-1 (ct) or J
      if (ct)
        svd_gather(st.site(i), p.T1, m, perm_d, scl_d, dn, m, true, false, c->stream);
      else
        svd_gather(st.site(i), p.T2, n, perm_d, nullptr, dn, n, true, false, c->stream);
      // r (dl x dn) = U[:, 1:dn] S: J S (ct) or Y
      if (ct) {
        for (i64 k = 0; k < dn; ++k) scl[k] = sig[perm[k]];
        HIPCHK(hipMemcpyAsync(scl_d, scl.data(), dn * sizeof(double), hipMemcpyHostToDevice, c->stream));
        svd_gather(p.AL, p.T2, n, perm_d, scl_d, dn, dl, false, true, c->stream);
      } else {
        svd_gather(p.AL, p.T1, m, perm_d, nullptr, dn, dl, false, true, c->stream);
      }
      c->sync();   // host vectors scl / perm are reused next site
      st.dims[i] = dn;
      dnew = dn;
    }
    double r0[2];
    HIPCHK(hipMemcpyAsync(r0, p.AL, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    *err_glob = 1.0 - (r0[0] * r0[0] + r0[1] * r0[1]);
    invalidate(s->owner);
  });
}

// two-site variational sweeps (compression.jl:110-164, 355-486, 574-647, 1072-1273)
int mps_compress_two_site(mps_state_t A0, const mps_env_t* env, mps_env_t envop, mps_state_t Ain, int numv,
                          const double* coef, const mps_state_t* A, mps_mpo_t H, int mode, int maxsweeps,
                          double svd_tol, int64_t dlim) {
  return guard([&] {
    REQUIRE(A0 && Ain && A && coef && numv >= 1, "invalid arguments");
    REQUIRE(mode >= 0 && mode <= 2 && maxsweeps >= 0 && dlim >= 1 && svd_tol >= 0.0 && svd_tol < 1.0,
            "mode 0..2, maxsweeps >= 0, dlim >= 1, svd_tol in [0, 1)");
    REQUIRE(A0 != Ain, "two-site sweeps write a new state: A0 must not be Ain");
    State& a0 = A0->s;
    no_batch(a0.ctx);
    REQUIRE(a0.N >= 2, "two-site sweeps need at least two sites");
    const int nov = H ? numv - 1 : numv;
    REQUIRE(nov >= 0 && (nov == 0 || env), "environments for the overlap terms required");
    std::vector<State*> As(numv);
    std::vector<Env*> Es(nov);
    std::vector<cd> cf(numv);
    for (int l = 0; l < numv; ++l) {
      As[l] = &A[l]->s;
      REQUIRE(A[l] != A0, "A0 must not be one of A");
      check_states({&a0, As[l]});
      cf[l] = cd(coef[2 * l], coef[2 * l + 1]);
    }
    check_states({&a0, &Ain->s});
    for (int j = 0; j <= a0.N; ++j)
      REQUIRE(Ain->s.dims[j] <= (j < a0.N ? a0.cap_l[j] : a0.cap_r[j - 1]), "A0's capacity is smaller than Ain's bonds");
    for (int l = 0; l < nov; ++l) {
      Es[l] = &env[l]->e;
      check_ovl_env(*Es[l], a0);
    }
    if (H) {
      REQUIRE(envop, "operator environment required");
      check_op_env(envop->e, a0);
      check_mpo(H->o, a0);
    }
    // environments must hold A0's capacity bonds against each target's bonds
    std::vector<i64> capd(a0.N + 1);
    for (int j = 0; j <= a0.N; ++j) capd[j] = j < a0.N ? a0.cap_l[j] : a0.cap_r[a0.N - 1];
    State cap_like;
    cap_like.N = a0.N;
    cap_like.dims = capd;
    for (int l = 0; l < nov; ++l) check_env_fits(*Es[l], cap_like, *As[l], nullptr);
    if (H) check_env_fits(envop->e, cap_like, *As[nov], &H->o);
    Ctx* c = a0.ctx;
    set_device(c);
    i64 Dm = a0.max_dim(), dm = 1, wm = H ? H->o.max_w() : 1;
    for (int l = 0; l < numv; ++l) Dm = std::max(Dm, As[l]->max_dim());
    for (int j = 0; j < a0.N; ++j) dm = std::max(dm, a0.d[j]);
    size_t before = c->ws_bytes;
    Plan p(c, Dm, dm, wm, 2 * std::max(nov, 1), two_site_extra_bytes(Dm, dm));
    track_ws(A0->owner, before);
    two_site_sweeps(p, &a0, Es.data(), H ? &envop->e : nullptr, &Ain->s, nov, cf.data(), As.data(),
                    H ? &H->o : nullptr, mode, maxsweeps, svd_tol, dlim);
    invalidate(A0->owner);
  });
}

// steady_state_search! (steadystate.jl:6-86)
int mps_steady_state_search(mps_state_t At, mps_mpo_t W, mps_env_t envop, int maxsweeps, double converge_rat,
                            double eigs_tol, double* evalue_out, int* sweeps_out) {
  return guard([&] {
    REQUIRE(At && W && envop, "invalid arguments");
    State& s = At->s;
    no_batch(s.ctx);
    check_op_env(envop->e, s);
    check_mpo(W->o, s);
    check_env_fits(envop->e, s, s, &W->o);
    REQUIRE(maxsweeps >= 0 && eigs_tol > 0.0, "maxsweeps >= 0 and eigs_tol > 0 required");
    Ctx* c = s.ctx;
    set_device(c);
    constexpr int kmax = 48;
    size_t before = c->ws_bytes;
    Plan p(c, {&s}, {&W->o}, 0, steady_state_extra_bytes(s.max_site(), kmax));
    track_ws(At->owner, before);
    p.op = Op{CMP_F64, F64};
    int sw = 0;
    const double ae = steady_state_search(p, &s, &W->o, &envop->e, maxsweeps, converge_rat, eigs_tol, kmax, &sw);
    if (evalue_out) *evalue_out = ae;
    if (sweeps_out) *sweeps_out = sw;
  });
}

// ----------------------------------------------------------------- RK4
int mps_rk4stp_apply_H_half(double dt, mps_mpo_t H1, mps_mpo_t H2, mps_mpo_t H3, mps_state_t psi0,
                            mps_state_t psi, mps_state_t psi1, mps_state_t psi2, mps_state_t psi3,
                            mps_env_t env1, mps_env_t env2, mps_env_t env3, mps_env_t envop) {
  return guard([&] {
    mps_state_t out = psi0 ? psi0 : psi;
    check_states({&psi->s, &psi1->s, &psi2->s, &psi3->s, &out->s});
    for (auto e : {env1, env2, env3}) check_ovl_env(e->e, psi->s);
    check_op_env(envop->e, psi->s);
    for (auto h : {H1, H2, H3}) check_mpo(h->o, psi->s);
    check_same_bonds({&psi->s, &psi1->s, &psi2->s, &psi3->s, &out->s});
    for (auto e : {env1, env2, env3}) check_env_fits(e->e, psi->s, psi->s, nullptr);
    for (auto h : {H1, H2, H3}) check_env_fits(envop->e, psi->s, psi->s, &h->o);
    mps_ctx_s* x = psi->owner;
    Ctx* c = &x->c;
    set_device(c);
    RaggedCall rc(c, {&psi->s, &psi1->s, &psi2->s, &psi3->s, &out->s}, {&H1->o, &H2->o, &H3->o},
                  {&env1->e, &env2->e, &env3->e, &envop->e});
    mps_ctx_s::Key key{{H1, H2, H3, psi0, psi, psi1, psi2, psi3, env1, env2, env3, envop}, dt, c->prec,
                       c->qr_tol, x->ws_gen, c->refine};
    auto run = [&](Plan& p) {
      rk4_half(p, dt, &H1->o, &H2->o, &H3->o, &out->s, &psi->s, &psi1->s, &psi2->s, &psi3->s, &env1->e,
               &env2->e, &env3->e, &envop->e);
    };
    size_t before = c->ws_bytes;
    Plan p(c, {&psi->s, &psi1->s, &psi2->s, &psi3->s, &out->s}, {&H1->o, &H2->o, &H3->o}, 4);
    track_ws(x, before);
    if (c->refine > 0) {
      p.rtmp = internal_like(x, x->rtmp, psi->s);
      if (out == psi) p.rkeep = internal_like(x, x->rkeep, psi->s);
    }
    if (p.conc) {
      const Env& like = envop->e;
      if (!x->shadow || x->shadow->N != like.N || x->shadow->off != like.off || x->shadow->total != like.total) {
        x->shadow.reset();
        auto e = std::make_unique<Env>();
        e->ctx = c;
        e->N = like.N;
        e->op = like.op;
        e->d1 = like.d1;
        e->wd = like.wd;
        e->d2 = like.d2;
        e->off = like.off;
        e->total = like.total;
        e->buf = (double2*)c->alloc(like.total * sizeof(double2), e->own);
        x->shadow = std::move(e);
      }
      p.shadow = x->shadow.get();
    }
    key.ws_gen = x->ws_gen;
    // Asynchronous on the context's stream (no host value is produced): the
    // caller's next call on this context, or a scalar-returning call, orders
    // after it; coefficients are kernel arguments, never host memory.
    // Captured graphs serve the launch-bound small solves; a plan large enough
    // for the side stream runs eagerly (its launches are hidden behind the
    // kernels, and measured, a second live graph with side-stream branches in
    // the same context ran at half speed).
    // A graph is keyed by the handles and holds their device pointers; a
    // handle's destruction clears the graphs of its own context only.  So a
    // call whose MPOs, environments or states belong to another context runs
    // eagerly: a graph of it could outlive one of them.
    bool own_all = true;
    for (auto h : {H1, H2, H3}) own_all = own_all && h->owner == x;
    for (auto e : {env1, env2, env3, envop}) own_all = own_all && e->owner == x;
    for (auto s : {out, psi, psi1, psi2, psi3}) own_all = own_all && s->owner == x;
    const bool had_spec = x->spec;
    x->spec = false;
    if (!c->graphs || c->profiling || p.conc || !own_all) {
      if (p.conc && p.shadow) {
        const bool hit = had_spec && g_epoch.load() == x->spec_epoch + 1 && x->spec_H1 == H1 &&
                         x->spec_psi == psi && x->spec_envop == envop && x->spec_prec == c->prec;
        rk4_half(p, dt, &H1->o, &H2->o, &H3->o, &out->s, &psi->s, &psi1->s, &psi2->s, &psi3->s, &env1->e,
                 &env2->e, &env3->e, &envop->e, hit, true);
        x->spec = true;
        x->spec_epoch = g_epoch.load();
        x->spec_H1 = H1;
        x->spec_psi = out;
        x->spec_envop = envop;
        x->spec_prec = c->prec;
        return;
      }
      run(p);
      if (c->profiling) c->sync();
      return;
    }
    auto it = x->graphs.find(key);
    if (it == x->graphs.end()) {
      int& n = x->seen[key];
      if (n++ == 0) {   // first call: eager (sizes the workspace)
        run(p);
        return;
      }
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      c->capturing = true;
      try {
        run(p);
      } catch (...) {
        c->capturing = false;
        hipGraph_t dummy = nullptr;
        (void)hipStreamEndCapture(c->stream, &dummy);
        if (dummy) (void)hipGraphDestroy(dummy);
        throw;
      }
      c->capturing = false;
      HIPCHK(hipStreamEndCapture(c->stream, &g));
      hipGraphExec_t ex = nullptr;
      HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
      (void)hipGraphDestroy(g);
      it = x->graphs.emplace(key, ex).first;
    }
    HIPCHK(hipGraphLaunch(it->second, c->stream));
  });
}

// ------------------------------------------------------------ scalar products
static void scal_common(mps_state_t A, mps_state_t B, int mode, double* out) {
  check_states({&A->s, &B->s});
  no_batch(A->s.ctx);
  Ctx* c = A->s.ctx;
  set_device(c);
  size_t before = c->ws_bytes;
  Plan p(c, {&A->s, &B->s}, {}, 0);
  track_ws(A->owner, before);
  scal_prod(p, &A->s, &B->s, mode, 0);
  HIPCHK(hipMemcpyAsync(c->hscal, c->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  out[0] = c->hscal[0];
  out[1] = c->hscal[1];
}

int mps_scal_prod(mps_state_t A, mps_state_t B, double* out) {
  return guard_ro([&] { scal_common(A, B, 0, out); });
}

int mps_scal_prod_no_conj(mps_state_t A, mps_state_t B, double* out) {
  return guard_ro([&] { scal_common(A, B, 2, out); });
}

// scal_prod_no_conj in the fixed buffer `scratch`: no device allocation, no
// split-K partials (each GEMM sums K in one pass), result through the
// context's pinned host scalars.
int mps_scal_prod_no_conj_noalloc(mps_state_t A, mps_state_t B, double* out) {
  return guard_ro([&] {
    REQUIRE(A && B && out, "invalid arguments");
    State& a = A->s;
    State& b = B->s;
    check_states({&a, &b});
    Ctx* c = a.ctx;
    if (c->bat.nb != 1) throw Error(MPS_E_STATE, "not available on a batched context (use a view of one copy)");
    if (c->profiling) throw Error(MPS_E_STATE, "kernel timing is on; it creates events on every launch");
    i64 env = 1, al = 1;
    for (int j = 0; j <= a.N; ++j) env = std::max(env, a.dims[j] * b.dims[j]);
    for (int j = 0; j < a.N; ++j) al = std::max(al, a.Dl(j) * b.d[j] * b.Dr(j));
    auto r256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t eb = r256((size_t)env * sizeof(double2)), ab = r256((size_t)al * sizeof(double2));
    const size_t need = 2 * eb + ab + 256;
    REQUIRE(need <= kNoAllocScratchBytes,
            "scal_prod_no_conj without allocation needs " + std::to_string(need) +
                " bytes of the library's fixed device buffer, which holds " +
                std::to_string(kNoAllocScratchBytes) + " (use mps_scal_prod_no_conj)");
    set_device(c);
    char* base = scratch_on(c->device);
    std::lock_guard<std::mutex> lk(g_scratch_call_mu[c->device]);
    Work w;
    w.c = c;   // no split-K buffer: GEMMs are not split
    scal_prod_ws(w, &a, &b, 2, base, base + eb, base + 2 * eb, base + 2 * eb + ab);
    HIPCHK(hipMemcpyAsync(c->hscal, base + 2 * eb + ab, sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    out[0] = c->hscal[0];
    out[1] = c->hscal[1];
  });
}

int mps_scal_op_prod(mps_state_t A, mps_mpo_t O, mps_state_t B, double* out) {
  return guard_ro([&] {
    check_states({&A->s, &B->s});
    no_batch(A->s.ctx);
    check_mpo(O->o, A->s);
    Ctx* c = A->s.ctx;
    set_device(c);
    size_t before = c->ws_bytes;
    Plan p(c, {&A->s, &B->s}, {&O->o}, 0);
    track_ws(A->owner, before);
    scal_op_prod(p, &A->s, &O->o, &B->s, 0);
    HIPCHK(hipMemcpyAsync(c->hscal, c->dscal, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    out[0] = c->hscal[0];
    out[1] = c->hscal[1];
  });
}

// measure_excitations! (spin_model_quantum_trajectory_rk4.jl:220-231)
int mps_measure_local(mps_state_t A, const double* op, double* out) {
  return guard_ro([&] {
    State& s = A->s;
    no_batch(s.ctx);
    Ctx* c = s.ctx;
    set_device(c);
    size_t before = c->ws_bytes;
    i64 dmax = 1;
    for (int j = 0; j < s.N; ++j) dmax = std::max(dmax, s.d[j]);
    // per-site d x d matrices M_n (N * dmax^2 complex) in the plan's extra buffer
    Plan p(c, {&s}, {}, 0, (size_t)s.N * dmax * dmax * sizeof(double2));
    track_ws(A->owner, before);
    double2* Ms = (double2*)p.extra;
    double2* cur = (double2*)p.envA;
    double2* nxt = (double2*)p.envB;
    launch_set_scalar(cur, 1.0, 0.0, c->stream, c->bat);
    for (int n = s.N - 1; n >= 0; --n) {
      const i64 Dl = s.Dl(n), d = s.d[n], Dr = s.Dr(n);
      // Aenv = prod_LR(A[n], env)
      prod_LR_32(p.work, p.AL, s.site(n), cur, Dl, d, Dr, Dr, false);
      local_exp_matrix(p.work, Ms + (size_t)n * dmax * dmax, p.AL, s.site(n), Dl, d, Dr);
      // env = update_renv(A[n], Aenv)
      update_renv(p.work, nxt, s.site(n), p.AL, Dl, d, Dr, Dl);
      std::swap(cur, nxt);
    }
    std::vector<double2> h((size_t)s.N * dmax * dmax);
    HIPCHK(hipMemcpyAsync(h.data(), Ms, h.size() * sizeof(double2), hipMemcpyDeviceToHost, c->stream));
    c->sync();
    for (int n = 0; n < s.N; ++n) {
      const i64 d = s.d[n];
      const double2* M = h.data() + (size_t)n * dmax * dmax;
      cd t = 0.0;
      for (i64 a = 0; a < d; ++a)
        for (i64 b = 0; b < d; ++b)
          t += cd(op[2 * (a + d * b)], op[2 * (a + d * b) + 1]) * cd(M[b + d * a].x, M[b + d * a].y);
      out[n] = t.real();
    }
  });
}

// apply_site_operator! (MatrixProductStates.jl:135-143)
int mps_apply_site_operator(mps_state_t A, int site, const double* op) {
  return guard([&] {
    State& s = A->s;
    no_batch(s.ctx);
    REQUIRE(site >= 0 && site < s.N, "site out of range");
    Ctx* c = s.ctx;
    set_device(c);
    size_t before = c->ws_bytes;
    const i64 Dl = s.Dl(site), d = s.d[site], Dr = s.Dr(site);
    Plan p(c, {&s}, {}, 0, (size_t)d * d * sizeof(double2));
    track_ws(A->owner, before);
    HIPCHK(hipMemcpyAsync(p.extra, op, d * d * sizeof(double2), hipMemcpyHostToDevice, c->stream));
    // tmp[a, s, b] = sum_t A[a, t, b] op[s, t]   (batch over b)
    Gemm g;
    g.M = Dl;
    g.N = d;
    g.K = d;
    g.batch = Dr;
    g.A = s.site(site); g.am = im1(1); g.ak = im1(Dl); g.a_bs = Dl * d;
    g.B = p.extra; g.bk = im1(d); g.bn = im1(1);
    g.C = p.tmp; g.cm = im1(1); g.cn = im1(Dl); g.c_bs = Dl * d;
    run_gemm(c, g, CMP_F64, p.splitk, p.splitk_bytes);
    launch_copy_c(s.site(site), p.tmp, Dl * d * Dr, c->stream, c->bat);
    c->sync();
  });
}

// ----------------------------------------------------------------- Arnoldi
// Host dense helpers for the numv x numv Krylov matrices (the reference
// evaluates expm(N \ H) on the host too, timeevolution.jl:114).
static std::vector<cd> solve(std::vector<cd> A, std::vector<cd> B, int n) {
  // column-major A (n x n), B (n x n): returns A \ B (partial pivoting)
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(A[i + n * k]) > std::abs(A[piv + n * k])) piv = i;
    REQUIRE(std::abs(A[piv + n * k]) > 0.0, "singular Krylov overlap matrix");
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k + n * j], A[piv + n * j]);
      for (int j = 0; j < n; ++j) std::swap(B[k + n * j], B[piv + n * j]);
    }
    for (int i = k + 1; i < n; ++i) {
      const cd f = A[i + n * k] / A[k + n * k];
      for (int j = k; j < n; ++j) A[i + n * j] -= f * A[k + n * j];
      for (int j = 0; j < n; ++j) B[i + n * j] -= f * B[k + n * j];
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = n - 1; i >= 0; --i) {
      cd s = B[i + n * j];
      for (int k = i + 1; k < n; ++k) s -= A[i + n * k] * B[k + n * j];
      B[i + n * j] = s / A[i + n * i];
    }
  return B;
}

static std::vector<cd> matmul(const std::vector<cd>& A, const std::vector<cd>& B, int n) {
  std::vector<cd> C((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      for (int i = 0; i < n; ++i) C[i + n * j] += A[i + n * k] * B[k + n * j];
  return C;
}

// exp(M) by scaling and squaring of a Taylor series (truncation below 1e-17 relative).
static std::vector<cd> expm(std::vector<cd> M, int n) {
  double nrm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::abs(M[i + n * j]);
    nrm = std::max(nrm, s);
  }
  int sq = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    sq++;
  }
  const double f = std::ldexp(1.0, -sq);
  for (auto& x : M) x *= f;
  std::vector<cd> E((size_t)n * n, 0.0), T((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) E[i + n * i] = T[i + n * i] = 1.0;
  for (int k = 1; k <= 30; ++k) {
    T = matmul(T, M, n);
    for (auto& x : T) x /= (double)k;
    for (size_t i = 0; i < E.size(); ++i) E[i] += T[i];
  }
  for (int s = 0; s < sq; ++s) E = matmul(E, E, n);
  return E;
}

int mps_arnoldi_expm(mps_mpo_t A, int numv, const mps_state_t* rhov, mps_env_t /*env*/, mps_env_t envop,
                     double* coef_out) {
  return guard([&] {
    REQUIRE(numv >= 1 && rhov, "numv >= 1 Krylov states required");
    no_batch(A->o.ctx);
    for (int l = 0; l < numv; ++l) check_states({&rhov[0]->s, &rhov[l]->s});
    check_mpo(A->o, rhov[0]->s);
    check_op_env(envop->e, rhov[0]->s);
    for (int l = 0; l < numv; ++l) check_same_bonds({&rhov[0]->s, &rhov[l]->s});
    check_env_fits(envop->e, rhov[0]->s, rhov[0]->s, &A->o);
    std::vector<cd> N((size_t)numv * numv, 0.0), H((size_t)numv * numv, 0.0);
    for (int i = 0; i < numv; ++i) N[i + numv * i] = 1.0;
    auto sop = [&](mps_state_t a, mps_state_t b) {
      double o[2];
      int rc = mps_scal_op_prod(a, A, b, o);
      if (rc != MPS_OK) throw Error(rc, g_err);
      return cd(o[0], o[1]);
    };
    auto sp = [&](mps_state_t a, mps_state_t b) {
      double o[2];
      int rc = mps_scal_prod(a, b, o);
      if (rc != MPS_OK) throw Error(rc, g_err);
      return cd(o[0], o[1]);
    };
    H[0] = sop(rhov[0], rhov[0]);
    for (int j = 1; j < numv; ++j) {
      std::vector<double> cf(2 * (j + 1));
      std::vector<mps_state_t> vs(j + 1);
      for (int l = 0; l < j; ++l) {
        const cd v = -H[l + numv * (j - 1)];
        cf[2 * l] = v.real();
        cf[2 * l + 1] = v.imag();
        vs[l] = rhov[l];
      }
      cf[2 * j] = 1.0;
      cf[2 * j + 1] = 0.0;
      vs[j] = rhov[j - 1];
      int rc = mps_compress_sum_var_apply_H(rhov[j], envop, j + 1, cf.data(), vs.data(), A, 1);
      if (rc != MPS_OK) throw Error(rc, g_err);
      double no;
      rc = mps_normalize_lo(rhov[j], &no);
      if (rc != MPS_OK) throw Error(rc, g_err);
      H[j + numv * j] = sop(rhov[j], rhov[j]);
      for (int l = 0; l < j; ++l) {
        H[j + numv * l] = sop(rhov[j], rhov[l]);
        H[l + numv * j] = sop(rhov[l], rhov[j]);
        N[j + numv * l] = sp(rhov[j], rhov[l]);
        N[l + numv * j] = std::conj(N[j + numv * l]);
      }
    }
    std::vector<cd> E = expm(solve(N, H, numv), numv);
    std::vector<double> cf(2 * numv);
    for (int l = 0; l < numv; ++l) {
      cf[2 * l] = E[l].real();
      cf[2 * l + 1] = E[l].imag();
    }
    if (coef_out) std::memcpy(coef_out, cf.data(), cf.size() * sizeof(double));
    int rc = mps_compress_sum_var(rhov[0], rhov[0], numv, cf.data(), rhov, 1);
    if (rc != MPS_OK) throw Error(rc, g_err);
    double no;
    rc = mps_normalize_lo(rhov[0], &no);
    if (rc != MPS_OK) throw Error(rc, g_err);
  });
}

}  // extern "C"
