// Context, device containers and the L1 contraction chain of mps_base.jl.
#pragma once
#include "common.h"
#include <memory>

namespace mpship {

using i64 = int64_t;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int num_cus = 256;
  double qr_tol = 0.0;
  int qr_chain = -1;     // QR form: 1 single launch where it applies, 0 launch per panel, -1 process default
  int prec = 0;          // 0 FP64, 1 FP32 operator terms
  int refine = 0;        // FP64 half-sweeps refitting the last RK4 stage (even)
  bool graphs = true;
  // workspace arena (grow only, outside graph capture)
  char* ws = nullptr;
  size_t ws_bytes = 0;
  // Lockstep batch: a batched context holds bat.nb copies of every buffer it
  // allocates (states, MPOs, environments, workspace, device scalars) in one
  // slab, copy b at byte offset b * bat.bs, and runs every launch for all
  // copies.  Copy 0 is carved from [slab, slab + bat.bs) by a bump allocator.
  Bat bat;
  char* slab = nullptr;
  size_t slab_used = 0;
  // Ragged pool (mps_ctx_set_ragged): the copies hold chains of different
  // lengths n[b] <= nmax, sorted longest first, in one slot space of nmax
  // sites.  Copy b keeps its first h[b] = n[b]/2 sites in slots 0..h[b]-1 and
  // the rest right-aligned in slots nmax-(n[b]-h[b])..nmax-1; the slots in
  // between are its gap.  With the dims rule's edges at both ends of the slot
  // space, every copy present at a slot has that slot's dims, and the copies
  // present at slot q are the prefix 0..kact[q]-1.  The H-applying sweeps
  // launch each site's work for that prefix only and carry the gap copies'
  // environments across their gap.
  struct Ragged {
    int nmax = 0;
    std::vector<int> n, h, kact;
    int slot_of_site(int b, int j) const { return j < h[b] ? j : j + nmax - n[b]; }
    int slot_of_bond(int b, int i) const { return i <= h[b] ? i : i + nmax - n[b]; }
  };
  std::shared_ptr<const Ragged> rag;
  bool rag_ok = false;   // set by the entry points whose sweeps handle a ragged pool
  // device memory for a container: from the slab (owned = false) or hipMalloc
  void* alloc(size_t bytes, bool& owned);
  // second stream for GEMM chains independent of the main chain (forked
  // from and joined back into `stream` within one call; default priority)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void ensure_side();
  // third stream: builds the next RK4 stage's operator environments site by
  // site behind the current sweep (default priority)
  hipStream_t aux = nullptr;
  hipEvent_t ev_site = nullptr, ev_aux = nullptr;
  // workgroup cap of the aux stream's GEMMs: they run while the main stream
  // factors a site (a QR on a few CUs at a time) and must leave CUs to it
  int aux_max_wg = 0;
  bool aux_wg_env = false;   // MPS_AUX_MAX_WG set: its cap holds for every plan
  int aux_cap = 0;           // the cap of the current plan's aux-stream GEMMs (0: none)
  void ensure_aux();
  double* dscal = nullptr;      // 256 doubles of device scalars
  double* hscal = nullptr;      // pinned host mirror
  bool capturing = false;
  // live kernel timing (HIP events on this stream around every GEMM / QR launch)
  bool profiling = false;
  struct ProfRec {
    hipEvent_t a, b;
    int kind;          // 0 GEMM, 1 QR
    double flops;      // algorithmic flops of the launch (8 per complex MAC)
    double bytes;      // algorithmic HBM bytes (operands read once, result written once)
  };
  std::vector<ProfRec> prof;
  // GEMM recording for the autotuner (mps_ctx_gemm_record): every contraction
  // launched eagerly while on, with its operands and workspace as launched
  struct RecGemm {
    Gemm g;
    Compute cmp;
    void* ws;
    size_t ws_bytes;
    hipStream_t s;
  };
  bool gemm_rec = false;
  std::vector<RecGemm> rec;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  hipEvent_t next_event();
  void prof_begin(hipEvent_t& a);
  void prof_end(hipEvent_t a, int kind, double flops, double bytes);

  ~Ctx();
  void sync() { HIPCHK(hipStreamSynchronize(stream)); }
  void ensure_ws(size_t bytes);
};

// N sites, per-site capacity fixed at creation (dims may only shrink).
struct State {
  Ctx* ctx = nullptr;
  int N = 0;
  std::vector<i64> d, dims, cap_l, cap_r, off;
  double2* buf = nullptr;
  size_t total = 0;
  bool own = true;   // buf from hipMalloc (false: a batched context's slab, or a view)
  ~State();
  double2* site(int j) const { return buf + off[j]; }
  i64 Dl(int j) const { return dims[j]; }
  i64 Dr(int j) const { return dims[j + 1]; }
  i64 size(int j) const { return dims[j] * d[j] * dims[j + 1]; }
  i64 max_site() const;
  i64 max_dim() const;
};

struct MPO {
  Ctx* ctx = nullptr;
  int N = 0;
  std::vector<i64> d, w, off;   // w: N+1 bond dims
  double2* buf = nullptr;
  size_t total = 0;
  bool own = true;
  ~MPO();
  double2* site(int j) const { return buf + off[j]; }
  i64 wl(int j) const { return w[j]; }
  i64 wr(int j) const { return w[j + 1]; }
  i64 max_w() const;
};

struct Env {
  Ctx* ctx = nullptr;
  int N = 0;
  bool op = false;
  std::vector<i64> d1, wd, d2, off;   // N+1 slots
  double2* buf = nullptr;
  size_t total = 0;
  bool own = true;
  ~Env();
  double2* slot(int j) const { return buf + off[j]; }
};

// ----------------------------------------------------------- workspace
// Carves named scratch regions out of the context arena for one call.
struct Scratch {
  Ctx* c;
  size_t used = 0;
  std::vector<std::pair<void**, size_t>> req;
  explicit Scratch(Ctx* ctx) : c(ctx) {}
  void add(void** p, size_t bytes) { req.emplace_back(p, (bytes + 255) & ~size_t(255)); }
  void commit();
};

// compute options for one contraction
struct Op {
  Compute cmp = CMP_F64;
  DType tmp = F64;   // storage of T1/T2 intermediates
};

// on stream s (nullptr: the context's stream); splitk == nullptr: K is not split
void run_gemm(Ctx* c, Gemm& g, Compute cmp, void* splitk, size_t splitk_bytes, hipStream_t s = nullptr);

// --------------------------------------------- L1 contractions (mps_base.jl)
struct Work {
  Ctx* c;
  void* splitk = nullptr;
  size_t splitk_bytes = 0;
  void* T1 = nullptr;
  hipStream_t stream = nullptr;   // nullptr: the context's stream
};

void prod_AOR(Work& w, const Op& op, void* T2, const void* A, i64 Dl, i64 d, i64 Dr, const void* O,
              i64 wl, i64 wr, const void* R, i64 Dr0);
void prod_AOL(Work& w, const Op& op, void* AOL, const void* A, i64 Dl, i64 d, i64 Dr, const void* O,
              i64 wl, i64 wr, const void* L, i64 Dl0);
void prod_LR_34(Work& w, const Op& op, void* out, const void* L, const void* T2, i64 Dl, i64 wl, i64 Dl0,
                i64 d, i64 Dr0, bool accumulate);
void prod_LR_43(Work& w, const Op& op, void* out, const void* AOL, const void* R, i64 Dl0, i64 d, i64 Dr,
                i64 wr, i64 Dr0, bool accumulate);
void update_re(Work& w, const Op& op, void* out, const void* T2, const void* A0, i64 Dl, i64 wl, i64 d,
               i64 Dr0, i64 Dl0);
void update_le(Work& w, const Op& op, void* out, const void* AOL, const void* A0, i64 Dl0, i64 d, i64 Dr,
               i64 wr, i64 Dr0);
void prod_LR_23(Work& w, void* out, const void* L, const void* R, i64 l1, i64 r1, i64 r2, i64 r3,
                bool transposeL, bool accumulate);
void prod_LR_32(Work& w, void* out, const void* L, const void* R, i64 l1, i64 l2, i64 l3, i64 r2,
                bool accumulate, bool transposeR = false);
void update_renv(Work& w, void* out, const void* A0, const void* AR, i64 a1, i64 d, i64 a2, i64 ar1);
void update_lenv(Work& w, void* out, const void* A0, const void* AL, i64 a1, i64 d, i64 a3, i64 al3,
                 int mode);

}  // namespace mpship
