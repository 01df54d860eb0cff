// L1 contraction chain of src/mps_base.jl on the MFMA GEMM.
//
// Each function restates one reference contraction.  The reference builds
// every product as reshape -> permutedims (a copy) -> GEMM; here the
// permutations are folded into the GEMM's index maps, so each contraction
// is one or two GEMM launches reading the reference's own column-major
// layouts.  Layout derivations are in the comments (0-based, first index
// fastest).
#include "contract.h"

#include <algorithm>
#include <cstdlib>

namespace mpship {

hipEvent_t Ctx::next_event() {
  if (ev_used == ev_pool.size()) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    ev_pool.push_back(e);
  }
  return ev_pool[ev_used++];
}

void Ctx::prof_begin(hipEvent_t& a) {
  a = next_event();
  HIPCHK(hipEventRecord(a, stream));
}

void Ctx::prof_end(hipEvent_t a, int kind, double flops, double bytes) {
  hipEvent_t b = next_event();
  HIPCHK(hipEventRecord(b, stream));
  prof.push_back(ProfRec{a, b, kind, flops, bytes});
}

void Ctx::ensure_side() {
  if (side) return;
  REQUIRE(!capturing, "side stream creation during graph capture");
  HIPCHK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
}

void Ctx::ensure_aux() {
  if (aux) return;
  REQUIRE(!capturing, "aux stream creation during graph capture");
  HIPCHK(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
  // half the CUs' worth of workgroups where the sites' QR is a launch per
  // panel, which needs free CUs on every XCD (measured at N = 100, D = 128,
  // d = 9 with that QR: 229.6 ms per RK4 step at 128 on 256 CUs, 241
  // uncapped, 238.7 at 96, 232.6 at 160); Plan::init lifts the cap where the
  // sites' QR is the single launch (183.5 ms uncapped vs 200.0 at 128, over 20
  // steps after 5 warm-up steps); MPS_AUX_MAX_WG overrides both (0: no cap)
  aux_max_wg = std::max(1, num_cus / 2);
  if (const char* e = getenv("MPS_AUX_MAX_WG")) {
    aux_max_wg = std::max(0, atoi(e));
    aux_wg_env = true;
  }
  aux_cap = aux_max_wg;
  HIPCHK(hipEventCreateWithFlags(&ev_site, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&ev_aux, hipEventDisableTiming));
}

void* Ctx::alloc(size_t bytes, bool& owned) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (slab != nullptr) {
    REQUIRE(slab_used + bytes <= (size_t)bat.bs, "batched context: slab too small for this allocation");
    void* p = slab + slab_used;
    slab_used += bytes;
    owned = false;
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) throw Error(-3, "device allocation failed");
  owned = true;
  return p;
}

Ctx::~Ctx() {
  if (side) (void)hipStreamSynchronize(side);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (aux) (void)hipStreamSynchronize(aux);
  if (ev_site) (void)hipEventDestroy(ev_site);
  if (ev_aux) (void)hipEventDestroy(ev_aux);
  if (aux) (void)hipStreamDestroy(aux);
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  if (slab) {
    (void)hipStreamSynchronize(stream);
    (void)hipFree(slab);   // holds the workspace and the device scalars too
  } else {
    if (ws) (void)hipFree(ws);
    if (dscal) (void)hipFree(dscal);
  }
  if (hscal) (void)hipHostFree(hscal);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

void Ctx::ensure_ws(size_t bytes) {
  if (bytes <= ws_bytes) return;
  REQUIRE(slab == nullptr, "batched context: workspace reserved at creation is too small for this call");
  REQUIRE(!capturing, "workspace growth during graph capture");
  HIPCHK(hipStreamSynchronize(stream));
  if (side) HIPCHK(hipStreamSynchronize(side));
  if (aux) HIPCHK(hipStreamSynchronize(aux));
  if (ws) HIPCHK(hipFree(ws));
  ws = nullptr;
  size_t nb = bytes + bytes / 4 + (1 << 20);
  if (hipMalloc((void**)&ws, nb) != hipSuccess) {
    ws = nullptr;
    ws_bytes = 0;
    throw Error(-3, "workspace allocation failed");
  }
  ws_bytes = nb;
}

void Scratch::commit() {
  size_t tot = 0;
  for (auto& r : req) tot += r.second;
  c->ensure_ws(tot);
  size_t o = 0;
  for (auto& r : req) {
    *r.first = c->ws + o;
    o += r.second;
  }
  used = tot;
}

State::~State() {
  if (buf && own) (void)hipFree(buf);
}
MPO::~MPO() {
  if (buf && own) (void)hipFree(buf);
}
Env::~Env() {
  if (buf && own) (void)hipFree(buf);
}

i64 State::max_site() const {
  i64 m = 1;
  for (int j = 0; j < N; ++j) m = std::max(m, cap_l[j] * d[j] * cap_r[j]);
  return m;
}
i64 State::max_dim() const {
  i64 m = 1;
  for (int j = 0; j <= N; ++j) m = std::max(m, dims[j]);
  for (int j = 0; j < N; ++j) m = std::max(m, std::max(cap_l[j], cap_r[j]));
  return m;
}
i64 MPO::max_w() const {
  i64 m = 1;
  for (int j = 0; j <= N; ++j) m = std::max(m, w[j]);
  return m;
}

static double esize(DType t) { return t == F64 ? 16.0 : 8.0; }

void run_gemm(Ctx* c, Gemm& g, Compute cmp, void* splitk, size_t splitk_bytes, hipStream_t s) {
  g.bat = c->bat;
  g.tunable = true;
  if (s != nullptr && s == c->aux) g.max_wg = c->aux_cap;
  if (splitk == nullptr) g.max_splits = 1;   // no partials buffer: K is not split
  if (c->gemm_rec && !c->capturing) c->rec.push_back(Ctx::RecGemm{g, cmp, splitk, splitk_bytes, s ? s : c->stream});
  if (!c->profiling || c->capturing) {
    launch_gemm(g, cmp, splitk, splitk_bytes, c->num_cus, s ? s : c->stream);
    return;
  }
  REQUIRE(s == nullptr || s == c->stream, "profiled GEMMs run on the context's stream");
  hipEvent_t a;
  c->prof_begin(a);
  launch_gemm(g, cmp, splitk, splitk_bytes, c->num_cus, c->stream);
  // every lockstep copy the launch covers counts
  const double flops = 8.0 * (double)g.M * g.N * g.K * g.batch * g.bat.nb;
  const double bytes = (double)g.bat.nb * (double)g.batch * ((double)g.M * g.K * esize(g.ta) * (g.a_bs ? 1 : 0) +
                                          (double)g.K * g.N * esize(g.tb) * (g.b_bs ? 1 : 0) +
                                          (double)g.M * g.N * esize(g.tc)) +
                       (double)g.bat.nb * ((g.a_bs ? 0.0 : (double)g.M * g.K * esize(g.ta)) +
                                           (g.b_bs ? 0.0 : (double)g.K * g.N * esize(g.tb)));
  c->prof_end(a, 0, flops, bytes);
}

static inline Gemm mk(i64 M, i64 N, i64 K) {
  Gemm g;
  g.M = M;
  g.N = N;
  g.K = K;
  return g;
}

// prod_AOR (mps_base.jl:7-19)
//   T1[a,s,w',c] = sum_a3 A[a,s,a3] R[a3,w',c]                      (line 14)
//   T2[a,w,s',c] = sum_{s,w'} T1[a,s,w',c] O[w,s',w',s]             (lines 15-17)
// T2 is the reference's AOR[a1, wl, sigma_out, r3] layout.
void prod_AOR(Work& w, const Op& op, void* T2, const void* A, i64 Dl, i64 d, i64 Dr, const void* O,
              i64 wl, i64 wr, const void* R, i64 Dr0) {
  Gemm g1 = mk(Dl * d, wr * Dr0, Dr);
  g1.A = A; g1.am = im1(1); g1.ak = im1(Dl * d); g1.ta = F64;
  g1.B = R; g1.bk = im1(1); g1.bn = im1(Dr); g1.tb = F64;
  g1.C = w.T1; g1.cm = im1(1); g1.cn = im1(Dl * d); g1.tc = op.tmp;
  run_gemm(w.c, g1, op.cmp, w.splitk, w.splitk_bytes, w.stream);
  // batch over c; m = a; k = s + d*w'; n = w + wl*s'
  Gemm g2 = mk(Dl, wl * d, d * wr);
  g2.batch = Dr0;
  g2.A = w.T1; g2.am = im1(1); g2.ak = im1(Dl); g2.a_bs = Dl * d * wr; g2.ta = op.tmp;
  // O[w,s',w',s] at w + wl*s' + wl*d*w' + wl*d*wr*s
  g2.B = O; g2.bk = im2(d, wl * d * wr, wl * d); g2.bn = im1(1); g2.b_bs = 0; g2.tb = F64;
  g2.C = T2; g2.cm = im1(1); g2.cn = im1(Dl); g2.c_bs = Dl * wl * d; g2.tc = op.tmp;
  run_gemm(w.c, g2, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// prod_AOL (mps_base.jl:26-39)
//   T1[w,b,s,a'] = sum_a conj(L[a,w,b]) A[a,s,a']                   (line 33)
//   AOL[b,s',a',w'] = sum_{w,s} T1[w,b,s,a'] O[w,s',w',s]           (lines 34-37)
void prod_AOL(Work& w, const Op& op, void* AOL, const void* A, i64 Dl, i64 d, i64 Dr, const void* O,
              i64 wl, i64 wr, const void* L, i64 Dl0) {
  Gemm g1 = mk(wl * Dl0, d * Dr, Dl);
  g1.A = L; g1.am = im1(Dl); g1.ak = im1(1); g1.conjA = true; g1.ta = F64;
  g1.B = A; g1.bk = im1(1); g1.bn = im1(Dl); g1.tb = F64;
  g1.C = w.T1; g1.cm = im1(1); g1.cn = im1(wl * Dl0); g1.tc = op.tmp;
  run_gemm(w.c, g1, op.cmp, w.splitk, w.splitk_bytes, w.stream);
  // batch over a'; m = b; k = w + wl*s; n = s' + d*w'
  Gemm g2 = mk(Dl0, d * wr, wl * d);
  g2.batch = Dr;
  g2.A = w.T1; g2.am = im1(wl); g2.ak = im2(wl, 1, wl * Dl0); g2.a_bs = wl * Dl0 * d; g2.ta = op.tmp;
  // O[w,s',w',s] at w + wl*s' + wl*d*w' + wl*d*wr*s
  g2.B = O; g2.bk = im2(wl, 1, wl * d * wr); g2.bn = im1(wl); g2.tb = F64;
  // AOL[b,s',a',w'] at b + Dl0*s' + Dl0*d*a' + Dl0*d*Dr*w'
  g2.C = AOL; g2.cm = im1(1); g2.cn = im2(d, Dl0, Dl0 * d * Dr); g2.c_bs = Dl0 * d; g2.tc = op.tmp;
  run_gemm(w.c, g2, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// prod_LR! with 3-index L, 4-index R (mps_base.jl:53-61):
//   out[b,s',c] (+)= sum_{a,w} conj(L[a,w,b]) T2[a,w,s',c]
void prod_LR_34(Work& w, const Op& op, void* out, const void* L, const void* T2, i64 Dl, i64 wl, i64 Dl0,
                i64 d, i64 Dr0, bool accumulate) {
  Gemm g = mk(Dl0, d * Dr0, Dl * wl);
  g.A = L; g.am = im1(Dl * wl); g.ak = im1(1); g.conjA = true; g.ta = F64;
  g.B = T2; g.bk = im1(1); g.bn = im1(Dl * wl); g.tb = op.tmp;
  g.C = out; g.cm = im1(1); g.cn = im1(Dl0); g.tc = F64;
  if (accumulate) g.beta_re = 1.0;
  run_gemm(w.c, g, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// prod_LR! with 4-index L, 3-index R (mps_base.jl:73-81):
//   out[b,s',c] (+)= sum_{a',w'} AOL[b,s',a',w'] R[a',w',c]
void prod_LR_43(Work& w, const Op& op, void* out, const void* AOL, const void* R, i64 Dl0, i64 d, i64 Dr,
                i64 wr, i64 Dr0, bool accumulate) {
  Gemm g = mk(Dl0 * d, Dr0, Dr * wr);
  g.A = AOL; g.am = im1(1); g.ak = im1(Dl0 * d); g.ta = op.tmp;
  g.B = R; g.bk = im1(1); g.bn = im1(Dr * wr); g.tb = F64;
  g.C = out; g.cm = im1(1); g.cn = im1(Dl0 * d); g.tc = F64;
  if (accumulate) g.beta_re = 1.0;
  run_gemm(w.c, g, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// update_re! (mps_base.jl:151-160): R'[a,w,b] = sum_{s',c} T2[a,w,s',c] conj(A0[b,s',c])
void update_re(Work& w, const Op& op, void* out, const void* T2, const void* A0, i64 Dl, i64 wl, i64 d,
               i64 Dr0, i64 Dl0) {
  Gemm g = mk(Dl * wl, Dl0, d * Dr0);
  g.A = T2; g.am = im1(1); g.ak = im1(Dl * wl); g.ta = op.tmp;
  g.B = A0; g.bk = im1(Dl0); g.bn = im1(1); g.conjB = true; g.tb = F64;
  g.C = out; g.cm = im1(1); g.cn = im1(Dl * wl); g.tc = F64;
  run_gemm(w.c, g, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// update_le! (mps_base.jl:232-241): L'[a',w',c] = sum_{b,s'} conj(AOL[b,s',a',w']) A0[b,s',c]
void update_le(Work& w, const Op& op, void* out, const void* AOL, const void* A0, i64 Dl0, i64 d, i64 Dr,
               i64 wr, i64 Dr0) {
  Gemm g = mk(Dr * wr, Dr0, Dl0 * d);
  g.A = AOL; g.am = im1(Dl0 * d); g.ak = im1(1); g.conjA = true; g.ta = op.tmp;
  g.B = A0; g.bk = im1(1); g.bn = im1(Dl0 * d); g.tb = F64;
  g.C = out; g.cm = im1(1); g.cn = im1(Dr * wr); g.tc = F64;
  run_gemm(w.c, g, op.cmp, w.splitk, w.splitk_bytes, w.stream);
}

// prod_LR with 2-index L (mps_base.jl:85-118): out[l1,r2,r3] = L * R, or L^T * R (prod_LR_co)
void prod_LR_23(Work& w, void* out, const void* L, const void* R, i64 l1, i64 r1, i64 r2, i64 r3,
                bool transposeL, bool accumulate) {
  Gemm g = mk(l1, r2 * r3, r1);
  g.A = L;
  if (transposeL) {   // L stored r1 x l1
    g.am = im1(r1);
    g.ak = im1(1);
  } else {            // L stored l1 x r1
    g.am = im1(1);
    g.ak = im1(l1);
  }
  g.B = R; g.bk = im1(1); g.bn = im1(r1);
  g.C = out; g.cm = im1(1); g.cn = im1(l1);
  if (accumulate) g.beta_re = 1.0;
  run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes, w.stream);
}

// prod_LR with 3-index L, 2-index R (mps_base.jl:122-137): out[l1,l2,r2] = L[l1*l2, l3] * R[l3, r2]
// (transposeR: R stored r2 x l3, used as R^T -- the C.' of rightorth, orthonormalisation.jl:209)
void prod_LR_32(Work& w, void* out, const void* L, const void* R, i64 l1, i64 l2, i64 l3, i64 r2,
                bool accumulate, bool transposeR) {
  Gemm g = mk(l1 * l2, r2, l3);
  g.A = L; g.am = im1(1); g.ak = im1(l1 * l2);
  g.B = R;
  if (transposeR) {
    g.bk = im1(r2);
    g.bn = im1(1);
  } else {
    g.bk = im1(1);
    g.bn = im1(l3);
  }
  g.C = out; g.cm = im1(1); g.cn = im1(l1 * l2);
  if (accumulate) g.beta_re = 1.0;
  run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes, w.stream);
}

// update_renv! (mps_base.jl:193-199): out[a, c_l] = sum_{s,c} AR[a,s,c] conj(A0[c_l,s,c])
void update_renv(Work& w, void* out, const void* A0, const void* AR, i64 a1, i64 d, i64 a2, i64 ar1) {
  Gemm g = mk(ar1, a1, d * a2);
  g.A = AR; g.am = im1(1); g.ak = im1(ar1);
  g.B = A0; g.bk = im1(a1); g.bn = im1(1); g.conjB = true;
  g.C = out; g.cm = im1(1); g.cn = im1(ar1);
  run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes, w.stream);
}

// update_lenv! family (mps_base.jl:267-325), A0[a1,d,a3], AL[a1,d,al3]
//   mode 0 (update_lenv):    out[a3,al3] = A0^H AL
//   mode 1 (update_lenv_co): out[al3,a3] = AL^T conj(A0)
//   mode 2 (update_lenv_nc): out[a3,al3] = A0^T AL
void update_lenv(Work& w, void* out, const void* A0, const void* AL, i64 a1, i64 d, i64 a3, i64 al3,
                 int mode) {
  const i64 K = a1 * d;
  if (mode == 1) {
    Gemm g = mk(al3, a3, K);
    g.A = AL; g.am = im1(K); g.ak = im1(1);
    g.B = A0; g.bk = im1(1); g.bn = im1(K); g.conjB = true;
    g.C = out; g.cm = im1(1); g.cn = im1(al3);
    run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes, w.stream);
    return;
  }
  Gemm g = mk(a3, al3, K);
  g.A = A0; g.am = im1(K); g.ak = im1(1); g.conjA = (mode == 0);
  g.B = AL; g.bk = im1(1); g.bn = im1(K);
  g.C = out; g.cm = im1(1); g.cn = im1(a3);
  run_gemm(w.c, g, CMP_F64, w.splitk, w.splitk_bytes, w.stream);
}

}  // namespace mpship
