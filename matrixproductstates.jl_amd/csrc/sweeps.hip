// L2 gauge moves and L3 sweeps / time steppers on device-resident states.
//
// Each routine follows the reference's schedule line by line (cited per
// function); every contraction is a contract.hip GEMM launch, every QR the
// in-place Householder kernel, and boundary coefficients are written by a
// one-thread kernel, so a whole RK4 step is a fixed sequence of launches on
// the context stream (capturable as one hipGraph).
#include "contract.h"
#include "sweeps.h"
#include <cmath>
#include <complex>

namespace mpship {

using cd = std::complex<double>;

// -------------------------------------------------------------- planning
Plan::Plan(Ctx* c, std::initializer_list<const State*> states, std::initializer_list<const MPO*> mpos,
           int n_ovl, size_t extra_bytes)
    : ctx(c), scr(c) {
  i64 Dm = 1, dm = 1, wm = 1;
  for (const State* s : states)
    if (s) {
      Dm = std::max(Dm, s->max_dim());
      for (int j = 0; j < s->N; ++j) dm = std::max(dm, s->d[j]);
    }
  for (const MPO* o : mpos)
    if (o) wm = std::max(wm, o->max_w());
  init(Dm, dm, wm, n_ovl, extra_bytes);
}

// split-K partials per lockstep copy: splits * blocks < 4*CUs + blocks, and a
// GEMM splits only while blocks (all copies counted) < CUs
static size_t splitk_bound(int num_cus, int nb) {
  return (size_t)((5 * num_cus + nb - 1) / nb + 64) * 64 * 64 * sizeof(double2);
}

// the regions of a plan, in order: add(slot, bytes)
template <typename F>
static void plan_layout(Plan* p, i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes, int num_cus, int nb,
                        bool conc, F&& add) {
  const size_t cz = sizeof(double2);
  const size_t site = (size_t)Dm * dm * Dm;
  const size_t big = site * wm;
  add(p ? &p->T1 : nullptr, big * cz);
  add(p ? &p->T2 : nullptr, big * cz);
  add(p ? &p->T3 : nullptr, big * cz);
  add(p ? &p->AL : nullptr, site * cz);
  add(p ? &p->tmp : nullptr, site * cz);
  for (int l = 0; l < std::max(n_ovl, 0); ++l) add(p ? &p->ovl[l] : nullptr, site * cz);
  add(p ? &p->tau : nullptr, qr_workspace_bytes(Dm * dm, Dm, num_cus));
  add(p ? &p->R : nullptr, (size_t)Dm * Dm * dm * cz);
  add(p ? &p->envA : nullptr, big * cz);
  add(p ? &p->envB : nullptr, big * cz);
  add(p ? &p->splitk : nullptr, splitk_bound(num_cus, nb));
  if (conc) add(p ? &p->splitk2 : nullptr, splitk_bound(num_cus, nb));
  if (conc) add(p ? &p->splitk3 : nullptr, splitk_bound(num_cus, nb));
  if (conc) add(p ? &p->Ta : nullptr, big * cz);
  if (conc) add(p ? &p->ALa : nullptr, site * cz);
  if (extra_bytes) add(p ? &p->extra : nullptr, extra_bytes);
}

size_t plan_bytes(i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes, int num_cus, int nb) {
  size_t tot = 0;
  plan_layout(nullptr, Dm, dm, wm, n_ovl, extra_bytes, num_cus, nb, false,
              [&](void**, size_t b) { tot += (b + 255) & ~size_t(255); });
  return tot;
}

void Plan::init(i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes) {
  Ctx* c = ctx;
  if (c->rag && !c->rag_ok)
    throw Error(-4, "not available on a ragged pool (RK4 and the H-applying half-sweeps are; "
                    "other calls go through a view of one copy)");
  const size_t site = (size_t)Dm * dm * Dm;
  ovl.assign(std::max(n_ovl, 0), nullptr);
  // the side stream pays off once a site's products fill a good part of the
  // GPU (measured: D = 128, d = 9 gains; bond-16/32 chains served many to a
  // GPU lose, their streams then outnumber the hardware queues)
  conc = n_ovl > 0 && !c->profiling && site >= 32768 && c->bat.nb == 1;
  if (conc) {
    c->ensure_side();
    c->ensure_aux();
    // the aux stream's GEMMs run beside the sites' QRs: capped where those
    // are a launch per panel, uncapped beside the single-launch QR
    c->aux_cap = (!c->aux_wg_env && qr_single_launch_applies(Dm * dm, Dm, c->bat, c->qr_chain)) ? 0 : c->aux_max_wg;
  }
  plan_layout(this, Dm, dm, wm, n_ovl, extra_bytes, c->num_cus, c->bat.nb, conc,
              [&](void** slot, size_t b) { scr.add(slot, b); });
  splitk_bytes = splitk_bound(c->num_cus, c->bat.nb);
  scr.commit();
  work.c = c;
  work.splitk = splitk;
  work.splitk_bytes = splitk_bytes;
  work.T1 = T1;
  side = work;
  if (conc) {
    side.splitk = splitk2;
    side.T1 = nullptr;   // the overlap terms need no T1
    side.stream = c->side;
    aux_w = work;
    aux_w.splitk = splitk3;
    aux_w.T1 = T3;
    aux_w.stream = c->aux;
  }
  op.cmp = (c->prec == 1) ? CMP_F32_3M : CMP_F64;
  op.tmp = (c->prec == 1) ? F32 : F64;
}

// ------------------------------------------------------------ gauge moves
// leftorth! (orthonormalisation.jl:164-171): QR of reshape(A, Dl*d, Dr)
static void run_qr(Plan& p, void* base, i64 m, i64 n, i64 rs, i64 cs, void* Rout) {
  Ctx* c = p.ctx;
  const bool prof = c->profiling && !c->capturing;
  hipEvent_t a = nullptr;
  if (prof) c->prof_begin(a);
  launch_qr(base, m, n, rs, cs, c->qr_tol, p.tau, c->num_cus, Rout, c->stream, c->bat, c->qr_chain);
  // zgeqrf + zungqr: 2 x 8 (m n^2 - n^3 / 3) flops; panel read and written once
  if (prof) c->prof_end(a, 1, 16.0 * ((double)m * n * n - (double)n * n * n / 3.0), 32.0 * (double)m * n);
}

void qr_left(Plan& p, State* s, int j, void* Rout) {
  const i64 Dl = s->Dl(j), d = s->d[j], Dr = s->Dr(j);
  REQUIRE(Dl * d >= Dr, "leftorth!: Dl*d < Dr");
  run_qr(p, s->site(j), Dl * d, Dr, 1, Dl * d, Rout);
}

// rightorth! (orthonormalisation.jl:227-234): QR of the plain transpose of
// reshape(A, Dl, d*Dr); element (r, c) of the transpose is A[c + Dl*r], and
// Q written through the same view is reshape(Q.', Dl, d, Dr).
void qr_right(Plan& p, State* s, int j, void* Rout) {
  const i64 Dl = s->Dl(j), d = s->d[j], Dr = s->Dr(j);
  REQUIRE(d * Dr >= Dl, "rightorth!: d*Dr < Dl");
  if (Dl * d * Dr < 32768) {
    run_qr(p, s->site(j), d * Dr, Dl, Dl, 1, Rout);
    return;
  }
  // Large sites: the transposed view reads and writes whole rows a column
  // apart (one cache line per element in the QR's column passes), so the
  // panel is transposed into a column-major copy, factored there, and Q
  // transposed back (the same arithmetic on the same values).
  Ctx* c = p.ctx;
  launch_transpose_c(p.tmp, s->site(j), Dl, d * Dr, c->stream, c->bat);
  run_qr(p, p.tmp, d * Dr, Dl, 1, d * Dr, Rout);
  launch_transpose_c(s->site(j), p.tmp, d * Dr, Dl, c->stream, c->bat);
}

// leftorthogonalizeQR! (orthonormalisation.jl:43-56); norm^2 left in dscal[slot]
void left_orthogonalize(Plan& p, State* s, int slot) {
  for (int i = 0; i + 1 < s->N; ++i) {
    qr_left(p, s, i, p.R);
    // mps[i+1] = prod_LR(r, mps[i+1]) : r is Dr x Dr
    const i64 Dr = s->Dr(i), d1 = s->d[i + 1], Dr1 = s->Dr(i + 1);
    prod_LR_23(p.work, p.tmp, p.R, s->site(i + 1), Dr, Dr, d1, Dr1, false, false);
    launch_copy_c(s->site(i + 1), p.tmp, Dr * d1 * Dr1, p.ctx->stream, p.ctx->bat);
  }
  const int n = s->N - 1;
  launch_sum_abs2(s->site(n), s->size(n), p.ctx->dscal + slot, p.ctx->stream, p.ctx->bat);
}

// rightorthogonalizeQR! (orthonormalisation.jl:77-90)
void right_orthogonalize(Plan& p, State* s, int slot) {
  for (int i = s->N - 1; i >= 1; --i) {
    qr_right(p, s, i, p.R);
    // mps[i-1] = prod_LR(mps[i-1], C.') with C = R (Dl x Dl) -> multiply by R^T
    const i64 Dl = s->Dl(i), Dlm = s->Dl(i - 1), dm = s->d[i - 1];
    prod_LR_32(p.work, p.tmp, s->site(i - 1), p.R, Dlm, dm, Dl, Dl, false, true);
    launch_copy_c(s->site(i - 1), p.tmp, Dlm * dm * Dl, p.ctx->stream, p.ctx->bat);
  }
  launch_sum_abs2(s->site(0), s->size(0), p.ctx->dscal + slot, p.ctx->stream, p.ctx->bat);
}

double read_scalar(Ctx* c, int slot) {
  HIPCHK(hipMemcpyAsync(c->hscal, c->dscal + slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  return c->hscal[0];
}

void Plan::fork() {
  if (!conc) return;
  HIPCHK(hipEventRecord(ctx->ev_fork, ctx->stream));
  HIPCHK(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
}

void Plan::join() {
  if (!conc) return;
  HIPCHK(hipEventRecord(ctx->ev_join, ctx->side));
  HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
}

// ---------------------------------------------------------------- helpers
static void set1(Ctx* c, double2* p, cd v) { launch_set_scalar(p, v.real(), v.imag(), c->stream, c->bat); }
static void set1(const Work& w, double2* p, cd v) {
  launch_set_scalar(p, v.real(), v.imag(), w.stream ? w.stream : w.c->stream, w.c->bat);
}

// --------------------------------------------------------- ragged pools
// Inside a sweep's step at slot j, the context launches for the copies
// present there (the prefix 0..kact[j]-1); outside it, for all copies.
struct SiteScope {
  Ctx* c;
  int nb;
  SiteScope(Ctx* ctx, int j) : c(ctx), nb(ctx->bat.nb) {
    if (c->rag) c->bat.nb = c->rag->kact[j];
  }
  ~SiteScope() { c->bat.nb = nb; }
  SiteScope(const SiteScope&) = delete;
  SiteScope& operator=(const SiteScope&) = delete;
};

// The copies in their gap at slot j (kact[j]..nb-1) carry the environments a
// sweep is updating one slot on: from slot j to j+1 in a sweep to the right,
// from j+1 to j in a sweep to the left.  Across its whole gap a copy thus
// hands the environment of its fold bond from one side of the gap to the
// other, where its next site reads it.
static void gap_pass(Plan& p, int j, bool to_right, Env* op, Env* const* ovl, int n_ovl) {
  Ctx* c = p.ctx;
  if (!c->rag) return;
  const int k = c->rag->kact[j], nb = (int)c->rag->n.size();
  if (k >= nb) return;
  const int64_t o = (int64_t)k * c->bat.bs;
  Bat g;
  g.nb = nb - k;
  g.bs = c->bat.bs;
  const int src = to_right ? j : j + 1, dst = to_right ? j + 1 : j;
  auto pass = [&](Env* e) {
    const i64 n = e->d1[src] * e->wd[src] * e->d2[src];
    REQUIRE(n == e->d1[dst] * e->wd[dst] * e->d2[dst], "ragged pool: environment slots differ inside a gap");
    launch_copy_c(boff(e->slot(dst), o), boff(e->slot(src), o), n, c->stream, g);
  };
  if (op) pass(op);
  for (int l = 0; l < n_ovl; ++l) pass(ovl[l]);
}

static void check_same_dims(const State* a, const State* b, const char* what) {
  REQUIRE(a->N == b->N, std::string(what) + ": chain lengths differ");
  for (int j = 0; j < a->N; ++j) REQUIRE(a->d[j] == b->d[j], std::string(what) + ": local dims differ");
}

static void check_env(const Env* e, const State* s, bool op, const char* what) {
  REQUIRE(e->op == op, std::string(what) + ": wrong environment kind");
  REQUIRE(e->N == s->N, std::string(what) + ": environment length mismatch");
}

// ------------------------------------------------------------ L3 sweeps
// op-env left initialisation shared by the _left sweeps:
//   envop[1][1] = 1; update_le!(envop[jj+1], Ain[jj], H[jj], A[jj], envop[jj])
static void init_left_op(Plan& p, Env* envop, State* Ain, State* A, MPO* H, cd right_coef) {
  Ctx* c = p.ctx;
  const int N = Ain->N;
  set1(c, envop->slot(0), 1.0);
  for (int j = 0; j + 1 < N; ++j) {
    SiteScope ss(c, j);
    gap_pass(p, j, true, envop, nullptr, 0);
    // AOL = prod_AOL(A[j], H[j], envop[j]); envop[j+1] = update_le(Ain[j], AOL)
    prod_AOL(p.work, p.op, p.T2, A->site(j), A->Dl(j), A->d[j], A->Dr(j), H->site(j), H->wl(j),
             H->wr(j), envop->slot(j), Ain->Dl(j));
    update_le(p.work, p.op, envop->slot(j + 1), p.T2, Ain->site(j), Ain->Dl(j), Ain->d[j], A->Dr(j),
              H->wr(j), Ain->Dr(j));
  }
  set1(c, envop->slot(N), right_coef);
}

// op-env right initialisation shared by the _right sweeps:
//   envop[na+1][1] = 1; update_re!(envop[jj], Ain[jj], H[jj], A[jj], envop[jj+1])
static void init_right_op(Plan& p, Env* envop, State* Ain, State* A, MPO* H, cd left_coef) {
  Ctx* c = p.ctx;
  const int N = Ain->N;
  set1(c, envop->slot(N), 1.0);
  for (int j = N - 1; j >= 1; --j) {
    SiteScope ss(c, j);
    gap_pass(p, j, false, envop, nullptr, 0);
    prod_AOR(p.work, p.op, p.T2, A->site(j), A->Dl(j), A->d[j], A->Dr(j), H->site(j), H->wl(j),
             H->wr(j), envop->slot(j + 1), Ain->Dr(j));
    update_re(p.work, p.op, envop->slot(j), p.T2, Ain->site(j), A->Dl(j), H->wl(j), A->d[j],
              Ain->Dr(j), Ain->Dl(j));
  }
  set1(c, envop->slot(0), left_coef);
}

// ------------------------------------- next sweep's environments (aux stream)
// The aux stream continues from the main stream's current point / the main
// stream waits for everything queued on the aux stream.
static void aux_follow(Plan& p) {
  Ctx* c = p.ctx;
  HIPCHK(hipEventRecord(c->ev_site, c->stream));
  HIPCHK(hipStreamWaitEvent(c->aux, c->ev_site, 0));
}
static void aux_wait(Plan& p) {
  Ctx* c = p.ctx;
  HIPCHK(hipEventRecord(c->ev_aux, c->aux));
  HIPCHK(hipStreamWaitEvent(c->stream, c->ev_aux, 0));
}

// After a left sweep's QR of site j (A0[j] final): init_right_op's step for
// site j of the next (right) sweep, whose A and Ain are this sweep's A0.
static void nx_right_begin(Plan& p, const NextInit* nx, int N) {
  if (!nx) return;
  aux_follow(p);   // nx's environments are no longer in use by earlier work on the main stream
  set1(p.aux_w, nx->e->slot(N), 1.0);
  for (int k = 0; k < nx->n_ovl; ++k) set1(p.aux_w, nx->oe[k]->slot(N), 1.0);
}
static void nx_right_step(Plan& p, const NextInit* nx, State* A0, int j) {
  if (!nx) return;
  aux_follow(p);
  Env* e = nx->e;
  MPO* H = nx->H;
  const Op& op = nx->op ? *nx->op : p.op;
  prod_AOR(p.aux_w, op, p.Ta, A0->site(j), A0->Dl(j), A0->d[j], A0->Dr(j), H->site(j), H->wl(j), H->wr(j),
           e->slot(j + 1), A0->Dr(j));
  update_re(p.aux_w, op, e->slot(j), p.Ta, A0->site(j), A0->Dl(j), H->wl(j), A0->d[j], A0->Dr(j),
            A0->Dl(j));
  // update_renv!(oe[j], A0[j], oA[j], oe[j+1]), as the next sweep's own build
  for (int k = 0; k < nx->n_ovl; ++k) {
    State* B = nx->oA[k];
    prod_LR_32(p.aux_w, p.ALa, B->site(j), nx->oe[k]->slot(j + 1), B->Dl(j), B->d[j], B->Dr(j), A0->Dr(j),
               false);
    update_renv(p.aux_w, nx->oe[k]->slot(j), A0->site(j), p.ALa, A0->Dl(j), A0->d[j], A0->Dr(j), B->Dl(j));
  }
}
static void nx_right_end(Plan& p, const NextInit* nx) {
  if (!nx) return;
  set1(p.aux_w, nx->e->slot(0), nx->coef);
  for (int k = 0; k < nx->n_ovl; ++k) set1(p.aux_w, nx->oe[k]->slot(0), nx->ocoef[k]);
}
// After a right sweep's QR of site j: init_left_op's step for site j of the
// next (left) sweep.
static void nx_left_begin(Plan& p, const NextInit* nx) {
  if (!nx) return;
  aux_follow(p);
  set1(p.aux_w, nx->e->slot(0), 1.0);
  for (int k = 0; k < nx->n_ovl; ++k) set1(p.aux_w, nx->oe[k]->slot(0), 1.0);
}
static void nx_left_step(Plan& p, const NextInit* nx, State* A0, int j) {
  if (!nx) return;
  aux_follow(p);
  Env* e = nx->e;
  MPO* H = nx->H;
  const Op& op = nx->op ? *nx->op : p.op;
  prod_AOL(p.aux_w, op, p.Ta, A0->site(j), A0->Dl(j), A0->d[j], A0->Dr(j), H->site(j), H->wl(j), H->wr(j),
           e->slot(j), A0->Dl(j));
  update_le(p.aux_w, op, e->slot(j + 1), p.Ta, A0->site(j), A0->Dl(j), A0->d[j], A0->Dr(j), H->wr(j),
            A0->Dr(j));
  // update_lenv!(oe[j+1], A0[j], oA[j], oe[j]), as the next sweep's own build
  for (int k = 0; k < nx->n_ovl; ++k) {
    State* B = nx->oA[k];
    prod_LR_23(p.aux_w, p.ALa, nx->oe[k]->slot(j), B->site(j), A0->Dl(j), B->Dl(j), B->d[j], B->Dr(j), false,
               false);
    update_lenv(p.aux_w, nx->oe[k]->slot(j + 1), A0->site(j), p.ALa, A0->Dl(j), A0->d[j], A0->Dr(j), B->Dr(j),
                0);
  }
}
static void nx_left_end(Plan& p, const NextInit* nx, int N) {
  if (!nx) return;
  set1(p.aux_w, nx->e->slot(N), nx->coef);
  for (int k = 0; k < nx->n_ovl; ++k) set1(p.aux_w, nx->oe[k]->slot(N), nx->ocoef[k]);
}

// The next sweep's environment step for a site is queued one site late:
// behind the products of the following site, so the aux stream's GEMMs run
// while the main stream factors that site (a latency-bound QR on a few CUs)
// instead of competing with the main stream's products.
struct NxQueue {
  Plan& p;
  const NextInit* nx;
  State* A0;
  bool left;
  int pend = -1;
  void push(int j) {
    flush();
    pend = j;
  }
  void flush() {
    if (pend < 0 || !nx) return;
    SiteScope ss(p.ctx, pend);
    if (left)
      nx_left_step(p, nx, A0, pend);
    else
      nx_right_step(p, nx, A0, pend);
    pend = -1;
  }
};

// compress_var_apply_H_left! (compression.jl:321-347)
void compress_var_apply_H_left(Plan& p, State* A0, Env* envop, State* Ain, State* A, MPO* H,
                               const NextInit* nx, bool init_done) {
  const int N = Ain->N;
  if (init_done)
    aux_wait(p);
  else
    init_left_op(p, envop, Ain, A, H, 1.0);
  nx_right_begin(p, nx, N);
  NxQueue nq{p, nx, A0, false};
  for (int j = N - 1; j >= 0; --j) {
    SiteScope ss(p.ctx, j);
    gap_pass(p, j, false, envop, nullptr, 0);
    // AOR = prod_AOR(A[jj], H[jj], envop[jj+1]); prod_LR!(A0[jj], envop[jj], AOR)
    prod_AOR(p.work, p.op, p.T2, A->site(j), A->Dl(j), A->d[j], A->Dr(j), H->site(j), H->wl(j),
             H->wr(j), envop->slot(j + 1), A0->Dr(j));
    prod_LR_34(p.work, p.op, A0->site(j), envop->slot(j), p.T2, A->Dl(j), H->wl(j), A0->Dl(j), A->d[j],
               A0->Dr(j), false);
    nq.flush();
    if (j == 0) break;
    qr_right(p, A0, j, nullptr);
    nq.push(j);
    update_re(p.work, p.op, envop->slot(j), p.T2, A0->site(j), A->Dl(j), H->wl(j), A->d[j], A0->Dr(j),
              A0->Dl(j));
  }
  nq.flush();
  nx_right_end(p, nx);
}

// compress_sum_var_apply_H_left! (compression.jl:889-944).  The overlap
// terms (environment builds, per-site products, environment updates) run on
// the side stream; the operator term joins them before accumulating into A0.
void compress_sum_var_apply_H_left(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int numv,
                                   const cd* coef, State** A, MPO* H, const NextInit* nx, bool init_done) {
  const int N = Ain->N;
  Work& ow = p.side;
  if (init_done) aux_wait(p);   // environments built on the aux stream during the previous sweep
  p.fork();
  for (int ll = 0; ll + 1 < numv && !init_done; ++ll) {
    Env* e = env[ll];
    set1(ow, e->slot(0), 1.0);
    for (int j = 0; j + 1 < N; ++j) {
      SiteScope ss(p.ctx, j);
      gap_pass(p, j, true, nullptr, &e, 1);
      // update_lenv!(env[ll][jj+1], Ain[jj], A[ll][jj], env[ll][jj])
      prod_LR_23(ow, p.AL, e->slot(j), A[ll]->site(j), Ain->Dl(j), A[ll]->Dl(j), A[ll]->d[j],
                 A[ll]->Dr(j), false, false);
      update_lenv(ow, e->slot(j + 1), Ain->site(j), p.AL, Ain->Dl(j), Ain->d[j], Ain->Dr(j),
                  A[ll]->Dr(j), 0);
    }
    set1(ow, e->slot(N), coef[ll]);
  }
  State* Ah = A[numv - 1];
  if (!init_done) init_left_op(p, envop, Ain, Ah, H, coef[numv - 1]);
  nx_right_begin(p, nx, N);
  NxQueue nq{p, nx, A0, false};
  for (int j = N - 1; j >= 0; --j) {
    SiteScope ss(p.ctx, j);
    gap_pass(p, j, false, envop, env, numv - 1);
    const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
    // overlap terms: AR = prod_LR(A[ll][jj], env[ll][jj+1]); A0 (+)= prod_LR(env[ll][jj], AR)
    for (int ll = 0; ll + 1 < numv; ++ll) {
      prod_LR_32(ow, p.ovl[ll], A[ll]->site(j), env[ll]->slot(j + 1), A[ll]->Dl(j), d, A[ll]->Dr(j),
                 Dr0, false);
      prod_LR_23(ow, A0->site(j), env[ll]->slot(j), p.ovl[ll], Dl0, A[ll]->Dl(j), d, Dr0, false,
                 ll > 0);
    }
    // operator term: AOR = prod_AOR(A[numv][jj], H[jj], envop[jj+1]); A0 += prod_LR(envop[jj], AOR)
    prod_AOR(p.work, p.op, p.T2, Ah->site(j), Ah->Dl(j), d, Ah->Dr(j), H->site(j), H->wl(j), H->wr(j),
             envop->slot(j + 1), Dr0);
    p.join();
    prod_LR_34(p.work, p.op, A0->site(j), envop->slot(j), p.T2, Ah->Dl(j), H->wl(j), Dl0, d, Dr0,
               numv > 1);
    nq.flush();
    if (j == 0) break;
    qr_right(p, A0, j, nullptr);
    p.fork();
    // the operator environment first: the main stream's next products wait on it
    update_re(p.work, p.op, envop->slot(j), p.T2, A0->site(j), Ah->Dl(j), H->wl(j), d, Dr0, Dl0);
    for (int ll = 0; ll + 1 < numv; ++ll)
      update_renv(ow, env[ll]->slot(j), A0->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dl(j));
    nq.push(j);
  }
  nq.flush();
  nx_right_end(p, nx);
  p.join();
}

// compress_sum_var_apply_H_right! (compression.jl:1007-1063).  A0 may alias
// A[0] (both touched only by the side stream's overlap chain, in order).
void compress_sum_var_apply_H_right(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int numv,
                                    const cd* coef, State** A, MPO* H, const NextInit* nx, bool init_done) {
  const int N = Ain->N;
  Work& ow = p.side;
  if (init_done) aux_wait(p);   // environments built on the aux stream during the previous sweep
  p.fork();
  for (int ll = 0; ll + 1 < numv && !init_done; ++ll) {
    Env* e = env[ll];
    set1(ow, e->slot(N), 1.0);
    for (int j = N - 1; j >= 1; --j) {
      SiteScope ss(p.ctx, j);
      gap_pass(p, j, false, nullptr, &e, 1);
      // update_renv!(env[ll][jj], Ain[jj], A[ll][jj], env[ll][jj+1])
      prod_LR_32(ow, p.AL, A[ll]->site(j), e->slot(j + 1), A[ll]->Dl(j), A[ll]->d[j], A[ll]->Dr(j),
                 Ain->Dr(j), false);
      update_renv(ow, e->slot(j), Ain->site(j), p.AL, Ain->Dl(j), Ain->d[j], Ain->Dr(j),
                  A[ll]->Dl(j));
    }
    set1(ow, e->slot(0), coef[ll]);
  }
  State* Ah = A[numv - 1];
  if (!init_done) init_right_op(p, envop, Ain, Ah, H, coef[numv - 1]);
  nx_left_begin(p, nx);
  NxQueue nq{p, nx, A0, true};
  for (int j = 0; j < N; ++j) {
    SiteScope ss(p.ctx, j);
    gap_pass(p, j, true, envop, env, numv - 1);
    const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
    // overlap terms first: AL[ll] reads A[ll][jj] before A0[jj] is written
    for (int ll = 0; ll + 1 < numv; ++ll)
      prod_LR_23(ow, p.ovl[ll], env[ll]->slot(j), A[ll]->site(j), Dl0, A[ll]->Dl(j), d, A[ll]->Dr(j),
                 false, false);
    for (int ll = 0; ll + 1 < numv; ++ll)
      prod_LR_32(ow, A0->site(j), p.ovl[ll], env[ll]->slot(j + 1), Dl0, d, A[ll]->Dr(j), Dr0, ll > 0);
    // operator term: AOL = prod_AOL(A[numv][jj], H[jj], envop[jj]); A0 += prod_LR(AOL, envop[jj+1])
    prod_AOL(p.work, p.op, p.T2, Ah->site(j), Ah->Dl(j), d, Ah->Dr(j), H->site(j), H->wl(j), H->wr(j),
             envop->slot(j), Dl0);
    p.join();
    prod_LR_43(p.work, p.op, A0->site(j), p.T2, envop->slot(j + 1), Dl0, d, Ah->Dr(j), H->wr(j), Dr0,
               numv > 1);
    nq.flush();
    if (j == N - 1) break;
    qr_left(p, A0, j, nullptr);
    p.fork();
    // the operator environment first: the main stream's next products wait on it
    update_le(p.work, p.op, envop->slot(j + 1), p.T2, A0->site(j), Dl0, d, Ah->Dr(j), H->wr(j), Dr0);
    for (int ll = 0; ll + 1 < numv; ++ll)
      update_lenv(ow, env[ll]->slot(j + 1), A0->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dr(j), 0);
    nq.push(j);
  }
  nq.flush();
  nx_left_end(p, nx, N);
  p.join();
}

// compress_var_apply_H! (compression.jl:217-258)
void compress_var_apply_H(Plan& p, State* A0, Env* envop, State* Ain, State* A, MPO* H, int maxsweeps) {
  const int N = Ain->N;
  init_left_op(p, envop, Ain, A, H, 1.0);
  for (int sw = 0; sw < maxsweeps; ++sw) {
    for (int j = N - 1; j >= 1; --j) {
      prod_AOR(p.work, p.op, p.T2, A->site(j), A->Dl(j), A->d[j], A->Dr(j), H->site(j), H->wl(j),
               H->wr(j), envop->slot(j + 1), A0->Dr(j));
      prod_LR_34(p.work, p.op, A0->site(j), envop->slot(j), p.T2, A->Dl(j), H->wl(j), A0->Dl(j), A->d[j],
                 A0->Dr(j), false);
      qr_right(p, A0, j, nullptr);
      update_re(p.work, p.op, envop->slot(j), p.T2, A0->site(j), A->Dl(j), H->wl(j), A->d[j],
                A0->Dr(j), A0->Dl(j));
    }
    for (int j = 0; j < N; ++j) {
      prod_AOL(p.work, p.op, p.T2, A->site(j), A->Dl(j), A->d[j], A->Dr(j), H->site(j), H->wl(j),
               H->wr(j), envop->slot(j), A0->Dl(j));
      prod_LR_43(p.work, p.op, A0->site(j), p.T2, envop->slot(j + 1), A0->Dl(j), A->d[j], A->Dr(j),
                 H->wr(j), A0->Dr(j), false);
      if (j == N - 1) break;
      qr_left(p, A0, j, nullptr);
      update_le(p.work, p.op, envop->slot(j + 1), p.T2, A0->site(j), A0->Dl(j), A->d[j], A->Dr(j),
                H->wr(j), A0->Dr(j));
    }
  }
}

// compress_var! (compression.jl:65-103); env holds L[ket,bra] on the left
// (update_lenv_co!) and R[ket,bra] on the right.
void compress_var(Plan& p, State* A0, State* A, Env* env, int maxsweeps) {
  Ctx* c = p.ctx;
  const int N = A0->N;
  set1(c, env->slot(0), 1.0);
  for (int j = 0; j + 1 < N; ++j) {
    // update_lenv_co!(env[jj+1], A0[jj], A[jj], env[jj]): AL = prod_LR_co(env[jj], A[jj])
    prod_LR_23(p.work, p.AL, env->slot(j), A->site(j), A0->Dl(j), A->Dl(j), A->d[j], A->Dr(j), true,
               false);
    update_lenv(p.work, env->slot(j + 1), A0->site(j), p.AL, A0->Dl(j), A0->d[j], A0->Dr(j), A->Dr(j), 1);
  }
  set1(c, env->slot(N), 1.0);
  for (int sw = 0; sw < maxsweeps; ++sw) {
    for (int j = N - 1; j >= 0; --j) {
      const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
      // AR = prod_LR(A[jj], env[jj+1]); prod_LR_co!(A0[jj], env[jj], AR)  (jj=1: prod_LR!)
      prod_LR_32(p.work, p.AL, A->site(j), env->slot(j + 1), A->Dl(j), d, A->Dr(j), Dr0, false);
      prod_LR_23(p.work, A0->site(j), env->slot(j), p.AL, Dl0, A->Dl(j), d, Dr0, j > 0, false);
      if (j == 0) break;
      qr_right(p, A0, j, nullptr);
      update_renv(p.work, env->slot(j), A0->site(j), p.AL, Dl0, d, Dr0, A->Dl(j));
    }
    for (int j = 0; j < N; ++j) {
      const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
      // AL = prod_LR_co(env[jj], A[jj]); prod_LR!(A0[jj], AL, env[jj+1])
      prod_LR_23(p.work, p.AL, env->slot(j), A->site(j), Dl0, A->Dl(j), d, A->Dr(j), true, false);
      prod_LR_32(p.work, A0->site(j), p.AL, env->slot(j + 1), Dl0, d, A->Dr(j), Dr0, false);
      if (j == N - 1) break;
      qr_left(p, A0, j, nullptr);
      update_lenv(p.work, env->slot(j + 1), A0->site(j), p.AL, Dl0, d, Dr0, A->Dr(j), 1);
    }
  }
}

// compress_sum_var (compression.jl:496-565).  out must not alias any A[ll]
// (the reference allocates a fresh A0); Ain may alias out.
void compress_sum_var(Plan& p, State* out, State* Ain, int numv, const cd* coef, State** A,
                      int maxsweeps, std::vector<Env*>& FL) {
  Ctx* c = p.ctx;
  const int N = Ain->N;
  // FL[ll] slots 0..N hold the left envs (slot j for site j) and, from the
  // right sweep on, the right envs FR[ll][j] in slot j+1 -- as the reference's
  // FL/FR pair, but sharing storage the way the !-variants do.
  for (int ll = 0; ll < numv; ++ll) {
    Env* e = FL[ll];
    set1(c, e->slot(0), 1.0);
    for (int j = 0; j + 1 < N; ++j) {
      prod_LR_23(p.work, p.AL, e->slot(j), A[ll]->site(j), Ain->Dl(j), A[ll]->Dl(j), A[ll]->d[j],
                 A[ll]->Dr(j), false, false);
      update_lenv(p.work, e->slot(j + 1), Ain->site(j), p.AL, Ain->Dl(j), Ain->d[j], Ain->Dr(j),
                  A[ll]->Dr(j), 0);
    }
    set1(c, e->slot(N), coef[ll]);
  }
  // Ain (if distinct) is only read during initialisation; copy it into out.
  if (out != Ain)
    for (int j = 0; j < N; ++j) launch_copy_c(out->site(j), Ain->site(j), Ain->size(j), c->stream, c->bat);
  for (int sw = 0; sw < maxsweeps; ++sw) {
    for (int j = N - 1; j >= 1; --j) {
      const i64 Dl0 = out->Dl(j), Dr0 = out->Dr(j), d = out->d[j];
      for (int ll = 0; ll < numv; ++ll) {
        prod_LR_32(p.work, p.ovl[ll], A[ll]->site(j), FL[ll]->slot(j + 1), A[ll]->Dl(j), d, A[ll]->Dr(j),
                   Dr0, false);
        prod_LR_23(p.work, out->site(j), FL[ll]->slot(j), p.ovl[ll], Dl0, A[ll]->Dl(j), d, Dr0, false,
                   ll > 0);
      }
      qr_right(p, out, j, nullptr);
      for (int ll = 0; ll < numv; ++ll)
        update_renv(p.work, FL[ll]->slot(j), out->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dl(j));
    }
    for (int j = 0; j + 1 < N; ++j) {
      const i64 Dl0 = out->Dl(j), Dr0 = out->Dr(j), d = out->d[j];
      for (int ll = 0; ll < numv; ++ll) {
        prod_LR_23(p.work, p.ovl[ll], FL[ll]->slot(j), A[ll]->site(j), Dl0, A[ll]->Dl(j), d, A[ll]->Dr(j),
                   false, false);
        prod_LR_32(p.work, out->site(j), p.ovl[ll], FL[ll]->slot(j + 1), Dl0, d, A[ll]->Dr(j), Dr0,
                   ll > 0);
      }
      qr_left(p, out, j, nullptr);
      for (int ll = 0; ll < numv; ++ll)
        update_lenv(p.work, FL[ll]->slot(j + 1), out->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dr(j), 0);
    }
  }
  const int j = N - 1;
  const i64 Dl0 = out->Dl(j), Dr0 = out->Dr(j), d = out->d[j];
  for (int ll = 0; ll < numv; ++ll) {
    prod_LR_23(p.work, p.ovl[ll], FL[ll]->slot(j), A[ll]->site(j), Dl0, A[ll]->Dl(j), d, A[ll]->Dr(j),
               false, false);
    prod_LR_32(p.work, out->site(j), p.ovl[ll], FL[ll]->slot(j + 1), Dl0, d, A[ll]->Dr(j), Dr0, ll > 0);
  }
}

// compress_sum_var_apply_H! (compression.jl:805-887), by intent: overlap envs
// start at 1 on the left and coef[ll] on the right (as in _left!).
void compress_sum_var_apply_H(Plan& p, State* A0, Env* envop, int numv, const cd* coef, State** A,
                              MPO* H, int maxsweeps, std::vector<Env*>& env) {
  Ctx* c = p.ctx;
  const int N = A0->N;
  for (int ll = 0; ll + 1 < numv; ++ll) {
    Env* e = env[ll];
    set1(c, e->slot(0), 1.0);
    for (int j = 0; j + 1 < N; ++j) {
      prod_LR_23(p.work, p.AL, e->slot(j), A[ll]->site(j), A0->Dl(j), A[ll]->Dl(j), A[ll]->d[j],
                 A[ll]->Dr(j), false, false);
      update_lenv(p.work, e->slot(j + 1), A0->site(j), p.AL, A0->Dl(j), A0->d[j], A0->Dr(j), A[ll]->Dr(j),
                  0);
    }
    set1(c, e->slot(N), coef[ll]);
  }
  State* Ah = A[numv - 1];
  init_left_op(p, envop, A0, Ah, H, coef[numv - 1]);
  for (int sw = 0; sw < maxsweeps; ++sw) {
    for (int j = N - 1; j >= 1; --j) {
      const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
      for (int ll = 0; ll + 1 < numv; ++ll) {
        prod_LR_32(p.work, p.ovl[ll], A[ll]->site(j), env[ll]->slot(j + 1), A[ll]->Dl(j), d,
                   A[ll]->Dr(j), Dr0, false);
        prod_LR_23(p.work, A0->site(j), env[ll]->slot(j), p.ovl[ll], Dl0, A[ll]->Dl(j), d, Dr0, false,
                   ll > 0);
      }
      prod_AOR(p.work, p.op, p.T2, Ah->site(j), Ah->Dl(j), d, Ah->Dr(j), H->site(j), H->wl(j), H->wr(j),
               envop->slot(j + 1), Dr0);
      prod_LR_34(p.work, p.op, A0->site(j), envop->slot(j), p.T2, Ah->Dl(j), H->wl(j), Dl0, d, Dr0,
                 numv > 1);
      qr_right(p, A0, j, nullptr);
      for (int ll = 0; ll + 1 < numv; ++ll)
        update_renv(p.work, env[ll]->slot(j), A0->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dl(j));
      update_re(p.work, p.op, envop->slot(j), p.T2, A0->site(j), Ah->Dl(j), H->wl(j), d, Dr0, Dl0);
    }
    for (int j = 0; j + 1 < N; ++j) {
      const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
      for (int ll = 0; ll + 1 < numv; ++ll)
        prod_LR_23(p.work, p.ovl[ll], env[ll]->slot(j), A[ll]->site(j), Dl0, A[ll]->Dl(j), d, A[ll]->Dr(j),
                   false, false);
      for (int ll = 0; ll + 1 < numv; ++ll)
        prod_LR_32(p.work, A0->site(j), p.ovl[ll], env[ll]->slot(j + 1), Dl0, d, A[ll]->Dr(j), Dr0,
                   ll > 0);
      prod_AOL(p.work, p.op, p.T2, Ah->site(j), Ah->Dl(j), d, Ah->Dr(j), H->site(j), H->wl(j), H->wr(j),
               envop->slot(j), Dl0);
      prod_LR_43(p.work, p.op, A0->site(j), p.T2, envop->slot(j + 1), Dl0, d, Ah->Dr(j), H->wr(j), Dr0,
                 numv > 1);
      qr_left(p, A0, j, nullptr);
      for (int ll = 0; ll + 1 < numv; ++ll)
        update_lenv(p.work, env[ll]->slot(j + 1), A0->site(j), p.ovl[ll], Dl0, d, Dr0, A[ll]->Dr(j), 0);
      update_le(p.work, p.op, envop->slot(j + 1), p.T2, A0->site(j), Dl0, d, Ah->Dr(j), H->wr(j), Dr0);
    }
  }
  const int j = N - 1;
  const i64 Dl0 = A0->Dl(j), Dr0 = A0->Dr(j), d = A0->d[j];
  for (int ll = 0; ll + 1 < numv; ++ll)
    prod_LR_23(p.work, p.ovl[ll], env[ll]->slot(j), A[ll]->site(j), Dl0, A[ll]->Dl(j), d, A[ll]->Dr(j),
               false, false);
  for (int ll = 0; ll + 1 < numv; ++ll)
    prod_LR_32(p.work, A0->site(j), p.ovl[ll], env[ll]->slot(j + 1), Dl0, d, A[ll]->Dr(j), Dr0, ll > 0);
  prod_AOL(p.work, p.op, p.T2, Ah->site(j), Ah->Dl(j), d, Ah->Dr(j), H->site(j), H->wl(j), H->wr(j),
           envop->slot(j), Dl0);
  prod_LR_43(p.work, p.op, A0->site(j), p.T2, envop->slot(j + 1), Dl0, d, Ah->Dr(j), H->wr(j), Dr0,
             numv > 1);
}

// RK4stp_apply_H_half! (timeevolution.jl:178-196)
//
// With the aux stream (p.shadow set): stages 2-4 do not build their operator
// environments between sweeps; the aux stream builds each one behind the
// previous sweep, site by site as that sweep's output becomes final, into
// the shadow environment (stages 2, 4) or envop (stage 3).  The steps and
// their order per environment are those of init_{left,right}_op.
//
// Refinement (ctx->refine = n > 0, even): after stage 4, n FP64 half-sweeps
// (left, right, ...) refit stage 4's target -1/3 psi + 1/3 psi1 + 2/3 psi2 +
// 1/3 H3 psi3, each started from the previous one's result, exactly as a
// caller would by calling compress_sum_var_apply_H_left! / _right!
// (compression.jl:889-944, 1007-1063) again with stage 4's arguments; the
// last (right) sweep leaves `out` left canonical as stage 4 does.  The
// 10-argument form overwrites psi in stage 4, so psi is first copied into
// p.rkeep, which then stands for psi in stage 4 and the refit (the same
// values: stage 4's arithmetic is unchanged).
void rk4_half(Plan& p, double dt, MPO* H1, MPO* H2, MPO* H3, State* out, State* psi, State* psi1,
              State* psi2, State* psi3, Env* env1, Env* env2, Env* env3, Env* envop, bool s1_done, bool next) {
  const bool ahead = p.conc && p.shadow != nullptr;
  const int nref = p.ctx->refine;
  State* psi_t = psi;   // psi as the first term of stage 4 (and of the refit)
  if (nref > 0) {
    REQUIRE(nref % 2 == 0 && p.rtmp != nullptr && (out != psi || p.rkeep != nullptr),
            "RK4 refinement: an even number of half-sweeps and their states");
    if (out == psi) {
      p.rkeep->dims = psi->dims;
      for (int j = 0; j < psi->N; ++j)
        launch_copy_c(p.rkeep->site(j), psi->site(j), psi->size(j), p.ctx->stream, p.ctx->bat);
      psi_t = p.rkeep;
    }
    p.rtmp->dims = psi->dims;
  }
  Env* E2 = ahead ? p.shadow : envop;   // stages 2 and 4
  const cd c2[2] = {1.0, dt / 2}, c3[2] = {1.0, dt}, c4[4] = {-1.0 / 3, 1.0 / 3, 2.0 / 3, 1.0 / 3};
  // Overlap environments: stage 2's env1 (psi vs psi1) is built behind stage
  // 1; stage 3's env1 (psi vs psi2, coefficient 1 at both ends) is what
  // stage 2's own environment updates leave in env1, operation for
  // operation; stage 4's env1 (psi vs psi3) is what stage 3 leaves, up to the
  // coefficient in slot 0; env2 and env3 are built behind stage 3.
  NextInit n2{E2, H2, c2[1]}, n3{envop, H2, c3[1]}, n4{E2, H3, c4[3]};
  n2.n_ovl = 1;
  n2.oe[0] = env1;
  n2.oA[0] = psi;
  n2.ocoef[0] = c2[0];
  n4.n_ovl = 2;
  n4.oe[0] = env2;
  n4.oA[0] = psi1;
  n4.ocoef[0] = c4[1];
  n4.oe[1] = env3;
  n4.oA[1] = psi2;
  n4.ocoef[1] = c4[2];
  compress_var_apply_H_left(p, psi1, envop, psi, psi, H1, ahead ? &n2 : nullptr, ahead && s1_done);
  {
    Env* e[1] = {env1};
    State* A[2] = {psi, psi1};
    compress_sum_var_apply_H_right(p, psi2, e, E2, psi1, 2, c2, A, H2, ahead ? &n3 : nullptr, ahead);
  }
  {
    Env* e[1] = {env1};
    State* A[2] = {psi, psi2};
    compress_sum_var_apply_H_left(p, psi3, e, envop, psi2, 2, c3, A, H2, ahead ? &n4 : nullptr, ahead);
  }
  // a following call's stage 1 (init_left_op on out, H1, far-end
  // coefficient 1, the context's precision) into envop, which stage 4 and
  // the refit leave alone; queued behind the call's last sweep
  const Op ctx_op = p.op;
  NextInit n1{envop, H1, 1.0};
  n1.op = &ctx_op;
  {
    Env* e[3] = {env1, env2, env3};
    State* A[4] = {psi_t, psi1, psi2, psi3};
    if (ahead) set1(p.ctx, env1->slot(0), c4[0]);
    compress_sum_var_apply_H_right(p, out, e, E2, psi3, 4, c4, A, H3,
                                   (ahead && next && nref == 0) ? &n1 : nullptr, ahead);
  }
  if (nref > 0) {
    p.op = Op{CMP_F64, F64};
    Env* e[3] = {env1, env2, env3};
    State* A[4] = {psi_t, psi1, psi2, psi3};
    for (int k = 0; k < nref; ++k) {
      if (k % 2 == 0) {
        compress_sum_var_apply_H_left(p, p.rtmp, e, E2, out, 4, c4, A, H3);
      } else {
        const bool last = k == nref - 1;
        compress_sum_var_apply_H_right(p, out, e, E2, p.rtmp, 4, c4, A, H3,
                                       (ahead && next && last) ? &n1 : nullptr, false);
      }
    }
    p.op = ctx_op;
  }
  if (ahead && next) aux_wait(p);   // the call's work is complete on the context's stream
}

// ------------------------------------------------------ scalar products
// scal_prod / scal_prod_no_conj (MatrixProductStates.jl:345-379) -> dscal[slot..slot+1]
void scal_prod(Plan& p, State* A, State* B, int mode, int slot) {
  scal_prod_ws(p.work, A, B, mode, p.envA, p.envB, p.AL, p.ctx->dscal + slot);
}

// The same chain on caller-provided buffers: envA / envB hold the running
// environment (at most max_j Dl_A(j) Dl_B(j) complex), AL the per-site
// product (at most max_j Dl_A(j) d_j Dr_B(j) complex); the scalar goes to out.
void scal_prod_ws(Work& w, State* A, State* B, int mode, void* envA, void* envB, void* AL, void* out) {
  Ctx* c = w.c;
  double2* cur = (double2*)envA;
  double2* nxt = (double2*)envB;
  set1(w, cur, 1.0);
  for (int j = 0; j < A->N; ++j) {
    // AL = prod_LR(lenv, B[j]); lenv = update_lenv(A[j], AL)
    prod_LR_23(w, AL, cur, B->site(j), A->Dl(j), B->Dl(j), B->d[j], B->Dr(j), false, false);
    update_lenv(w, nxt, A->site(j), AL, A->Dl(j), A->d[j], A->Dr(j), B->Dr(j), mode);
    std::swap(cur, nxt);
  }
  launch_copy_c(out, cur, 1, w.stream ? w.stream : c->stream, c->bat);
}

// scal_op_prod (MatrixProductStates.jl:327-340): conj(<A|O|B>) -> dscal[slot..]
void scal_op_prod(Plan& p, State* A, MPO* O, State* B, int slot) {
  Ctx* c = p.ctx;
  double2* cur = (double2*)p.envA;
  double2* nxt = (double2*)p.envB;
  set1(c, cur, 1.0);
  for (int j = 0; j < A->N; ++j) {
    // lenv = update_le(A[j], O[j], B[j], lenv)
    prod_AOL(p.work, p.op, p.T2, B->site(j), B->Dl(j), B->d[j], B->Dr(j), O->site(j), O->wl(j),
             O->wr(j), cur, A->Dl(j));
    update_le(p.work, p.op, nxt, p.T2, A->site(j), A->Dl(j), A->d[j], B->Dr(j), O->wr(j), A->Dr(j));
    std::swap(cur, nxt);
  }
  launch_copy_c(c->dscal + slot, cur, 1, c->stream, c->bat);
}

}  // namespace mpship
