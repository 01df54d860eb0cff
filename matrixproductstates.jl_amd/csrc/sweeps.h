// L2/L3 algorithms on device-resident containers.
#pragma once
#include "contract.h"
#include <complex>
#include <initializer_list>

namespace mpship {

// Scratch carved for one high-level call, sized from the largest bond,
// local and MPO dimensions of the containers involved.
struct Plan {
  Ctx* ctx;
  Scratch scr;
  void *T1 = nullptr, *T2 = nullptr, *T3 = nullptr, *AL = nullptr, *tmp = nullptr;
  std::vector<void*> ovl;
  void *tau = nullptr, *R = nullptr, *envA = nullptr, *envB = nullptr;
  void* splitk = nullptr;
  size_t splitk_bytes = 0;
  void* extra = nullptr;   // caller-sized buffer (extra_bytes), for per-call arrays
  Work work{};
  Op op{};
  // Overlap-term chains (their environments and products) run on the
  // context's side stream with their own split-K partials, concurrently with
  // the operator chain on the main stream: fork() lets the side stream
  // continue from the main stream's current point, join() makes the main
  // stream wait for everything queued on the side stream.  Without overlap
  // terms, for small sites, or while kernels are being timed, `side` is
  // `work` and both are no-ops.
  Work side{};
  bool conc = false;
  void* splitk2 = nullptr;
  void fork();
  void join();
  // The aux stream (conc plans): RK4 stages 2-4 find their operator
  // environments built by it during the previous stage, into `shadow` or the
  // caller's envop alternately (see rk4_half).  `aux_w` has its own T1 and
  // split-K partials, `Ta` holds its products.
  Work aux_w{};
  void *Ta = nullptr, *ALa = nullptr, *splitk3 = nullptr;
  Env* shadow = nullptr;
  // RK4 refinement (ctx->refine > 0): a copy of psi taken before the 10-argument
  // form overwrites it, and the intermediate state of the refit sweeps
  State* rkeep = nullptr;
  State* rtmp = nullptr;
  Plan(Ctx* c, std::initializer_list<const State*> states, std::initializer_list<const MPO*> mpos,
       int n_ovl, size_t extra_bytes = 0);
  Plan(Ctx* c, i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes = 0) : ctx(c), scr(c) {
    init(Dm, dm, wm, n_ovl, extra_bytes);
  }
  void init(i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes);
};

// workspace bytes a plan of these sizes commits (per lockstep copy)
size_t plan_bytes(i64 Dm, i64 dm, i64 wm, int n_ovl, size_t extra_bytes, int num_cus, int nb);

void qr_left(Plan& p, State* s, int j, void* Rout);
void qr_right(Plan& p, State* s, int j, void* Rout);
void left_orthogonalize(Plan& p, State* s, int slot);
void right_orthogonalize(Plan& p, State* s, int slot);
double read_scalar(Ctx* c, int slot);

// The next sweep's operator-environment build, queued on the aux stream
// behind this sweep: after site j of this sweep's output is final, the
// environment step of the next sweep that reads it (update_re! for a right
// sweep, update_le! for a left one, with that sweep's A = Ain = the output).
struct NextInit {
  Env* e;                       // the next sweep's envop
  MPO* H;                       // the next sweep's operator
  std::complex<double> coef;    // its far-end boundary coefficient
  // overlap environments of the next sweep built the same way: oe[k]
  // between oA[k] and this sweep's output, far-end coefficient ocoef[k]
  int n_ovl = 0;
  Env* oe[3] = {nullptr, nullptr, nullptr};
  State* oA[3] = {nullptr, nullptr, nullptr};
  std::complex<double> ocoef[3] = {};
  // arithmetic of these builds (nullptr: the sweep's own, Plan::op)
  const Op* op = nullptr;
};

void compress_var_apply_H_left(Plan& p, State* A0, Env* envop, State* Ain, State* A, MPO* H,
                               const NextInit* nx = nullptr, bool init_done = false);
void compress_sum_var_apply_H_left(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int numv,
                                   const std::complex<double>* coef, State** A, MPO* H,
                                   const NextInit* nx = nullptr, bool init_done = false);
void compress_sum_var_apply_H_right(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int numv,
                                    const std::complex<double>* coef, State** A, MPO* H,
                                    const NextInit* nx = nullptr, bool init_done = false);
void compress_var_apply_H(Plan& p, State* A0, Env* envop, State* Ain, State* A, MPO* H, int maxsweeps);
void compress_var(Plan& p, State* A0, State* A, Env* env, int maxsweeps);
void compress_sum_var(Plan& p, State* out, State* Ain, int numv, const std::complex<double>* coef,
                      State** A, int maxsweeps, std::vector<Env*>& FL);
void compress_sum_var_apply_H(Plan& p, State* A0, Env* envop, int numv, const std::complex<double>* coef,
                              State** A, MPO* H, int maxsweeps, std::vector<Env*>& env);
// s1_done: stage 1's operator environments are already in envop (built
// behind the previous call's last sweep); next: build them for a following
// call with psi = out, the same H1 and envop, behind this call's last sweep
void rk4_half(Plan& p, double dt, MPO* H1, MPO* H2, MPO* H3, State* out, State* psi, State* psi1,
              State* psi2, State* psi3, Env* env1, Env* env2, Env* env3, Env* envop, bool s1_done = false,
              bool next = false);
// steady_state_search! (steadystate.jl:6-86) on the combined MPO W =
// applyMPOtoMPO(L, Lconj); p.extra holds steady_state_extra_bytes(max site, kmax)
size_t steady_state_extra_bytes(i64 site_elems, int kmax);
double steady_state_search(Plan& p, State* At, MPO* W, Env* envop, int maxsweeps, double converge_rat,
                           double eigs_tol, int kmax, int* sweeps_run);
// two-site sweeps with SVD truncation (compression.jl:110-164, 355-486,
// 574-647, 1072-1273, by intent); p.extra holds two_site_extra_bytes(Dm, dm),
// p.ovl 2 nov site buffers; mode 0 full sweeps, 1 one left, 2 one right
size_t two_site_extra_bytes(i64 Dm, i64 dm);
void two_site_sweeps(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int nov,
                     const std::complex<double>* coef, State** A, MPO* H, int mode, int maxsweeps,
                     double svd_tol, i64 dlim);
void scal_prod(Plan& p, State* A, State* B, int mode, int slot);
void scal_prod_ws(Work& w, State* A, State* B, int mode, void* envA, void* envB, void* AL, void* out);
void scal_op_prod(Plan& p, State* A, MPO* O, State* B, int slot);

}  // namespace mpship
