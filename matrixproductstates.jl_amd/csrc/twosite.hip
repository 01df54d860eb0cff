// Two-site variational compression with SVD truncation: compress_var_two_site,
// compress_var_apply_H_two_site{,_left,_right}, compress_sum_var_two_site and
// compress_sum_var_apply_H_two_site{,_left,_right} (compression.jl:110-164,
// 355-486, 574-647, 1072-1273).  The reference's versions call svdecon, which
// the package never defines, so they are built by intent: svdecon is the thin
// SVD M = U diag(S) V^H, and the sweeps are those of the reference.
//
// The target is sum_l c_l A_l (overlap terms) plus c H A_last (the operator
// term, when H is given).  Per bond (j, j+1) the projected two-site tensor
// T[b, s1, s2, f] is, for an overlap term, prod_LR(L_l[j], A_l[j]) times
// prod_LR(A_l[j+1], R_l[j+2]) and, for the operator term, prod_AOL(A[j], H[j],
// E[j]) times prod_AOR(A[j+1], H[j+1], E[j+2]) (A_prod_LR_two_site /
// A_prod_LOR_two_site, MatrixProductStates.jl:168-173, 292-295), one GEMM
// each, accumulated.  T (Dl d x d Dr) is factored by the one-sided Jacobi
// SVD of svd.hip; the new bond keeps min(#{k: cumsum(s^2)/sum s^2 < 1 - svd_tol}
// + 1, dlim) values; a left sweep keeps V^H on site j+1 (right isometry), a
// right sweep U on site j (left isometry), and the last bond of a half-sweep
// keeps U S (left) or S V^H (right) on its outer site.  Environments are
// updated from the new site with the products already formed, and the
// coefficients enter through the far-end boundary slots as in the one-site
// sum sweeps (compression.jl:903-913, 1021-1031).
#include "contract.h"
#include "sweeps.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <numeric>
#include <vector>

namespace mpship {

using cd = std::complex<double>;

size_t two_site_extra_bytes(i64 Dm, i64 dm) {
  const size_t mn = (size_t)(Dm * dm) * (Dm * dm);
  const size_t nmax = (size_t)Dm * dm + 64;
  return 3 * mn * sizeof(double2) + 3 * nmax * sizeof(double) + nmax * sizeof(int) + 4096;
}

namespace {

struct SvdBufs {
  double2 *T, *Y, *J;
  double *sig, *scl;
  int* perm;
  int* flag;
};

struct Factor {
  bool ct;          // Y = T^H (columns = m) instead of Y = T
  i64 m, n, rows, cols;
  int k;            // kept singular values
  std::vector<int> order;
  std::vector<double> sig;
};

// thin SVD of T (m x n) and the truncation rule
Factor factor(Ctx* c, const SvdBufs& b, i64 m, i64 n, double svd_tol, i64 dlim, i64 cap) {
  Factor f;
  f.m = m;
  f.n = n;
  f.ct = n >= m;
  f.rows = f.ct ? n : m;
  f.cols = f.ct ? m : n;
  svd_load(b.Y, b.T, f.rows, f.cols, f.ct, c->stream);
  svd_jacobi(b.Y, f.rows, f.cols, b.J, 4.0 * std::sqrt((double)f.rows) * 2.220446049250313e-16, b.flag,
             (int*)c->hscal, 100, c->stream);
  svd_colnorms(b.Y, f.rows, f.cols, b.sig, c->stream);
  f.sig.assign(f.cols, 0.0);
  HIPCHK(hipMemcpyAsync(f.sig.data(), b.sig, f.cols * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  c->sync();
  for (int k = 0; k < f.cols; ++k)
    if (f.sig[k] > 0.0) f.order.push_back(k);
  std::stable_sort(f.order.begin(), f.order.end(), [&](int a, int bb) { return f.sig[a] > f.sig[bb]; });
  REQUIRE(!f.order.empty(), "two-site sweep: zero two-site tensor");
  double tot = 0.0;
  for (int k : f.order) tot += f.sig[k] * f.sig[k];
  i64 cnt = 0;
  double cum = 0.0;
  for (int k : f.order) {
    cum += f.sig[k] * f.sig[k] / tot;
    if (cum < 1.0 - svd_tol) ++cnt;
  }
  f.k = (int)std::min<i64>(std::min<i64>(cnt + 1, dlim), std::min<i64>((i64)f.order.size(), cap));
  return f;
}

// out (m x k, column major) = U[:, 1:k] (scaled by S when withS)
void gather_U(Ctx* c, const SvdBufs& b, const Factor& f, void* out, bool withS) {
  std::vector<double> scl(f.k);
  for (int k = 0; k < f.k; ++k) {
    const double s = f.sig[f.order[k]];
    scl[k] = f.ct ? (withS ? s : 1.0) : (withS ? 1.0 : 1.0 / s);
  }
  HIPCHK(hipMemcpyAsync(b.perm, f.order.data(), f.k * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(b.scl, scl.data(), f.k * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // ct: U = J (m x m); otherwise U = W Sigma^-1 with W = Y (m x n)
  svd_gather(out, f.ct ? b.J : b.Y, f.ct ? f.cols : f.rows, b.perm, b.scl, f.k, f.m, false, true, c->stream);
  c->sync();   // host vectors
}

// out (k x n, column major) = V[:, 1:k]^H (scaled row by row by S when withS)
void gather_Vh(Ctx* c, const SvdBufs& b, const Factor& f, void* out, bool withS) {
  std::vector<double> scl(f.k);
  for (int k = 0; k < f.k; ++k) {
    const double s = f.sig[f.order[k]];
    scl[k] = f.ct ? (withS ? 1.0 : 1.0 / s) : (withS ? s : 1.0);
  }
  HIPCHK(hipMemcpyAsync(b.perm, f.order.data(), f.k * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(b.scl, scl.data(), f.k * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // ct: V = W Sigma^-1 with W = Y (n x m); otherwise V = J (n x n)
  svd_gather(out, f.ct ? b.Y : b.J, f.ct ? f.rows : f.cols, b.perm, b.scl, f.k, f.n, true, false, c->stream);
  c->sync();
}

struct Sweep {
  Plan& p;
  State* A0;
  Env** env;
  Env* envop;
  int nov;          // overlap terms
  State** A;        // nov overlap states, then the operator term's state
  const cd* coef;
  MPO* H;           // nullptr: no operator term
  SvdBufs b;
  Ctx* c() const { return p.ctx; }

  // T = two-site projection at bond (j, j+1); AL/AR per overlap term into
  // p.ovl, X / Y of the operator term into p.T2 / p.T3
  void project(int j) {
    const i64 Dl0 = A0->Dl(j), d1 = A0->d[j], d2 = A0->d[j + 1], Dr0 = A0->Dr(j + 1);
    bool first = true;
    for (int l = 0; l < nov; ++l) {
      State* S = A[l];
      void* AL = p.ovl[2 * l];
      void* AR = p.ovl[2 * l + 1];
      prod_LR_23(p.work, AL, env[l]->slot(j), S->site(j), Dl0, S->Dl(j), d1, S->Dr(j), false, false);
      prod_LR_32(p.work, AR, S->site(j + 1), env[l]->slot(j + 2), S->Dl(j + 1), d2, S->Dr(j + 1), Dr0, false);
      Gemm g;
      g.M = Dl0 * d1; g.N = d2 * Dr0; g.K = S->Dr(j);
      g.A = AL; g.am = im1(1); g.ak = im1(Dl0 * d1);
      g.B = AR; g.bk = im1(1); g.bn = im1(S->Dl(j + 1));
      g.C = b.T; g.cm = im1(1); g.cn = im1(Dl0 * d1);
      if (!first) g.beta_re = 1.0;
      run_gemm(c(), g, CMP_F64, p.splitk, p.splitk_bytes);
      first = false;
    }
    if (H) {
      State* S = A[nov];
      prod_AOL(p.work, p.op, p.T2, S->site(j), S->Dl(j), d1, S->Dr(j), H->site(j), H->wl(j), H->wr(j),
               envop->slot(j), Dl0);
      Work w3 = p.work;
      w3.T1 = p.T1;
      prod_AOR(w3, p.op, p.T3, S->site(j + 1), S->Dl(j + 1), d2, S->Dr(j + 1), H->site(j + 1), H->wl(j + 1),
               H->wr(j + 1), envop->slot(j + 2), Dr0);
      // T += X[b s1, c w] Y[c w, s2 f]
      Gemm g;
      g.M = Dl0 * d1; g.N = d2 * Dr0; g.K = S->Dr(j) * H->wr(j);
      g.A = p.T2; g.am = im1(1); g.ak = im1(Dl0 * d1); g.ta = p.op.tmp;
      g.B = p.T3; g.bk = im1(1); g.bn = im1(S->Dr(j) * H->wr(j)); g.tb = p.op.tmp;
      g.C = b.T; g.cm = im1(1); g.cn = im1(Dl0 * d1);
      if (!first) g.beta_re = 1.0;
      run_gemm(c(), g, p.op.cmp, p.splitk, p.splitk_bytes);
    }
  }

  // left sweep: bonds N-2 .. 0; the last keeps U S on site 0
  void left(double svd_tol, i64 dlim) {
    const int N = A0->N;
    for (int j = N - 2; j >= 0; --j) {
      project(j);
      const i64 Dl0 = A0->Dl(j), d1 = A0->d[j], d2 = A0->d[j + 1], Dr0 = A0->Dr(j + 1);
      Factor f = factor(c(), b, Dl0 * d1, d2 * Dr0, svd_tol, dlim, std::min(A0->cap_r[j], A0->cap_l[j + 1]));
      A0->dims[j + 1] = f.k;
      gather_Vh(c(), b, f, A0->site(j + 1), false);
      // right environments at slot j+1 from the new site j+1
      for (int l = 0; l < nov; ++l)
        update_renv(p.work, env[l]->slot(j + 1), A0->site(j + 1), p.ovl[2 * l + 1], f.k, d2, Dr0, A[l]->Dl(j + 1));
      if (H)
        update_re(p.work, p.op, envop->slot(j + 1), p.T3, A0->site(j + 1), A[nov]->Dl(j + 1), H->wl(j + 1), d2,
                  Dr0, f.k);
      if (j == 0) gather_U(c(), b, f, A0->site(0), true);
    }
  }

  // right sweep: bonds 0 .. N-2; the last keeps S V^H on site N-1
  void right(double svd_tol, i64 dlim) {
    const int N = A0->N;
    for (int j = 0; j + 1 < N; ++j) {
      project(j);
      const i64 Dl0 = A0->Dl(j), d1 = A0->d[j], d2 = A0->d[j + 1], Dr0 = A0->Dr(j + 1);
      Factor f = factor(c(), b, Dl0 * d1, d2 * Dr0, svd_tol, dlim, std::min(A0->cap_r[j], A0->cap_l[j + 1]));
      A0->dims[j + 1] = f.k;
      gather_U(c(), b, f, A0->site(j), false);
      for (int l = 0; l < nov; ++l)
        update_lenv(p.work, env[l]->slot(j + 1), A0->site(j), p.ovl[2 * l], Dl0, d1, f.k, A[l]->Dr(j), 0);
      if (H) update_le(p.work, p.op, envop->slot(j + 1), p.T2, A0->site(j), Dl0, d1, A[nov]->Dr(j), H->wr(j), f.k);
      if (j == N - 2) gather_Vh(c(), b, f, A0->site(N - 1), true);
    }
  }
};

}  // namespace

// mode 0: maxsweeps full sweeps (left, then right) after left environments
// from Ain; 1: one left half-sweep; 2: one right half-sweep after right
// environments from Ain (which must then be right canonical).
void two_site_sweeps(Plan& p, State* A0, Env** env, Env* envop, State* Ain, int nov, const cd* coef, State** A,
                     MPO* H, int mode, int maxsweeps, double svd_tol, i64 dlim) {
  Ctx* c = p.ctx;
  const int N = A0->N;
  char* ex = (char*)p.extra;
  const i64 Dm = A0->max_dim();
  i64 dm = 1;
  for (int j = 0; j < N; ++j) dm = std::max(dm, A0->d[j]);
  const size_t mn = (size_t)(Dm * dm) * (Dm * dm);
  SvdBufs b;
  b.T = (double2*)ex;
  b.Y = b.T + mn;
  b.J = b.Y + mn;
  b.sig = (double*)(b.J + mn);
  b.scl = b.sig + (Dm * dm + 64);
  b.perm = (int*)(b.scl + (Dm * dm + 64));
  b.flag = b.perm + (Dm * dm + 64);
  Sweep s{p, A0, env, envop, nov, A, coef, H, b};
  const cd cH = H ? coef[nov] : cd(0.0);
  if (mode == 2) {
    // right environments from Ain: slots N-1 .. 2; coefficients at slot 0
    for (int l = 0; l < nov; ++l) {
      Env* e = env[l];
      launch_set_scalar(e->slot(N), 1.0, 0.0, c->stream);
      for (int j = N - 1; j >= 2; --j) {
        prod_LR_32(p.work, p.AL, A[l]->site(j), e->slot(j + 1), A[l]->Dl(j), A[l]->d[j], A[l]->Dr(j), Ain->Dr(j),
                   false);
        update_renv(p.work, e->slot(j), Ain->site(j), p.AL, Ain->Dl(j), Ain->d[j], Ain->Dr(j), A[l]->Dl(j));
      }
      launch_set_scalar(e->slot(0), coef[l].real(), coef[l].imag(), c->stream);
    }
    if (H) {
      State* S = A[nov];
      launch_set_scalar(envop->slot(N), 1.0, 0.0, c->stream);
      for (int j = N - 1; j >= 2; --j) {
        prod_AOR(p.work, p.op, p.T2, S->site(j), S->Dl(j), S->d[j], S->Dr(j), H->site(j), H->wl(j), H->wr(j),
                 envop->slot(j + 1), Ain->Dr(j));
        update_re(p.work, p.op, envop->slot(j), p.T2, Ain->site(j), S->Dl(j), H->wl(j), S->d[j], Ain->Dr(j),
                  Ain->Dl(j));
      }
      launch_set_scalar(envop->slot(0), cH.real(), cH.imag(), c->stream);
    }
    A0->dims = Ain->dims;
    s.right(svd_tol, dlim);
    c->sync();
    return;
  }
  // left environments from Ain: slots 1 .. N-2; coefficients at slot N
  for (int l = 0; l < nov; ++l) {
    Env* e = env[l];
    launch_set_scalar(e->slot(0), 1.0, 0.0, c->stream);
    for (int j = 0; j + 2 < N; ++j) {
      prod_LR_23(p.work, p.AL, e->slot(j), A[l]->site(j), Ain->Dl(j), A[l]->Dl(j), A[l]->d[j], A[l]->Dr(j), false,
                 false);
      update_lenv(p.work, e->slot(j + 1), Ain->site(j), p.AL, Ain->Dl(j), Ain->d[j], Ain->Dr(j), A[l]->Dr(j), 0);
    }
    launch_set_scalar(e->slot(N), coef[l].real(), coef[l].imag(), c->stream);
  }
  if (H) {
    State* S = A[nov];
    launch_set_scalar(envop->slot(0), 1.0, 0.0, c->stream);
    for (int j = 0; j + 2 < N; ++j) {
      prod_AOL(p.work, p.op, p.T2, S->site(j), S->Dl(j), S->d[j], S->Dr(j), H->site(j), H->wl(j), H->wr(j),
               envop->slot(j), Ain->Dl(j));
      update_le(p.work, p.op, envop->slot(j + 1), p.T2, Ain->site(j), Ain->Dl(j), Ain->d[j], S->Dr(j), H->wr(j),
                Ain->Dr(j));
    }
    launch_set_scalar(envop->slot(N), cH.real(), cH.imag(), c->stream);
  }
  A0->dims = Ain->dims;
  if (mode == 1) {
    s.left(svd_tol, dlim);
  } else {
    for (int sw = 0; sw < maxsweeps; ++sw) {
      s.left(svd_tol, dlim);
      s.right(svd_tol, dlim);
    }
  }
  c->sync();
}

}  // namespace mpship
