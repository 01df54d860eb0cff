// steady_state_search! (steadystate.jl:6-86): one-site variational search
// of the state minimising <rho| L^dag L |rho>, i.e. the null vector of a
// Lindbladian L, by sweeps of local smallest-eigenvalue problems.
//
// The reference contracts L and Lconj (= L^dag) through four-index
// environments (update_lenv_two_op / update_renv_two_op and
// A_prod_L_OO_R_eigs, MatrixProductStates.jl:263-312); with the two MPO
// bonds combined into one (w1 fastest) these are exactly the one-operator
// environments and effective operator of the combined MPO
// W = applyMPOtoMPO(L, Lconj) (site: W[w1 + wl1 w2, s', .., s] =
// sum_t Lconj[w2, s', .., t] L[w1, t, .., s]), which the caller passes.
// The local problem H_eff x = prod_LR(E_L, prod_AOR(x, W[j], E_R)) is the
// same product the H-applying sweeps form.  Its smallest eigenpair comes,
// for local spaces of at most kDenseMax elements, from the dense operator
// (applied to unit vectors) and a one-sided Jacobi SVD on the device (the
// operator is positive semidefinite, so its right singular vectors are its
// eigenvectors; the null space of a Lindbladian makes the lowest eigenvalue
// sit far below the spread of the spectrum, where a Krylov method from v0
// stalls), and for larger spaces from a Lanczos iteration on device vectors
// (Krylov basis in the plan's extra buffer, classical Gram-Schmidt applied
// twice through two GEMMs per step, the small tridiagonal problem solved on
// the host by Jacobi rotations), started from the current site as the
// reference's eigs(...; v0 = At[jj]) is, and restarted from the Ritz vector.  Gauge moves follow the reference:
// the left sweep factors the adjoint of reshape(av, Dl, d Dr), keeps Q' and
// multiplies the next site by R^T (steadystate.jl:38-43, as written); the
// right sweep factors reshape(av, Dl d, Dr) and multiplies the next site by R.
#include "contract.h"
#include "sweeps.h"
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace mpship {

using cd = std::complex<double>;

namespace {

// eigenpairs of a real symmetric k x k matrix (row major) by cyclic Jacobi
void sym_eig(std::vector<double> A, int k, std::vector<double>& w, std::vector<double>& V) {
  V.assign((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) V[(size_t)i * k + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) off += A[(size_t)p * k + q] * A[(size_t)p * k + q];
    if (off < 1e-300) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[(size_t)p * k + q];
        if (std::abs(apq) < 1e-300) continue;
        const double app = A[(size_t)p * k + p], aqq = A[(size_t)q * k + q];
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::abs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {   // A <- J^T A J
          const double arp = A[(size_t)r * k + p], arq = A[(size_t)r * k + q];
          A[(size_t)r * k + p] = c * arp - s * arq;
          A[(size_t)r * k + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {
          const double apr = A[(size_t)p * k + r], aqr = A[(size_t)q * k + r];
          A[(size_t)p * k + r] = c * apr - s * aqr;
          A[(size_t)q * k + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const double vrp = V[(size_t)r * k + p], vrq = V[(size_t)r * k + q];
          V[(size_t)r * k + p] = c * vrp - s * vrq;
          V[(size_t)r * k + q] = s * vrp + c * vrq;
        }
      }
  }
  w.resize(k);
  for (int i = 0; i < k; ++i) w[i] = A[(size_t)i * k + i];
}

struct Site {
  i64 Dl, d, Dr;
  i64 n() const { return Dl * d * Dr; }
};

// Device vectors and the effective operator of site j.
// local problems up to this size are solved densely (the matrix formed by
// applying the operator to unit vectors, then a one-sided Jacobi SVD, whose
// right singular vectors are the eigenvectors of the PSD operator)
constexpr i64 kDenseMax = 1536;

struct Local {
  Plan& p;
  State* At;
  MPO* W;
  Env* envop;
  int j;
  Site sh;
  double2* V;      // Krylov basis, column k at V + k n (the dense matrices, n <= kDenseMax)
  double2* w;      // work vector
  double2* c;      // projections (kmax complex)
  int kmax;
  Ctx* ctx() const { return p.ctx; }
  // y = H_eff x
  void apply(const double2* x, double2* y) {
    prod_AOR(p.work, p.op, p.T2, x, sh.Dl, sh.d, sh.Dr, W->site(j), W->wl(j), W->wr(j), envop->slot(j + 1), sh.Dr);
    prod_LR_34(p.work, p.op, y, envop->slot(j), p.T2, sh.Dl, W->wl(j), sh.Dl, sh.d, sh.Dr, false);
  }
  // c[0..k) = V[:, 0..k)^H x  (one GEMM), copied to the host
  void project(const double2* x, int k, std::vector<cd>& out) {
    Gemm g;
    g.M = k; g.N = 1; g.K = sh.n();
    g.A = V; g.am = im1(sh.n()); g.ak = im1(1); g.conjA = true;
    g.B = x; g.bk = im1(1); g.bn = im1(sh.n());
    g.C = c; g.cm = im1(1); g.cn = im1(k);
    run_gemm(ctx(), g, CMP_F64, p.splitk, p.splitk_bytes);
    out.resize(k);
    HIPCHK(hipMemcpyAsync(ctx()->hscal, c, k * sizeof(double2), hipMemcpyDeviceToHost, ctx()->stream));
    ctx()->sync();
    for (int i = 0; i < k; ++i) out[i] = cd(ctx()->hscal[2 * i], ctx()->hscal[2 * i + 1]);
  }
  // x <- alpha x + V[:, 0..k) coef (coef from the host)
  void combine(double2* x, int k, const std::vector<cd>& coef, double alpha) {
    for (int i = 0; i < k; ++i) {
      ctx()->hscal[2 * i] = coef[i].real();
      ctx()->hscal[2 * i + 1] = coef[i].imag();
    }
    HIPCHK(hipMemcpyAsync(c, ctx()->hscal, k * sizeof(double2), hipMemcpyHostToDevice, ctx()->stream));
    Gemm g;
    g.M = sh.n(); g.N = 1; g.K = k;
    g.A = V; g.am = im1(1); g.ak = im1(sh.n());
    g.B = c; g.bk = im1(1); g.bn = im1(k);
    g.C = x; g.cm = im1(1); g.cn = im1(sh.n());
    g.beta_re = alpha;
    run_gemm(ctx(), g, CMP_F64, p.splitk, p.splitk_bytes);
    ctx()->sync();   // hscal is reused by the next projection
  }
  double norm(const double2* x) {
    Gemm g;
    g.M = 1; g.N = 1; g.K = sh.n();
    g.A = x; g.am = im1(sh.n()); g.ak = im1(1); g.conjA = true;
    g.B = x; g.bk = im1(1); g.bn = im1(sh.n());
    g.C = c; g.cm = im1(1); g.cn = im1(1);
    run_gemm(ctx(), g, CMP_F64, p.splitk, p.splitk_bytes);
    HIPCHK(hipMemcpyAsync(ctx()->hscal, c, sizeof(double2), hipMemcpyDeviceToHost, ctx()->stream));
    ctx()->sync();
    return std::sqrt(std::max(ctx()->hscal[0], 0.0));
  }
  void scale(double2* x, double f) { launch_scale_c(x, sh.n(), f, ctx()->stream); }

  // Rayleigh quotient <x|H|x> / <x|x>
  double rayleigh(const double2* x) {
    apply(x, w);
    Gemm g;
    g.M = 1; g.N = 1; g.K = sh.n();
    g.A = x; g.am = im1(sh.n()); g.ak = im1(1); g.conjA = true;
    g.B = w; g.bk = im1(1); g.bn = im1(sh.n());
    g.C = c; g.cm = im1(1); g.cn = im1(1);
    run_gemm(ctx(), g, CMP_F64, p.splitk, p.splitk_bytes);
    HIPCHK(hipMemcpyAsync(ctx()->hscal, c, sizeof(double2), hipMemcpyDeviceToHost, ctx()->stream));
    ctx()->sync();
    const double num = ctx()->hscal[0], nx = norm(x);
    return num / (nx * nx);
  }

  // Dense local solve: H (n x n) column by column, Y = H J with orthogonal
  // columns by one-sided Jacobi, the column of J with the smallest ||y_k||.
  double lowest_dense() {
    double2* site = At->site(j);
    const i64 n = sh.n();
    double2* H = V;
    double2* Y = H + (size_t)n * n;
    double2* J = Y + (size_t)n * n;
    double* sig = (double*)(J + (size_t)n * n);
    int* flag = (int*)(sig + n + 2);
    hipStream_t st = ctx()->stream;
    for (i64 i = 0; i < n; ++i) {
      launch_fill_zero(w, n, st);
      launch_set_scalar(w + i, 1.0, 0.0, st);
      apply(w, H + (size_t)i * n);
    }
    svd_load(Y, H, n, n, false, st);
    svd_jacobi(Y, n, n, J, 4.0 * std::sqrt((double)n) * 2.220446049250313e-16, flag, (int*)ctx()->hscal, 100, st);
    svd_colnorms(Y, n, n, sig, st);
    std::vector<double> s(n);
    HIPCHK(hipMemcpyAsync(s.data(), sig, n * sizeof(double), hipMemcpyDeviceToHost, st));
    ctx()->sync();
    const i64 k = std::min_element(s.begin(), s.end()) - s.begin();
    launch_copy_c(site, J + (size_t)k * n, n, st);
    return rayleigh(site);
  }

  // Smallest eigenpair of H_eff, from the current site; the eigenvector
  // (unit norm) replaces the site.  Returns the eigenvalue.
  double lowest(double tol, int max_restarts) {
    if (sh.n() <= kDenseMax) return lowest_dense();
    double2* site = At->site(j);
    const i64 n = sh.n();
    const int km = (int)std::min<i64>(kmax, n);
    double theta = 0.0;
    for (int restart = 0; restart <= max_restarts; ++restart) {
      double nr = norm(site);
      if (!(nr > 0.0)) {   // a zero start: the first unit vector
        launch_fill_zero(site, n, ctx()->stream);
        launch_set_scalar(site, 1.0, 0.0, ctx()->stream);
        nr = 1.0;
      }
      launch_copy_c(V, site, n, ctx()->stream);
      scale(V, 1.0 / nr);
      std::vector<double> T((size_t)km * km, 0.0), ev, Y;
      int k = 0;
      bool done = false;
      for (; k < km; ++k) {
        double2* vk = V + (size_t)k * n;
        apply(vk, w);
        std::vector<cd> h1, h2;
        project(w, k + 1, h1);
        std::vector<cd> m1(h1);
        for (auto& x : m1) x = -x;
        combine(w, k + 1, m1, 1.0);
        project(w, k + 1, h2);   // second Gram-Schmidt pass
        for (auto& x : h2) x = -x;
        combine(w, k + 1, h2, 1.0);
        T[(size_t)k * km + k] = h1[k].real() - h2[k].real();
        const double beta = norm(w);
        // Ritz values of the leading (k+1) x (k+1) block
        std::vector<double> Tk((size_t)(k + 1) * (k + 1));
        for (int a = 0; a <= k; ++a)
          for (int b = 0; b <= k; ++b) Tk[(size_t)a * (k + 1) + b] = T[(size_t)a * km + b];
        sym_eig(Tk, k + 1, ev, Y);
        int imin = 0;
        for (int a = 1; a <= k; ++a)
          if (ev[a] < ev[imin]) imin = a;
        double tmax = 0.0;
        for (double e : ev) tmax = std::max(tmax, std::abs(e));
        theta = ev[imin];
        const double resid = beta * std::abs(Y[(size_t)k * (k + 1) + imin]);
        if (resid <= tol * std::max(std::abs(theta), tmax) || beta <= 1e-14 * std::max(tmax, 1e-300) ||
            k + 1 == km) {
          // Ritz vector x = V y (written into the site)
          std::vector<cd> y(k + 1);
          for (int a = 0; a <= k; ++a) y[a] = Y[(size_t)a * (k + 1) + imin];
          launch_fill_zero(site, n, ctx()->stream);
          combine(site, k + 1, y, 1.0);
          scale(site, 1.0 / norm(site));
          done = resid <= tol * std::max(std::abs(theta), tmax) || beta <= 1e-14 * std::max(tmax, 1e-300);
          break;
        }
        if (k + 1 < km) {
          T[(size_t)k * km + k + 1] = T[(size_t)(k + 1) * km + k] = beta;
          launch_copy_c(V + (size_t)(k + 1) * n, w, n, ctx()->stream);
          scale(V + (size_t)(k + 1) * n, 1.0 / beta);
        }
      }
      if (done) return theta;
    }
    return theta;
  }
};

}  // namespace

size_t steady_state_extra_bytes(i64 site_elems, int kmax) {
  const size_t krylov = ((size_t)kmax * site_elems) * sizeof(double2);
  const i64 nd = std::min<i64>(site_elems, kDenseMax);
  const size_t dense = (3 * (size_t)nd * nd) * sizeof(double2) + (size_t)(nd + 8) * sizeof(double);
  return std::max(krylov, dense) + ((size_t)site_elems + 64) * sizeof(double2) + 4096;
}

double steady_state_search(Plan& p, State* At, MPO* W, Env* envop, int maxsweeps, double converge_rat,
                           double eigs_tol, int kmax, int* sweeps_run) {
  Ctx* c = p.ctx;
  const int N = At->N;
  const i64 nmax = At->max_site();
  double2* V = (double2*)p.extra;
  const i64 nd = std::min<i64>(nmax, kDenseMax);
  const size_t vspan = std::max((size_t)kmax * nmax, 3 * (size_t)nd * nd + (size_t)nd + 8);
  double2* w = V + vspan;
  double2* cvec = w + nmax;
  auto local = [&](int j) {
    Local L{p, At, W, envop, j, Site{At->Dl(j), At->d[j], At->Dr(j)}, V, w, cvec, kmax};
    return L;
  };
  // FL: left environments of At (left canonical) with W; FR[na] = 1
  launch_set_scalar(envop->slot(0), 1.0, 0.0, c->stream);
  for (int j = 0; j + 1 < N; ++j) {
    prod_AOL(p.work, p.op, p.T2, At->site(j), At->Dl(j), At->d[j], At->Dr(j), W->site(j), W->wl(j), W->wr(j),
             envop->slot(j), At->Dl(j));
    update_le(p.work, p.op, envop->slot(j + 1), p.T2, At->site(j), At->Dl(j), At->d[j], At->Dr(j), W->wr(j),
              At->Dr(j));
  }
  launch_set_scalar(envop->slot(N), 1.0, 0.0, c->stream);
  double ae = 0.0, aestart = 0.0;
  int sw = 0;
  for (; sw < maxsweeps; ++sw) {
    // sweep left
    for (int j = N - 1; j >= 1; --j) {
      ae = local(j).lowest(eigs_tol, 20);
      if (j == N - 1) aestart = ae;
      const i64 Dl = At->Dl(j), d = At->d[j], Dr = At->Dr(j);
      // av, rv = qr(reshape(av, Dl, d Dr)'); At[jj] = reshape(av', ...)
      launch_conj_c(At->site(j), Dl * d * Dr, c->stream);
      qr_right(p, At, j, p.R);
      launch_conj_c(At->site(j), Dl * d * Dr, c->stream);
      // FR[jj-1] = update_renv_two_op(At[jj], L, Lconj, At[jj], FR[jj])
      prod_AOR(p.work, p.op, p.T2, At->site(j), Dl, d, Dr, W->site(j), W->wl(j), W->wr(j), envop->slot(j + 1), Dr);
      update_re(p.work, p.op, envop->slot(j), p.T2, At->site(j), Dl, W->wl(j), d, Dr, Dl);
      // At[jj-1] = At[jj-1] * rv^T
      const i64 Dlm = At->Dl(j - 1), dm = At->d[j - 1];
      prod_LR_32(p.work, p.tmp, At->site(j - 1), p.R, Dlm, dm, Dl, Dl, false, true);
      launch_copy_c(At->site(j - 1), p.tmp, Dlm * dm * Dl, c->stream);
    }
    // sweep right
    for (int j = 0; j + 1 < N; ++j) {
      ae = local(j).lowest(eigs_tol, 20);
      const i64 Dl = At->Dl(j), d = At->d[j], Dr = At->Dr(j);
      qr_left(p, At, j, p.R);
      prod_AOL(p.work, p.op, p.T2, At->site(j), Dl, d, Dr, W->site(j), W->wl(j), W->wr(j), envop->slot(j), Dl);
      update_le(p.work, p.op, envop->slot(j + 1), p.T2, At->site(j), Dl, d, Dr, W->wr(j), Dr);
      // At[jj+1] = rv * At[jj+1]
      const i64 d1 = At->d[j + 1], Dr1 = At->Dr(j + 1);
      prod_LR_23(p.work, p.tmp, p.R, At->site(j + 1), Dr, Dr, d1, Dr1, false, false);
      launch_copy_c(At->site(j + 1), p.tmp, Dr * d1 * Dr1, c->stream);
    }
    ae = local(N - 1).lowest(eigs_tol, 20);
    if (std::abs(ae) > converge_rat * std::abs(aestart) || std::abs(ae) < 1e-10) {
      ++sw;
      break;
    }
  }
  c->sync();
  if (sweeps_run) *sweeps_run = sw;
  return ae;
}

}  // namespace mpship
