// BiCGStab(4), GPBiCGStab(2), GPBiCGStab(4) behind the reference entry points
// (ddscat_cuda_unified/cuda/ddscat_matvec_cuda.cu:1363-1570).  These keep
// the reference structure: vectors on the device, FP64 dot products reduced
// on the device and returned to the host, where the small minimisation
// systems (2x2 ... 5x5) are solved in complex<double>.  Gram matrices are
// reduced in one multi-dot launch instead of one reduction per entry.
#include "solvers.hpp"
#include "../../include/ddscat_matvec_hip.h"
#include <complex>
#include <vector>
#include <cmath>
#include <utility>

namespace ddhip {

using zc = std::complex<double>;

namespace {

struct Ops {
  Context& c;
  size_t n;
  hipStream_t st;
  Ops(Context& cc) : c(cc), n((size_t)3 * cc.nat0), st(cc.stream) {}
  // batched conjugated dots <a_i, b_i> = sum conj(a_i) b_i, one reduction launch
  int dotv(const std::vector<std::pair<const float2*, const float2*>>& ab, zc* out) {
    for (size_t i0 = 0; i0 < ab.size(); i0 += DOT_MAX) {
      DotSet d;
      const size_t m = std::min<size_t>(DOT_MAX, ab.size() - i0);
      for (size_t i = 0; i < m; ++i) d.add(ab[i0 + i].first, ab[i0 + i].second, DOT_C);
      if (dots(c, d, n, SL_D, st) || fetch_scalars(c, SL_D, (int)(2 * m), st)) return 1;
      for (size_t i = 0; i < m; ++i) out[i0 + i] = zc(c.hscal[SL_D + 2 * i], c.hscal[SL_D + 2 * i + 1]);
    }
    return 0;
  }
  int dot(const float2* a, const float2* b, zc& out) { return dotv({{a, b}}, &out); }
  int norm(const float2* a, double& out) {
    zc z;
    if (dot(a, a, z)) return 1;
    double v = z.real();
    out = std::sqrt(std::max(0.0, v));
    return 0;
  }
  int lc(float2* d, std::initializer_list<std::pair<const float2*, zc>> terms) {
    LinComb l;
    for (auto& t : terms) l.lit(t.first, t.second.real(), t.second.imag());
    return lincomb(c, d, l, n, st);
  }
  int axpy(float2* d, const float2* a, zc z) { return lc(d, {{d, 1.0}, {a, z}}); }
  int affine(float2* d, const float2* a, zc z) { return lc(d, {{a, 1.0}, {d, z}}); }  // d = a + z d
  int scale(float2* d, const float2* a, zc z) { return lc(d, {{a, z}}); }
  int copy(float2* d, const float2* a) { return lc(d, {{a, 1.0}}); }
  // no terms: exact zeros (0 * d would keep a NaN or Inf of uninitialised memory)
  int zero(float2* d) { return lincomb(c, d, LinComb{}, n, st); }
};

// Gaussian elimination with partial pivoting on a[n][ld] (row major), b[n]
bool solve_small(zc* a, zc* b, int n, int ld) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = std::abs(a[k * ld + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = std::abs(a[i * ld + k]);
      if (v > best) { best = v; piv = i; }
    }
    if (best < 1.e-30) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * ld + j], a[piv * ld + j]);
      std::swap(b[k], b[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      const zc f = a[i * ld + k] / a[k * ld + k];
      for (int j = k; j < n; ++j) a[i * ld + j] -= f * a[k * ld + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    zc v = b[i];
    for (int j = i + 1; j < n; ++j) v -= a[i * ld + j] * b[j];
    if (std::abs(a[i * ld + i]) < 1.e-30) return false;
    b[i] = v / a[i * ld + i];
  }
  return true;
}

// r = b - A x into r (A x left in c.sy)
int initial_residual(Context& c, Ops& o, const float2* x, float2* r, int& nc) {
  if (int rc = run_mv(c, 'N', x, c.sy, nc)) return rc;
  return o.lc(r, {{c.sb, 1.0}, {c.sy, -1.0}});
}

}  // namespace

// ============================================================ BiCGStab(4) ==
int solve_bicgstab4(Context& c, const SolverArgs& a) {
  const char* nm = "BiCGStab4";
  clear_error();
  if (!c.ready) { set_error(nm, "invalid argument"); return 130; }
  if (int rc = ensure_solver_workspace(c, 13)) return rc;
  Ops o(c);
  const size_t n = o.n;
  float2* W = c.wrk;
  float2* r[5] = {W, W + 3 * n, W + 4 * n, W + 5 * n, W + 6 * n};
  float2* u[5] = {W + n, W + 7 * n, W + 8 * n, W + 9 * n, W + 10 * n};
  float2* rt = W + 2 * n;
  float2* x = c.sx;
  if (int rc = solver_begin(c, a, 130, 131, nm, x)) return rc;
  int nc = 0, rc = 0;
  if ((rc = initial_residual(c, o, x, r[0], nc))) return rc;
  double bnorm = 0, rnorm = 0;
  if (o.norm(c.sb, bnorm) || o.norm(r[0], rnorm)) return 132;
  if (bnorm == 0.0) { set_error(nm, "zero RHS norm"); return 133; }
  double resid = rnorm / bnorm;
  if (resid <= a.tol) return solver_end(c, a, x, 0, resid, nc, 133);
  if (rnorm < 1.e-30) { set_error(nm, "cannot start from zero residual"); return 133; }
  if (o.scale(rt, r[0], 1.0 / rnorm) || o.zero(u[0])) return 134;
  zc rho(1, 0), alpha(0, 0), omega(1, 0), rho1, beta, den;
  zc tau[4][4], gp[5], gam[5], gpp[5];
  double sigma[5] = {0, 0, 0, 0, 0};
  int it;
  for (it = 1; it <= a.maxit; ++it) {
    rho = -omega * rho;
    for (int j = 0; j < 4; ++j) {
      if (std::abs(rho) < 1.e-30) { set_error(nm, "breakdown: rho is zero"); return 135; }
      if (o.dot(rt, r[j], rho1)) return 135;
      beta = alpha * rho1 / rho;
      rho = rho1;
      for (int i = 0; i <= j; ++i)
        if (o.affine(u[i], r[i], -beta)) return 136;
      if ((rc = run_mv(c, 'N', u[j], u[j + 1], nc))) return rc;
      if (o.dot(rt, u[j + 1], den)) return 137;
      if (std::abs(den) < 1.e-30) { set_error(nm, "breakdown: <rtilde,u> is zero"); return 137; }
      alpha = rho / den;
      for (int i = 0; i <= j; ++i)
        if (o.axpy(r[i], u[i + 1], -alpha)) return 138;
      if ((rc = run_mv(c, 'N', r[j], r[j + 1], nc))) return rc;
      if (o.axpy(x, u[0], alpha)) return 139;
    }
    for (auto& row : tau) for (auto& v : row) v = 0.0;
    for (int k = 0; k < 5; ++k) gp[k] = gam[k] = gpp[k] = 0.0;
    // modified Gram-Schmidt minimal-residual polynomial (cu:1398-1412)
    for (int j = 1; j <= 4; ++j) {
      for (int i = 1; i < j; ++i) {
        zc d;
        if (o.dot(r[i], r[j], d)) return 140;
        tau[j - 1][i - 1] = d / sigma[i];
        if (o.axpy(r[j], r[i], -tau[j - 1][i - 1])) return 140;
      }
      zc pr[2];
      if (o.dotv({{r[j], r[j]}, {r[j], r[0]}}, pr)) return 141;
      sigma[j] = pr[0].real();
      if (!(sigma[j] > 1.e-30)) { set_error(nm, "breakdown in MR polynomial"); return 141; }
      gp[j] = pr[1] / sigma[j];
    }
    omega = gam[4] = gp[4];
    for (int j = 3; j >= 1; --j) {
      gam[j] = gp[j];
      for (int i = j + 1; i <= 4; ++i) gam[j] -= tau[i - 1][j - 1] * gam[i];
    }
    for (int j = 1; j < 4; ++j) {
      gpp[j] = gam[j + 1];
      for (int i = j + 1; i < 4; ++i) gpp[j] += tau[i - 1][j - 1] * gam[i + 1];
    }
    if (o.axpy(x, r[0], gam[1])) return 142;
    if (o.axpy(r[0], r[4], -gp[4]) || o.axpy(u[0], u[4], -gam[4])) return 142;
    for (int j = 1; j < 4; ++j)
      if (o.axpy(x, r[j], gpp[j]) || o.axpy(r[0], r[j], -gp[j]) || o.axpy(u[0], u[j], -gam[j])) return 142;
    if (o.norm(r[0], rnorm)) return 143;
    resid = rnorm / bnorm;
    if ((it % RELIABLE_PERIOD) == 0 || resid <= a.tol) {
      Reliable rr;
      if ((rc = reliable_check(c, c.sb, x, r[0], bnorm, resid, a.tol, nc, rr))) return rc > 1 ? rc : 144;
      if (rr.converged) { resid = rr.true_res; break; }
      if (rr.restart) {
        if (o.lc(r[0], {{c.sb, 1.0}, {c.sy, -1.0}})) return 144;
        rnorm = rr.true_res * bnorm;
        if (!(rnorm > 0.0)) { set_error(nm, "reliable restart produced zero residual"); return 144; }
        if (o.scale(rt, r[0], 1.0 / rnorm) || o.zero(u[0])) return 144;
        rho = 1.0; alpha = 0.0; omega = 1.0;
        resid = rr.true_res;
      }
    }
  }
  // final true residual, reported without changing x (cu:1421-1423)
  if ((rc = initial_residual(c, o, x, r[0], nc))) return rc;
  if (o.norm(r[0], rnorm)) return 144;
  resid = rnorm / bnorm;
  return solver_end(c, a, x, it <= a.maxit ? it : a.maxit, resid, nc, 145);
}

// ========================================================= GPBiCGStab(2) ==
int solve_gpbicgstab2(Context& c, const SolverArgs& a) {
  const char* nm = "GPBiCGStab2";
  clear_error();
  if (!c.ready) { set_error(nm, "invalid argument"); return 150; }
  if (int rc = ensure_solver_workspace(c, 13)) return rc;
  Ops o(c);
  const size_t n = o.n;
  float2* W = c.wrk;
  float2 *r0 = W, *p0 = W + n, *rt = W + 2 * n, *z = W + 3 * n, *yv = W + 4 * n, *u = W + 5 * n, *hs = W + 6 * n,
         *hq0 = W + 7 * n, *hq1 = W + 8 * n, *work = W + 9 * n;
  float2* x = c.sx;
  if (int rc = solver_begin(c, a, 150, 151, nm, x)) return rc;
  int nc = 0, rc = 0;
  if ((rc = initial_residual(c, o, x, r0, nc))) return rc;
  double bnorm = 0, rnorm = 0;
  if (o.norm(c.sb, bnorm) || o.norm(r0, rnorm)) return 152;
  if (bnorm == 0.0) { set_error(nm, "zero RHS norm"); return 153; }
  if (o.copy(rt, r0) || o.copy(p0, r0) || o.zero(z)) return 153;
  bool fresh = true;
  double resid = rnorm / bnorm;
  zc rho, sigma, alpha, beta, zeta1, zeta2, eta;
  zc mat[3][3], rhs[3];
  int it;
  for (it = 1; it <= a.maxit && resid > a.tol; ++it) {
    if (fresh) {
      if (o.dot(rt, r0, rho)) return 154;
      if ((rc = run_mv(c, 'N', p0, work, nc))) return rc;
      if (o.dot(rt, work, sigma)) return 154;
      if (std::abs(sigma) < 1.e-30) { set_error(nm, "sigma1 zero"); return 154; }
      alpha = rho / sigma;
      if (o.axpy(x, p0, alpha) || o.axpy(r0, work, -alpha)) return 155;
      if ((rc = run_mv(c, 'N', r0, hs, nc))) return rc;
      if (o.dot(rt, hs, rho)) return 155;
      beta = rho / sigma;
      if (o.affine(p0, r0, -beta) || o.affine(work, hs, -beta)) return 156;
      if ((rc = run_mv(c, 'N', work, hq1, nc))) return rc;
      if (o.dot(rt, hq1, sigma)) return 156;
      if (std::abs(sigma) < 1.e-30) { set_error(nm, "sigma2 zero"); return 156; }
      alpha = rho / sigma;
      if (o.axpy(x, p0, alpha) || o.axpy(r0, work, -alpha) || o.axpy(hs, hq1, -alpha)) return 157;
      if ((rc = run_mv(c, 'N', hs, hq0, nc))) return rc;
      if (o.dot(rt, hq0, rho)) return 157;
      beta = rho / sigma;
      if (o.affine(p0, r0, -beta) || o.affine(work, hs, -beta) || o.affine(hq1, hq0, -beta)) return 158;
      zc g[5];
      if (o.dotv({{hs, hs}, {hs, hq0}, {hq0, hq0}, {hs, r0}, {hq0, r0}}, g)) return 159;
      mat[0][0] = g[0]; mat[0][1] = g[1]; mat[1][1] = g[2]; rhs[0] = g[3]; rhs[1] = g[4];
      mat[1][0] = std::conj(mat[0][1]);
      if (!solve_small(&mat[0][0], rhs, 2, 3)) { set_error(nm, "initial 2x2 minimization breakdown"); return 159; }
      zeta1 = rhs[0]; zeta2 = rhs[1];
      if (o.lc(z, {{r0, zeta1}, {hs, zeta2}}) || o.axpy(x, z, 1.0) || o.lc(yv, {{hs, zeta1}, {hq0, zeta2}}) ||
          o.lc(u, {{work, zeta1}, {hq1, zeta2}}) || o.axpy(r0, yv, -1.0) || o.axpy(p0, u, -1.0) || o.copy(hq0, work))
        return 160;
      fresh = false;
    } else {
      if (o.dot(rt, r0, rho)) return 161;
      if ((rc = run_mv(c, 'N', p0, work, nc))) return rc;
      if (o.dot(rt, work, sigma)) return 161;
      if (std::abs(sigma) < 1.e-30) { set_error(nm, "sigma1 zero"); return 161; }
      alpha = rho / sigma;
      if (o.axpy(x, p0, alpha) || o.axpy(z, u, -alpha) || o.lc(yv, {{yv, 1.0}, {hq0, -alpha}, {work, alpha}}) ||
          o.axpy(r0, work, -alpha) || o.axpy(hs, hq1, -alpha))
        return 162;
      if ((rc = run_mv(c, 'N', r0, hq1, nc))) return rc;
      if (o.dot(rt, hq1, rho)) return 162;
      beta = rho / sigma;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq1, -beta) || o.affine(hq0, hs, -beta) || o.affine(u, yv, -beta))
        return 163;
      if ((rc = run_mv(c, 'N', work, hs, nc))) return rc;
      if (o.dot(rt, hs, sigma)) return 163;
      if (std::abs(sigma) < 1.e-30) { set_error(nm, "sigma2 zero"); return 163; }
      alpha = rho / sigma;
      if (o.axpy(x, p0, alpha) || o.axpy(z, u, -alpha) || o.lc(yv, {{yv, 1.0}, {hq0, -alpha}, {work, alpha}}) ||
          o.axpy(r0, work, -alpha) || o.axpy(hq1, hs, -alpha))
        return 164;
      if ((rc = run_mv(c, 'N', hq1, hq0, nc))) return rc;
      if (o.dot(rt, hq0, rho)) return 164;
      beta = rho / sigma;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq1, -beta) || o.affine(hs, hq0, -beta) || o.affine(u, yv, -beta))
        return 165;
      zc g[9];
      if (o.dotv({{hq1, hq1}, {hq1, hq0}, {hq1, yv}, {hq0, hq0}, {hq0, yv}, {yv, yv}, {hq1, r0}, {hq0, r0}, {yv, r0}},
                 g))
        return 166;
      mat[0][0] = g[0]; mat[0][1] = g[1]; mat[0][2] = g[2]; mat[1][1] = g[3]; mat[1][2] = g[4]; mat[2][2] = g[5];
      rhs[0] = g[6]; rhs[1] = g[7]; rhs[2] = g[8];
      mat[1][0] = std::conj(mat[0][1]); mat[2][0] = std::conj(mat[0][2]); mat[2][1] = std::conj(mat[1][2]);
      if (!solve_small(&mat[0][0], rhs, 3, 3)) { set_error(nm, "3x3 minimization breakdown"); return 166; }
      zeta1 = rhs[0]; zeta2 = rhs[1]; eta = rhs[2];
      if (o.lc(z, {{z, eta}, {r0, zeta1}, {hq1, zeta2}}) || o.axpy(x, z, 1.0) ||
          o.lc(yv, {{yv, eta}, {hq1, zeta1}, {hq0, zeta2}}) || o.lc(u, {{u, eta}, {work, zeta1}, {hs, zeta2}}) ||
          o.axpy(r0, yv, -1.0) || o.axpy(p0, u, -1.0))
        return 167;
      // fixed boundary histories: hs = r1, hq0 = p1, hq1 = p2 (cu:1478)
      if (o.copy(hq0, hs) || o.copy(hs, hq1) || o.copy(hq1, hq0) || o.copy(hq0, work)) return 168;
    }
    if (o.norm(r0, rnorm)) return 169;
    resid = rnorm / bnorm;
    if ((it % RELIABLE_PERIOD) == 0 || resid <= a.tol) {
      Reliable rr;
      if ((rc = reliable_check(c, c.sb, x, r0, bnorm, resid, a.tol, nc, rr))) return rc > 1 ? rc : 170;
      if (rr.converged) { resid = rr.true_res; break; }
      if (rr.restart) {
        if (o.lc(r0, {{c.sb, 1.0}, {c.sy, -1.0}}) || o.copy(rt, r0) || o.copy(p0, r0) || o.zero(z)) return 171;
        fresh = true;
        resid = rr.true_res;
      }
    }
  }
  if (resid > a.tol) {
    if ((rc = initial_residual(c, o, x, r0, nc))) return rc;
    if (o.norm(r0, rnorm)) return 172;
    resid = rnorm / bnorm;
  }
  return solver_end(c, a, x, it <= a.maxit ? it : a.maxit, resid, nc, 173);
}

// ========================================================= GPBiCGStab(4) ==
int solve_gpbicgstab4(Context& c, const SolverArgs& a) {
  const char* nm = "GPBiCGStab4";
  clear_error();
  if (!c.ready) { set_error(nm, "invalid argument"); return 180; }
  if (int rc = ensure_solver_workspace(c, 14)) return rc;
  Ops o(c);
  const size_t n = o.n;
  float2* W = c.wrk;
  float2 *r0 = W, *p0 = W + n, *rt = W + 2 * n, *z = W + 3 * n, *yv = W + 4 * n, *u = W + 5 * n;
  float2* hs[3] = {W + 6 * n, W + 7 * n, W + 8 * n};
  float2* hq[4] = {W + 9 * n, W + 10 * n, W + 11 * n, W + 12 * n};
  float2* work = W + 13 * n;
  float2* x = c.sx;
  if (int rc = solver_begin(c, a, 180, 181, nm, x)) return rc;
  int nc = 0, rc = 0;
  if ((rc = initial_residual(c, o, x, r0, nc))) return rc;
  double bnorm = 0, rnorm = 0;
  if (o.norm(c.sb, bnorm) || o.norm(r0, rnorm)) return 182;
  if (bnorm == 0.0) { set_error(nm, "zero RHS norm"); return 183; }
  if (o.copy(rt, r0) || o.copy(p0, r0) || o.zero(z)) return 183;
  bool fresh = true;
  double resid = rnorm / bnorm;
  zc rho, sigma, alpha, beta, zeta[4], eta;
  zc mat[5][5], rhs[5];
  // swap hs[k] <-> hq[3-k] and set hq0 = work (cu:1510-1513)
  auto restore = [&]() -> int {
    if (o.copy(hq[0], hs[0]) || o.copy(hs[0], hq[3]) || o.copy(hq[3], hq[0])) return 1;
    if (o.copy(hq[0], hs[1]) || o.copy(hs[1], hq[2]) || o.copy(hq[2], hq[0])) return 1;
    if (o.copy(hq[0], hs[2]) || o.copy(hs[2], hq[1]) || o.copy(hq[1], hq[0]) || o.copy(hq[0], work)) return 1;
    return 0;
  };
  auto sig = [&](const float2* v, int code, const char* what) -> int {
    if (o.dot(rt, v, sigma)) return code;
    if (std::abs(sigma) < 1.e-30) { set_error(nm, what); return code; }
    alpha = rho / sigma;
    return 0;
  };
  auto rhobeta = [&](const float2* v, int code) -> int {
    if (o.dot(rt, v, rho)) return code;
    beta = rho / sigma;
    return 0;
  };
  int it;
  for (it = 1; it <= a.maxit && resid > a.tol; ++it) {
    if (fresh) {
      float2* rr[5] = {r0, hq[3], hq[2], hq[1], hq[0]};
      float2* pp[5] = {p0, work, hs[2], hs[1], hs[0]};
      if (o.dot(rt, rr[0], rho)) return 184;
      for (int j = 1; j <= 4; ++j) {
        if ((rc = run_mv(c, 'N', pp[j - 1], pp[j], nc))) return rc;
        if ((rc = sig(pp[j], 184, "initial sigma zero"))) return rc;
        if (o.axpy(x, pp[0], alpha)) return 185;
        for (int i = 0; i < j; ++i)
          if (o.axpy(rr[i], pp[i + 1], -alpha)) return 185;
        if ((rc = run_mv(c, 'N', rr[j - 1], rr[j], nc))) return rc;
        if ((rc = rhobeta(rr[j], 185))) return rc;
        for (int i = 0; i <= j; ++i)
          if (o.affine(pp[i], rr[i], -beta)) return 186;
      }
      std::vector<std::pair<const float2*, const float2*>> q;
      for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) q.push_back({rr[i + 1], rr[j + 1]});
        q.push_back({rr[i + 1], rr[0]});
      }
      zc g[20];
      if (o.dotv(q, g)) return 187;
      for (auto& row : mat) for (auto& v : row) v = 0.0;
      for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) mat[i][j] = g[5 * i + j];
        rhs[i] = g[5 * i + 4];
      }
      if (!solve_small(&mat[0][0], rhs, 4, 5)) { set_error(nm, "initial 4x4 minimization breakdown"); return 187; }
      for (int i = 0; i < 4; ++i) zeta[i] = rhs[i];
      if (o.lc(z, {{rr[0], zeta[0]}, {rr[1], zeta[1]}, {rr[2], zeta[2]}, {rr[3], zeta[3]}}) || o.axpy(x, z, 1.0) ||
          o.lc(yv, {{rr[1], zeta[0]}, {rr[2], zeta[1]}, {rr[3], zeta[2]}, {rr[4], zeta[3]}}) ||
          o.lc(u, {{pp[1], zeta[0]}, {pp[2], zeta[1]}, {pp[3], zeta[2]}, {pp[4], zeta[3]}}) ||
          o.axpy(r0, yv, -1.0) || o.axpy(p0, u, -1.0) || restore())
        return 189;
      fresh = false;
    } else {
      if (o.dot(rt, r0, rho)) return 190;
      auto common = [&](float2* mvout) -> int {  // x, z, y, r0 updates shared by every step
        if (o.axpy(x, p0, alpha) || o.axpy(z, u, -alpha) || o.lc(yv, {{yv, 1.0}, {hq[0], -alpha}, {work, alpha}}) ||
            o.axpy(r0, work, -alpha))
          return 1;
        (void)mvout;
        return 0;
      };
      // j = 1
      if ((rc = run_mv(c, 'N', p0, work, nc))) return rc;
      if ((rc = sig(work, 190, "sigma1 zero"))) return rc;
      if (common(work)) return 191;
      for (int i = 0; i < 3; ++i) if (o.axpy(hs[i], hq[i + 1], -alpha)) return 191;
      if ((rc = run_mv(c, 'N', r0, hq[3], nc))) return rc;
      if ((rc = rhobeta(hq[3], 191))) return rc;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq[3], -beta)) return 192;
      for (int i = 0; i < 3; ++i) if (o.affine(hq[i], hs[i], -beta)) return 192;
      if (o.affine(u, yv, -beta)) return 192;
      // j = 2
      if ((rc = run_mv(c, 'N', work, hs[2], nc))) return rc;
      if ((rc = sig(hs[2], 193, "sigma2 zero"))) return rc;
      if (common(work) || o.axpy(hq[3], hs[2], -alpha)) return 194;
      for (int i = 0; i < 2; ++i) if (o.axpy(hs[i], hq[i + 1], -alpha)) return 194;
      if ((rc = run_mv(c, 'N', hq[3], hq[2], nc))) return rc;
      if ((rc = rhobeta(hq[2], 194))) return rc;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq[3], -beta) || o.affine(hs[2], hq[2], -beta)) return 195;
      for (int i = 0; i < 2; ++i) if (o.affine(hq[i], hs[i], -beta)) return 195;
      if (o.affine(u, yv, -beta)) return 195;
      // j = 3
      if ((rc = run_mv(c, 'N', hs[2], hs[1], nc))) return rc;
      if ((rc = sig(hs[1], 196, "sigma3 zero"))) return rc;
      if (common(work) || o.axpy(hq[3], hs[2], -alpha) || o.axpy(hq[2], hs[1], -alpha) || o.axpy(hs[0], hq[1], -alpha))
        return 197;
      if ((rc = run_mv(c, 'N', hq[2], hq[1], nc))) return rc;
      if ((rc = rhobeta(hq[1], 197))) return rc;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq[3], -beta) || o.affine(hs[2], hq[2], -beta) ||
          o.affine(hs[1], hq[1], -beta) || o.affine(hq[0], hs[0], -beta) || o.affine(u, yv, -beta))
        return 198;
      // j = 4
      if ((rc = run_mv(c, 'N', hs[1], hs[0], nc))) return rc;
      if ((rc = sig(hs[0], 199, "sigma4 zero"))) return rc;
      if (common(work) || o.axpy(hq[3], hs[2], -alpha) || o.axpy(hq[2], hs[1], -alpha) || o.axpy(hq[1], hs[0], -alpha))
        return 200;
      if ((rc = run_mv(c, 'N', hq[1], hq[0], nc))) return rc;
      if ((rc = rhobeta(hq[0], 200))) return rc;
      if (o.affine(p0, r0, -beta) || o.affine(work, hq[3], -beta) || o.affine(hs[2], hq[2], -beta) ||
          o.affine(hs[1], hq[1], -beta) || o.affine(hs[0], hq[0], -beta) || o.affine(u, yv, -beta))
        return 201;
      float2* rr[5] = {r0, hq[3], hq[2], hq[1], hq[0]};
      float2* pp[5] = {p0, work, hs[2], hs[1], hs[0]};
      std::vector<std::pair<const float2*, const float2*>> q;
      for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) q.push_back({rr[i + 1], rr[j + 1]});
        q.push_back({rr[i + 1], yv});
        q.push_back({rr[i + 1], rr[0]});
      }
      q.push_back({yv, yv});
      q.push_back({yv, rr[0]});
      zc g[26];
      if (o.dotv(q, g)) return 202;
      for (auto& row : mat) for (auto& v : row) v = 0.0;
      for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) mat[i][j] = g[6 * i + j];
        mat[i][4] = g[6 * i + 4];
        mat[4][i] = std::conj(mat[i][4]);
        rhs[i] = g[6 * i + 5];
      }
      mat[4][4] = g[24];
      rhs[4] = g[25];
      if (!solve_small(&mat[0][0], rhs, 5, 5)) { set_error(nm, "5x5 minimization breakdown"); return 202; }
      for (int i = 0; i < 4; ++i) zeta[i] = rhs[i];
      eta = rhs[4];
      if (o.lc(z, {{z, eta}, {rr[0], zeta[0]}, {rr[1], zeta[1]}, {rr[2], zeta[2]}, {rr[3], zeta[3]}}) ||
          o.axpy(x, z, 1.0) ||
          o.lc(yv, {{yv, eta}, {rr[1], zeta[0]}, {rr[2], zeta[1]}, {rr[3], zeta[2]}, {rr[4], zeta[3]}}) ||
          o.lc(u, {{u, eta}, {pp[1], zeta[0]}, {pp[2], zeta[1]}, {pp[3], zeta[2]}, {pp[4], zeta[3]}}) ||
          o.axpy(r0, yv, -1.0) || o.axpy(p0, u, -1.0) || restore())
        return 204;
    }
    if (o.norm(r0, rnorm)) return 205;
    resid = rnorm / bnorm;
    if ((it % RELIABLE_PERIOD) == 0 || resid <= a.tol) {
      Reliable rr;
      if ((rc = reliable_check(c, c.sb, x, r0, bnorm, resid, a.tol, nc, rr))) return rc > 1 ? rc : 206;
      if (rr.converged) { resid = rr.true_res; break; }
      if (rr.restart) {
        if (o.lc(r0, {{c.sb, 1.0}, {c.sy, -1.0}}) || o.copy(rt, r0) || o.copy(p0, r0) || o.zero(z)) return 206;
        fresh = true;
        resid = rr.true_res;
      }
    }
  }
  if ((rc = initial_residual(c, o, x, r0, nc))) return rc;
  if (o.norm(r0, rnorm)) return 206;
  resid = rnorm / bnorm;
  return solver_end(c, a, x, it <= a.maxit ? it : a.maxit, resid, nc, 207);
}

}  // namespace ddhip

using namespace ddhip;

#define DDH_SOLVER_ENTRY(name, fn)                                                                        \
  extern "C" DDSCAT_HIP_API int name(const void* b, void* x, const void* adia, const void* aoff, int ndim,  \
                                     int maxit, float tol, int* itno, float* tole, int* ncompte) {         \
    SolverArgs a{b, x, adia, aoff, ndim, maxit, tol, itno, tole, ncompte};                                 \
    return fn(ctx(), a);                                                                                   \
  }

DDH_SOLVER_ENTRY(ddscat_cuda_solve_bicgstab4_f32, solve_bicgstab4)
DDH_SOLVER_ENTRY(ddscat_cuda_solve_gpbicgstab2_f32, solve_gpbicgstab2)
DDH_SOLVER_ENTRY(ddscat_cuda_solve_gpbicgstab4_f32, solve_gpbicgstab4)
