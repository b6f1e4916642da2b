// Host-side pieces of the adaptive Krylov exponential (no HIP): the Padé
// scaling-and-squaring exponential of the small Hessenberg matrix, the
// a-posteriori error estimate of a Krylov step, the proposer's features and
// the hindsight labels its training uses.
//
// The reference's exponential integrator takes a fixed step with a fixed or
// tolerance-stopped Krylov dimension (crates/rapidfem-td/src/propagator.rs:
// 186-326, session.rs:700-726) and uses a Taylor-18 exponential
// (propagator.rs:69-114, restated in expm_host for the expmv path). The
// adaptive (h, m) controller built on these pieces has no reference
// counterpart.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <thread>
#include <vector>

#include "rfhip.h"
#include "rfhip_krylov.hpp"

namespace rfhip {

namespace {

// o = x * y for n x n row-major matrices.
void matmul(const double *x, const double *y, double *o, int n) {
  std::fill(o, o + (size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      const double xik = x[(size_t)i * n + k];
      if (xik == 0.0) continue;
      const double *yr = y + (size_t)k * n;
      double *orow = o + (size_t)i * n;
      for (int j = 0; j < n; ++j) orow[j] += xik * yr[j];
    }
}

// Solves q x = p in place of p (n x n, n right-hand sides) by Gaussian
// elimination with partial pivoting; q is overwritten.
void solve_inplace(double *q, double *p, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    double best = std::fabs(q[(size_t)c * n + c]);
    for (int r = c + 1; r < n; ++r) {
      const double v = std::fabs(q[(size_t)r * n + c]);
      if (v > best) best = v, piv = r;
    }
    if (piv != c) {
      std::swap_ranges(q + (size_t)c * n, q + (size_t)(c + 1) * n, q + (size_t)piv * n);
      std::swap_ranges(p + (size_t)c * n, p + (size_t)(c + 1) * n, p + (size_t)piv * n);
    }
    const double inv = 1.0 / q[(size_t)c * n + c];
    for (int r = c + 1; r < n; ++r) {
      const double f = q[(size_t)r * n + c] * inv;
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) q[(size_t)r * n + k] -= f * q[(size_t)c * n + k];
      for (int k = 0; k < n; ++k) p[(size_t)r * n + k] -= f * p[(size_t)c * n + k];
    }
  }
  for (int c = n - 1; c >= 0; --c) {
    const double inv = 1.0 / q[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) p[(size_t)c * n + k] *= inv;
    for (int r = 0; r < c; ++r) {
      const double f = q[(size_t)r * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) p[(size_t)r * n + k] -= f * p[(size_t)c * n + k];
    }
  }
}

}  // namespace

// exp(a) for a row-major n x n matrix by the [13/13] Padé approximant with
// scaling and squaring (N. J. Higham, "The scaling and squaring method for
// the matrix exponential revisited", SIAM J. Matrix Anal. Appl. 26 (2005)
// 1179-1193, Algorithm 2.3 at degree 13 only): a / 2^s with
// ||a||_1 / 2^s <= theta_13, r13 = (V - U)^-1 (V + U) from the even and odd
// parts of the approximant, then s squarings.
void expm_pade13(const double *a, int n, double *out, std::vector<double> &work) {
  static const double b[14] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0,
                               1187353796428800.0,  129060195264000.0,   10559470521600.0,
                               670442572800.0,      33522128640.0,       1323241920.0,
                               40840800.0,          960960.0,            16380.0,
                               182.0,               1.0};
  const double theta13 = 5.371920351148152;
  const size_t nn = (size_t)n * n;
  double norm1 = 0.0;
  for (int j = 0; j < n; ++j) {
    double col = 0.0;
    for (int i = 0; i < n; ++i) col += std::fabs(a[(size_t)i * n + j]);
    norm1 = std::max(norm1, col);
  }
  int s = 0;
  if (norm1 > theta13) s = std::max(0, (int)std::ceil(std::log2(norm1 / theta13)));
  const double scale = std::ldexp(1.0, -s);
  work.resize(8 * nn);
  double *A = work.data(), *A2 = A + nn, *A4 = A2 + nn, *A6 = A4 + nn, *T1 = A6 + nn, *T2 = T1 + nn, *U = T2 + nn,
         *V = U + nn;
  for (size_t k = 0; k < nn; ++k) A[k] = a[k] * scale;
  matmul(A, A, A2, n);
  matmul(A2, A2, A4, n);
  matmul(A4, A2, A6, n);
  // U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I]
  for (size_t k = 0; k < nn; ++k) T1[k] = b[13] * A6[k] + b[11] * A4[k] + b[9] * A2[k];
  matmul(A6, T1, T2, n);
  for (size_t k = 0; k < nn; ++k) T2[k] += b[7] * A6[k] + b[5] * A4[k] + b[3] * A2[k];
  for (int i = 0; i < n; ++i) T2[(size_t)i * n + i] += b[1];
  matmul(A, T2, U, n);
  // V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
  for (size_t k = 0; k < nn; ++k) T1[k] = b[12] * A6[k] + b[10] * A4[k] + b[8] * A2[k];
  matmul(A6, T1, V, n);
  for (size_t k = 0; k < nn; ++k) V[k] += b[6] * A6[k] + b[4] * A4[k] + b[2] * A2[k];
  for (int i = 0; i < n; ++i) V[(size_t)i * n + i] += b[0];
  // (V - U) X = (V + U)
  for (size_t k = 0; k < nn; ++k) {
    T1[k] = V[k] - U[k];
    T2[k] = V[k] + U[k];
  }
  solve_inplace(T1, T2, n);
  for (int q = 0; q < s; ++q) {
    matmul(T2, T2, T1, n);
    std::swap(T1, T2);
  }
  std::copy(T2, T2 + nn, out);
}

// beta * hnext * |[exp(h H_j)]_{j-1,0}|, H_j the leading j x j block of the
// Hessenberg matrix h_full (row stride ld) and hnext = h_full[j][j-1]: the
// a-posteriori estimate of propagator.rs:278-294 (there relative, here
// scaled by the norm of the vector being propagated).
double krylov_estimate(const double *hess, int ld, int j, double beta, double h, KrylovScratch &ws) {
  const double hnext = hess[(size_t)j * ld + (j - 1)];
  if (hnext == 0.0) return 0.0;
  ws.m.resize((size_t)j * j);
  ws.e.resize((size_t)j * j);
  for (int a = 0; a < j; ++a)
    for (int c = 0; c < j; ++c) ws.m[(size_t)a * j + c] = h * hess[(size_t)a * ld + c];
  expm_pade13(ws.m.data(), j, ws.e.data(), ws.work);
  const double v = beta * hnext * std::fabs(ws.e[(size_t)(j - 1) * j]);
  return std::isfinite(v) ? v : std::numeric_limits<double>::infinity();
}

// Largest h in (0, h_cap] with estimate(h) <= rate * h, by bisection on log h
// (iters halvings of the bracket) after bracketing from h_ref by doubling
// or halving; 0 when even h_ref * 2^-60 fails. A breakdown column
// (hnext = 0) admits h_cap.
double max_admissible_h(const double *hess, int ld, int j, double beta, double rate, double h_ref, double h_cap,
                        int iters, KrylovScratch &ws) {
  if (hess[(size_t)j * ld + (j - 1)] == 0.0) return h_cap;
  auto ok = [&](double h) { return krylov_estimate(hess, ld, j, beta, h, ws) <= rate * h; };
  const double h0 = h_ref > 0.0 ? std::min(h_ref, h_cap) : h_cap;
  double lo, hi;
  if (ok(h0)) {
    lo = h0;
    for (;;) {
      if (lo >= h_cap) return h_cap;
      hi = std::min(2.0 * lo, h_cap);
      if (!ok(hi)) break;
      lo = hi;
    }
  } else {
    hi = h0;
    lo = 0.5 * h0;
    for (int k = 0; !ok(lo); ++k) {
      if (k >= 60) return 0.0;
      hi = lo;
      lo *= 0.5;
    }
  }
  for (int q = 0; q < iters; ++q) {
    const double mid = std::sqrt(lo * hi);
    if (ok(mid))
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Features of one step's proposal, from the first two Arnoldi columns of
// the step (hess with row stride ld; cols >= 2) and the run so far, in units
// of the run's characteristic step hs (the CFL step when the caller gives
// it): the decay of the propagated norm since the start, log |H[1][0]| hs,
// the first Hessenberg entries times hs, the decay rate over the previous
// step, the fraction of the run done, log |H[1][0]| hs of the first basis,
// the size of the system, the tolerance and the time left.
void proposal_features(const double *hess, int ld, const ProposalHistory &hist, double beta, double tol,
                       double t_left, double hs, int64_t n_dof, float *f) {
  const double h00 = hess[0], h01 = hess[1], h10 = hess[(size_t)ld], h11 = hess[(size_t)ld + 1],
               h21 = hess[(size_t)2 * ld + 1];
  auto lg = [](double x) { return std::log(std::max(x, 1e-300)); };
  f[0] = (float)(hist.beta0 > 0.0 ? lg(beta / hist.beta0) : 0.0);
  f[1] = (float)lg(std::fabs(h10) * hs);
  f[2] = (float)(h00 * hs);
  f[3] = (float)(h11 * hs);
  f[4] = (float)(h21 * hs);
  f[5] = (float)(h01 * hs);
  f[6] = (float)(hist.h_prev > 0.0 && hist.beta_prev > 0.0 ? lg(beta / hist.beta_prev) / (hist.h_prev / hs) : 0.0);
  f[7] = (float)(hist.t_end > 0.0 ? std::clamp((hist.t_end - t_left) / hist.t_end, 0.0, 1.0) : 0.0);
  f[8] = (float)lg((hist.h10_0 > 0.0 ? hist.h10_0 : std::fabs(h10)) * hs);
  f[9] = (float)(std::log10((double)std::max<int64_t>(n_dof, 1)) / 8.0);
  f[10] = (float)std::log10(tol);
  f[11] = (float)lg(std::min(t_left / hs, 1e6));
}

}  // namespace rfhip

using namespace rfhip;

extern "C" {

int rfhip_expm_pade(const double *a, int32_t n, double *out) {
  if (!a || !out || n < 1 || n > 4096) {
    set_error("rfhip_expm_pade: bad arguments");
    return RFHIP_EINVAL;
  }
  for (size_t k = 0; k < (size_t)n * n; ++k)
    if (!std::isfinite(a[k])) {
      set_error("rfhip_expm_pade: non-finite entry");
      return RFHIP_EINVAL;
    }
  std::vector<double> work;
  expm_pade13(a, n, out, work);
  return RFHIP_OK;
}

int rfhip_krylov_labels(const double *hess, int32_t m_log, int64_t count, const int32_t *dim_built, const double *beta,
                        const double *rate, const double *h_ref, const double *h_cap, const int32_t *m_grid,
                        int32_t n_grid, const double *cost_model, int32_t nthreads, double *h_best,
                        int32_t *m_best, double *eff_best) {
  if (!hess || !dim_built || !beta || !rate || !h_ref || !h_cap || !m_grid || !cost_model || !h_best || !m_best ||
      m_log < 1 || m_log > 1024 || count < 0 || n_grid < 1) {
    set_error("rfhip_krylov_labels: bad arguments");
    return RFHIP_EINVAL;
  }
  for (int g = 0; g < n_grid; ++g)
    if (m_grid[g] < 1 || m_grid[g] > m_log) {
      set_error("rfhip_krylov_labels: grid dimension outside [1, m_log]");
      return RFHIP_EINVAL;
    }
  const int nt = std::max(1, std::min<int>(nthreads, 64));
  std::atomic<int64_t> next{0};
  auto work = [&] {
    KrylovScratch ws;
    for (int64_t q; (q = next.fetch_add(1)) < count;) {
      const double *H = hess + (size_t)q * (m_log + 1) * m_log;
      double best_eff = -1.0, best_h = 0.0;
      int best_m = m_grid[0];
      for (int g = 0; g < n_grid; ++g) {
        const int j = m_grid[g];
        if (j > dim_built[q]) continue;
        const double hj = max_admissible_h(H, m_log, j, beta[q], rate[q], h_ref[q], h_cap[q], 12, ws);
        const double cost = cost_model[0] + cost_model[1] * j + cost_model[2] * 0.5 * j * (j + 1);
        const double eff = hj / cost;
        if (eff > best_eff) best_eff = eff, best_h = hj, best_m = j;
      }
      h_best[q] = best_h;
      m_best[q] = best_m;
      if (eff_best) eff_best[q] = best_eff;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  return RFHIP_OK;
}

}  // extern "C"
