// Host-side Krylov helpers shared by the adaptive exponential integrator
// (rfhip_api.hip) and its training labels (krylov_host.cpp). No HIP.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace rfhip {

void set_error(const std::string &msg);

// Reusable host buffers of the small-matrix exponentials.
struct KrylovScratch {
  std::vector<double> m, e, work;
};

// What the proposer's features read of the previous step.
// What the features take from the run so far. None of it depends on the
// controller's own choices of step or dimension (the previous step enters
// only as the interval of a decay rate), so a network trained on one
// controller's logs sees the same features under another.
struct ProposalHistory {
  double t_end = 0.0;      // the run's horizon
  double beta0 = 0.0;      // norm of the initial vector
  double h10_0 = 0.0;      // |H[1][0]| of the first step's basis
  double h_prev = 0.0;     // previous step, 0 before the first
  double beta_prev = 0.0;  // norm of the vector it propagated
};

constexpr int kProposalFeatures = 12;

// exp(a), n x n row-major, by the degree-13 Padé approximant with scaling
// and squaring.
void expm_pade13(const double *a, int n, double *out, std::vector<double> &work);
// beta * H[j][j-1] * |[exp(h H_j)]_{j-1,0}| for the leading j x j block of a
// Hessenberg matrix with row stride ld (row j must exist).
double krylov_estimate(const double *hess, int ld, int j, double beta, double h, KrylovScratch &ws);
// Largest h <= h_cap with krylov_estimate(h) <= rate * h.
double max_admissible_h(const double *hess, int ld, int j, double beta, double rate, double h_ref, double h_cap,
                        int iters, KrylovScratch &ws);
// The 12 proposal features (needs the first two Arnoldi columns), in units
// of the characteristic step hs.
void proposal_features(const double *hess, int ld, const ProposalHistory &hist, double beta, double tol,
                       double t_left, double hs, int64_t n_dof, float *f);

}  // namespace rfhip
