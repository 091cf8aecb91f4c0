// Device vector engine for the plain-form minimiser and dimer: one fused kernel computes
//   out = c0 v0 + c1 v1 + c2 v2 (+ per-component mean correction)
// and up to four dot products (or the three component sums of `out`), reduced deterministically
// (fixed grid, per-block trees, the last block sums the block partials in order) into device scalar
// slots, followed by an optional scalar post-op (the L-BFGS two-loop coefficients), so whole chains
// of vector algebra run without host synchronisation.
#pragma once

#include "eam_host.hpp"

namespace flyg {

  // Coefficient value = k * f(*s) with f = identity, or 1/sqrt when inv_sqrt (s == nullptr: value = k).
  struct Coef {
    double k = 0;
    double const* s = nullptr;
    int inv_sqrt = 0;
  };

  inline Coef cst(double k) { return Coef{k, nullptr, 0}; }
  inline Coef dev(double const* s, double k = 1.0) { return Coef{k, s, 0}; }
  inline Coef dev_rsqrt(double const* s, double k = 1.0) { return Coef{k, s, 1}; }

  // Post-ops applied by the finalising block (acc = the reduced sums):
  enum PostOp : int {
    POST_NONE = 0,
    POST_HIST = 1,    // lb.rho[j] = 1/|acc0|; lb.h0 = (1/rho[j]) / acc1
    POST_ALPHA = 2,   // lb.alpha[j] = lb.rho[j] * acc0
    POST_BETA = 3,    // lb.coef[j] = lb.alpha[j] - lb.rho[j] * acc0
  };

  // Scalars of one device L-BFGS core.
  struct LBScalars {
    double rho[32];
    double alpha[32];
    double coef[32];
    double h0;
  };

  struct VecArgs {
    long n = 0;
    double* out = nullptr;
    double const* v[3] = {nullptr, nullptr, nullptr};
    Coef c[3];
    double const* mean3 = nullptr;   // out[i] += mean_k * mean3[i % 3]
    double mean_k = 0;
    // dots: acc_k = sum X_k[i] Y_k[i]; a null X_k/Y_k means `out`.  ndot = number of dots.
    int ndot = 0;
    double const* X[4] = {nullptr, nullptr, nullptr, nullptr};
    double const* Y[4] = {nullptr, nullptr, nullptr, nullptr};
    int comp_sums = 0;   // acc_{0,1,2} = component sums of out (instead of dots)
    double* slots = nullptr;   // receives the reduced sums (4 doubles)
    int post = POST_NONE;
    int pj = 0;
    LBScalars* lb = nullptr;
  };

  void vec(flyg_ctx_s& ctx, VecArgs const& a);

}  // namespace flyg
