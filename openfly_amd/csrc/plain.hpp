// Plain form of the minimiser and dimer: the reference's control flow on the host, every 3N vector
// resident on the device, vector algebra in fused kernels with device-side scalars (vecops.hpp).
#pragma once

#include <memory>

#include "vecops.hpp"

namespace flyg {

  struct DVec {
    DBuf<double> b;
    double* p(long n) {
      b.reserve(static_cast<std::size_t>(n));
      return b.get();
    }
    double* get() const { return b.get(); }
  };

  // StepLBFGS (lbfgs_core.hpp:37-119, lbfgs_core.cpp:29-87) with history on the device.
  class DevLBFGS {
  public:
    DevLBFGS(flyg_ctx_s* ctx, int m);
    void clear() { m_k = 0; }
    // Returns q ~ H^-1 g (device, length N); x and g are device vectors of length N.
    double* newton_step(double const* x, double const* g, long N);

  private:
    flyg_ctx_s* m_ctx;
    int m_m;
    int m_k = 0;
    long m_N = -1;
    std::vector<DVec> m_s, m_y;
    DVec m_px, m_pg, m_q, m_r;
    DBuf<LBScalars> m_sc;
    DBuf<double> m_slots;
  };

  // Host-side copy of a small scalar result (synchronises the stream).
  double read_scalar(flyg_ctx_s& ctx, double const* d, int count = 1, double* out = nullptr);

}  // namespace flyg

struct flyg_steplbfgs_s {
  flyg_ctx_s* ctx;
  std::unique_ptr<flyg::DevLBFGS> core;
  flyg::DVec x, g;
};

struct flyg_lbfgs_s {
  flyg_ctx_s* ctx;
  flyg_lbfgs_options opt;
  flyg::BoxDesc box;
  double basis[9];
  int periodic[3];
  std::unique_ptr<flyg_list_s> nl;
  std::unique_ptr<flyg::DevLBFGS> core;
  flyg::DVec x, g, tmp;
  flyg::DBuf<double> slots;
  flyg_step_observer obs = nullptr;   // per-iteration positions to the caller (Options::fout)
  void* obs_user = nullptr;
  std::vector<double> obs_host;
};

struct flyg_dimer_s {
  flyg_ctx_s* ctx;
  flyg_dimer_options opt;
  flyg_rotor_options ropt;
  double basis[9];
  int periodic[3];
  std::unique_ptr<flyg_list_s> nl;
  std::unique_ptr<flyg::DevLBFGS> core;    // translation, m = opt.n
  std::unique_ptr<flyg::DevLBFGS> rcore;   // rotation, m = ropt.n
  flyg::DVec x, ax, eff, g0, g1, g1p, dg, delta_g, axisp, delta_prev, u, tmp, inmin;
  std::vector<flyg::DVec> hist_d;   // sp_k - in_min
  flyg::DBuf<double> slots;
  long grads = 0;
  flyg_step_observer obs = nullptr;
  void* obs_user = nullptr;
  std::vector<double> obs_host;
};

// Rotor (rotor.hpp:42-139): the plain form's rotor state (the dimer's, without its own list).
struct flyg_rotor_s {
  flyg_ctx_s* ctx;
  flyg_dimer_s d;
};

namespace flyg {
  double rotor_eff_gradient(flyg_dimer_s& d, flyg_list_s& nl, flyg_eam_s& e, uint32_t const* t, uint8_t const* f,
                            long n, double* d_axis, double* d_eff, int count_frozen);
  bool lbfgs_minimise(flyg_lbfgs_s& m, flyg_eam_s& e, double* d_pos, double* d_grad, uint32_t const* d_types,
                      uint8_t const* d_frozen, long n, int* iterations);
  int dimer_find_sp(flyg_dimer_s& d, flyg_eam_s& e, double* d_pos, double* d_axis, uint32_t const* d_types,
                    uint8_t const* d_frozen, long n, double const* d_in_min, double const* d_hist, int n_hist,
                    double theta_tol, int count_frozen, flyg_search_stats* st);
  std::unique_ptr<flyg_list_s> make_list(flyg_ctx_s* ctx, double const* basis, int const* periodic, double rc);
}  // namespace flyg
