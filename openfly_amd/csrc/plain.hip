// Plain-form StepLBFGS, LBFGS::minimise, Rotor::eff_gradient and Dimer::find_sp (FP64).
//
// Control flow restates lbfgs_core.cpp:29-87, lbfgs.cpp:29-115, rotor.cpp:33-164 and dimer.cpp:28-171;
// scalars that steer the host-side control flow (norms, curvatures, projections) are read back once
// per decision, everything else stays on the device.
#include <cmath>
#include <utility>

#include "plain.hpp"

namespace flyg {

  double read_scalar(flyg_ctx_s& ctx, double const* d, int count, double* out) {
    HIPCHK(hipMemcpyAsync(ctx.hscal, d, count * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    if (out)
      for (int k = 0; k < count; ++k) out[k] = ctx.hscal[k];
    return ctx.hscal[0];
  }

  namespace {
    void copy(flyg_ctx_s& ctx, double* dst, double const* src, long n) {
      HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx.stream));
    }

    // Options::fout of the reference (lbfgs.cpp:75-77, dimer.cpp:88-90): the positions at the start of
    // iteration `it`, copied to the host, go to the caller's observer.
    void observe(flyg_ctx_s& ctx, flyg_step_observer fn, void* user, int it, double const* d_x, long n,
                 std::vector<double>& host) {
      if (!fn) return;
      host.resize(static_cast<std::size_t>(3 * n));
      HIPCHK(hipMemcpyAsync(host.data(), d_x, 3 * n * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
      HIPCHK(hipStreamSynchronize(ctx.stream));
      if (int const rc = fn(user, it, host.data(), n)) {
        throw Error(FLYG_STOPPED, "the step observer stopped the run at iteration " + std::to_string(it) +
                                      " (returned " + std::to_string(rc) + ")");
      }
    }

    VecArgs lin(long n, double* out, double const* a, Coef ca, double const* b = nullptr, Coef cb = {},
                double const* c = nullptr, Coef cc = {}) {
      VecArgs v;
      v.n = n;
      v.out = out;
      v.v[0] = a;
      v.c[0] = ca;
      v.v[1] = b;
      v.c[1] = cb;
      v.v[2] = c;
      v.c[2] = cc;
      return v;
    }

    VecArgs dots(long n, double* slots, std::initializer_list<std::pair<double const*, double const*>> xy) {
      VecArgs v;
      v.n = n;
      v.slots = slots;
      for (auto const& p : xy) {
        v.X[v.ndot] = p.first;
        v.Y[v.ndot] = p.second;
        ++v.ndot;
      }
      return v;
    }

    // Adds dot products (null = the output vector) to a lincomb.
    VecArgs with_dots(VecArgs v, double* slots, std::initializer_list<std::pair<double const*, double const*>> xy) {
      v.slots = slots;
      for (auto const& p : xy) {
        v.X[v.ndot] = p.first;
        v.Y[v.ndot] = p.second;
        ++v.ndot;
      }
      return v;
    }
  }  // namespace

  // ================================================================================================
  // DevLBFGS (two-loop recursion, Nocedal & Wright Alg. 7.4; rho = 1/|s.y|; H0 = (1/rho)/(y.y) from the
  // newest pair; the first call returns g).
  // ================================================================================================

  DevLBFGS::DevLBFGS(flyg_ctx_s* ctx, int m) : m_ctx(ctx), m_m(m), m_s(static_cast<std::size_t>(m)), m_y(static_cast<std::size_t>(m)) {
    require(m > 0 && m <= 32, "L-BFGS history length must be in [1, 32]");
    m_sc.reserve(1);
    m_slots.reserve(8);
  }

  double* DevLBFGS::newton_step(double const* x, double const* g, long N) {
    flyg_ctx_s& ctx = *m_ctx;
    if (m_k > 0) require(N == m_N, "Call to newton_step() with a different number of atoms than the history");
    if (N != m_N) {
      for (auto& v : m_s) v.p(N);
      for (auto& v : m_y) v.p(N);
      m_px.p(N);
      m_pg.p(N);
      m_q.p(N);
      m_r.p(N);
      m_N = N;
    }
    LBScalars* sc = m_sc.get();
    double* slots = m_slots.get();
    int prev = m_k > 0 ? (m_k - 1) % m_m : 0;
    if (m_k > 0) {
      vec(ctx, lin(N, m_s[prev].get(), x, cst(1), m_px.get(), cst(-1)));
      VecArgs a = with_dots(lin(N, m_y[prev].get(), g, cst(1), m_pg.get(), cst(-1)), slots,
                            {{m_s[prev].get(), nullptr}, {nullptr, nullptr}});
      a.post = POST_HIST;
      a.pj = prev;
      a.lb = sc;
      vec(ctx, a);
    }
    copy(ctx, m_px.get(), x, N);
    copy(ctx, m_pg.get(), g, N);
    copy(ctx, m_q.get(), g, N);
    int incur = m_k <= m_m ? 0 : m_k - m_m;
    int bound = m_k <= m_m ? m_k : m_m;
    double* q = m_q.get();
    double* r = m_r.get();
    int jprev = -1;
    for (int i = bound - 1; i >= 0; --i) {   // loop 1: newest -> oldest
      int j = (i + incur) % m_m;
      VecArgs a;
      if (jprev < 0) {
        a = dots(N, slots, {{m_s[j].get(), q}});
      } else {
        a = with_dots(lin(N, q, q, cst(1), m_y[jprev].get(), dev(&sc->alpha[jprev], -1.0)), slots, {{m_s[j].get(), nullptr}});
      }
      a.post = POST_ALPHA;
      a.pj = j;
      a.lb = sc;
      vec(ctx, a);
      jprev = j;
    }
    if (bound == 0) {
      copy(ctx, r, q, N);
    } else {
      vec(ctx, lin(N, q, q, cst(1), m_y[jprev].get(), dev(&sc->alpha[jprev], -1.0)));
      int j0 = incur % m_m;
      VecArgs a = with_dots(lin(N, r, q, dev(&sc->h0)), slots, {{m_y[j0].get(), nullptr}});
      a.post = POST_BETA;
      a.pj = j0;
      a.lb = sc;
      vec(ctx, a);
      for (int i = 0; i < bound; ++i) {   // loop 2: oldest -> newest
        int j = (i + incur) % m_m;
        VecArgs b = lin(N, r, r, cst(1), m_s[j].get(), dev(&sc->coef[j]));
        if (i + 1 < bound) {
          int jn = (i + 1 + incur) % m_m;
          b = with_dots(b, slots, {{m_y[jn].get(), nullptr}});
          b.post = POST_BETA;
          b.pj = jn;
          b.lb = sc;
        }
        vec(ctx, b);
      }
    }
    ++m_k;
    return r;
  }

  std::unique_ptr<flyg_list_s> make_list(flyg_ctx_s* ctx, double const* basis, int const* periodic, double rc) {
    auto nl = std::make_unique<flyg_list_s>();
    nl->ctx = ctx;
    nl->box = make_box(basis, periodic);
    nl->grid = make_grid(nl->box, rc);
    nl->rc = rc;
    return nl;
  }

  namespace {
    void check_flags(flyg_ctx_s& ctx) {
      unsigned fl = ctx.take_flags();
      if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
      if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
    }
  }  // namespace

  // ================================================================================================
  // LBFGS::minimise (lbfgs.cpp:29-115)
  // ================================================================================================

  bool lbfgs_minimise(flyg_lbfgs_s& m, flyg_eam_s& e, double* d_pos, double* d_grad, uint32_t const* d_types,
                      uint8_t const* d_frozen, long n, int* iterations) {
    flyg_ctx_s& ctx = *m.ctx;
    flyg_lbfgs_options const& o = m.opt;
    long N = 3 * n;
    double skin = std::fmax(std::pow(o.skin_frac, 1. / 3.) - 1, 0.0) * e.tab.rc;
    double rc = e.tab.rc + skin;
    if (!m.nl || m.nl->rc != rc) m.nl = make_list(&ctx, m.basis, m.periodic, rc);
    if (!m.core) m.core = std::make_unique<DevLBFGS>(&ctx, o.n);
    m.core->clear();
    m.slots.reserve(8);
    double* x = d_pos;
    double* g = d_grad;
    double* sl = m.slots.get();

    list_rebuild(*m.nl, x, n);
    eam_gradient(e, *m.nl, d_types, d_frozen, g, FLYG_FP64);
    vec(ctx, dots(N, sl, {{g, g}}));
    double mag_g = read_scalar(ctx, sl);
    check_flags(ctx);
    double trust = o.min_trust, acc = 0;
    for (int i = 0; i < o.iter_max; ++i) {
      if (iterations) *iterations = i;
      observe(ctx, m.obs, m.obs_user, i, x, n, m.obs_host);
      if (mag_g < o.f2norm * o.f2norm) return false;
      double* q = m.core->newton_step(x, g, N);
      vec(ctx, dots(N, sl, {{q, q}}));
      double nrm = std::sqrt(read_scalar(ctx, sl));
      double scale = std::fmin(1.0, trust / nrm);
      vec(ctx, lin(N, q, q, cst(scale)));
      acc += nrm * scale;
      vec(ctx, lin(N, x, x, cst(1), q, cst(-1)));
      if (acc > 0.5 * skin) {
        list_rebuild(*m.nl, x, n);
        acc = 0;
      } else {
        list_update(*m.nl, q, n);
      }
      eam_gradient(e, *m.nl, d_types, d_frozen, g, FLYG_FP64);
      vec(ctx, dots(N, sl, {{g, q}, {g, g}}));
      double two[2];
      read_scalar(ctx, sl, 2, two);
      check_flags(ctx);
      double proj = two[0];
      mag_g = two[1];
      if (proj < -o.proj_tol) {
        trust = std::fmax(o.min_trust, o.shrink_trust * trust);
      } else if (proj > o.proj_tol) {
        trust = std::fmin(o.max_trust, o.grow_trust * trust);
      }
    }
    if (iterations) *iterations = o.iter_max;
    return true;
  }

  // ================================================================================================
  // Rotor::eff_gradient (rotor.cpp:49-164) and Dimer::find_sp (dimer.cpp:28-171)
  // ================================================================================================

  namespace {

    void gradient_at(flyg_dimer_s& d, flyg_list_s& nl, flyg_eam_s& e, uint32_t const* t, uint8_t const* f, double* out) {
      eam_gradient(e, nl, t, f, out, FLYG_FP64);
      ++d.grads;
    }

  }  // namespace

  // Rotor::eff_gradient on the list nl (ready at the positions; moved by +-delta_r and back, never rebuilt).
  // Axis in/out in d.ax, effective gradient to d.eff.  Returns the curvature; leaves |g_eff|^2 in d.slots[0].
  double eff_gradient(flyg_dimer_s& d, flyg_list_s& nl, flyg_eam_s& e, uint32_t const* t, uint8_t const* f, long n,
                      int count_frozen) {
      flyg_ctx_s& ctx = *d.ctx;
      flyg_rotor_options const& o = d.ropt;
      long N = 3 * n;
      double const dr = o.delta_r;
      double* sl = d.slots.get();
      double* ax = d.ax.get();
      double* g0 = d.g0.get();
      double* g1 = d.g1.get();
      double* g1p = d.g1p.get();
      double* dg = d.dg.get();
      double* dgp = d.delta_g.get();
      double* axp = d.axisp.get();
      double* dprev = d.delta_prev.get();
      double* u = d.u.get();
      double const inv_n = 1.0 / static_cast<double>(n);

      d.rcore->clear();
      gradient_at(d, nl, e, t, f, g0);
      vec(ctx, lin(N, dprev, ax, cst(-dr)));
      list_update(nl, dprev, n);
      gradient_at(d, nl, e, t, f, g1);

      double curv = 0;
      for (int i = 0;; ++i) {
        // Torque: delta_g = (g1 - g0) - ((g1 - g0).n) n, minus its mean when nothing is frozen.
        vec(ctx, with_dots(lin(N, dg, g1, cst(1), g0, cst(-1)), sl + 0, {{nullptr, ax}}));
        VecArgs a = lin(N, dgp, dg, cst(1), ax, dev(sl + 0, -1.0));
        if (count_frozen == 0) {
          a.comp_sums = 1;
          a.slots = sl + 4;
        }
        vec(ctx, a);
        if (count_frozen == 0) {
          VecArgs b = lin(N, dgp, nullptr, cst(0));
          b.mean3 = sl + 4;
          b.mean_k = -inv_n;
          vec(ctx, b);
        }
        double* theta = d.rcore->newton_step(ax, dgp, N);
        vec(ctx, dots(N, sl + 0, {{theta, ax}}));
        vec(ctx, with_dots(lin(N, theta, theta, cst(1), ax, dev(sl + 0, -1.0)), sl + 1, {{nullptr, nullptr}}));
        vec(ctx, with_dots(lin(N, theta, theta, dev_rsqrt(sl + 1)), sl + 2, {{dg, nullptr}, {dg, ax}}));
        double bc[2];
        read_scalar(ctx, sl + 2, 2, bc);
        double b_1 = bc[0] / dr;
        double c_x0 = bc[1] / dr;
        double theta_1 = -0.5 * std::atan(b_1 / std::abs(c_x0));
        if (std::abs(theta_1) < o.theta_tol || (i >= o.iter_max_rot && c_x0 < 0) || i > 10 * o.iter_max_rot) {
          curv = c_x0;
          break;
        }
        // Trial rotation; the list moves from the previous end to the trial end.
        vec(ctx, lin(N, axp, ax, cst(std::cos(theta_1)), theta, cst(std::sin(theta_1))));
        vec(ctx, lin(N, u, axp, cst(-dr), dprev, cst(-1)));
        vec(ctx, lin(N, dprev, axp, cst(-dr)));
        list_update(nl, u, n);
        gradient_at(d, nl, e, t, f, g1p);
        vec(ctx, with_dots(lin(N, u, g1p, cst(1), g0, cst(-1)), sl + 0, {{nullptr, axp}}));
        double c_x1 = read_scalar(ctx, sl + 0) / dr;
        double a_1 = (c_x0 - c_x1 + b_1 * std::sin(2 * theta_1)) / (1 - std::cos(2 * theta_1));
        double theta_min = 0.5 * std::atan(b_1 / a_1);
        if (a_1 * std::cos(2 * theta_min) - a_1 + b_1 * std::sin(2 * theta_min) > 0) theta_min += M_PI / 2;
        vec(ctx, lin(N, ax, ax, cst(std::cos(theta_min)), theta, cst(std::sin(theta_min))));
        double w1 = std::sin(theta_1 - theta_min) / std::sin(theta_1);
        double w2 = std::sin(theta_min) / std::sin(theta_1);
        double w3 = 1 - std::cos(theta_min) - std::sin(theta_min) * std::tan(0.5 * theta_1);
        vec(ctx, lin(N, g1, g1, cst(w1), g1p, cst(w2), g0, cst(w3)));
        if (std::abs(theta_min) < o.theta_tol) {
          vec(ctx, with_dots(lin(N, dg, g1, cst(1), g0, cst(-1)), sl + 0, {{nullptr, ax}}));
          curv = read_scalar(ctx, sl + 0) / dr;
          break;
        }
      }
      // Move the list back to the centre.
      vec(ctx, lin(N, u, dprev, cst(-1)));
      list_update(nl, u, n);

      vec(ctx, dots(N, sl + 0, {{g0, ax}}));
      double k = (!o.relax_in_convex && curv > 0) ? -1.0 : -2.0;
      double* eff = d.eff.get();
      if (!o.relax_in_convex && curv > 0) {
        VecArgs a = lin(N, eff, ax, dev(sl + 0, k));
        if (count_frozen == 0) {
          a.comp_sums = 1;
          a.slots = sl + 4;
        }
        vec(ctx, a);
      } else {
        VecArgs a = lin(N, eff, g0, cst(1), ax, dev(sl + 0, k));
        if (count_frozen == 0) {
          a.comp_sums = 1;
          a.slots = sl + 4;
        }
        vec(ctx, a);
      }
      if (count_frozen == 0) {
        VecArgs b = lin(N, eff, nullptr, cst(0));
        b.mean3 = sl + 4;
        b.mean_k = -inv_n;
        vec(ctx, with_dots(b, sl + 0, {{nullptr, nullptr}}));
      } else {
        vec(ctx, dots(N, sl + 0, {{eff, eff}}));
      }
      return curv;
    }

  double rotor_eff_gradient(flyg_dimer_s& d, flyg_list_s& nl, flyg_eam_s& e, uint32_t const* t, uint8_t const* f,
                            long n, double* d_axis, double* d_eff, int count_frozen) {
    flyg_ctx_s& ctx = *d.ctx;
    require(nl.built, "List used before rebuild");
    require(nl.n == n, "rotor size mismatch");
    require(nl.rc >= e.tab.rc + 2 * d.ropt.delta_r, "the list's cut-off is below the rotor's r_cut()");
    long N = 3 * n;
    if (!d.rcore) d.rcore = std::make_unique<DevLBFGS>(&ctx, d.ropt.n);
    d.slots.reserve(16);
    for (DVec* v : {&d.ax, &d.eff, &d.g0, &d.g1, &d.g1p, &d.dg, &d.delta_g, &d.axisp, &d.delta_prev, &d.u}) v->p(N);
    copy(ctx, d.ax.get(), d_axis, N);
    double curv = eff_gradient(d, nl, e, t, f, n, count_frozen);
    copy(ctx, d_axis, d.ax.get(), N);
    copy(ctx, d_eff, d.eff.get(), N);
    check_flags(ctx);
    return curv;
  }

  int dimer_find_sp(flyg_dimer_s& d, flyg_eam_s& e, double* d_pos, double* d_axis, uint32_t const* t,
                    uint8_t const* f, long n, double const* d_in_min, double const* d_hist, int n_hist,
                    double theta_tol, int count_frozen, flyg_search_stats* st) {
    flyg_ctx_s& ctx = *d.ctx;
    flyg_dimer_options const& o = d.opt;
    long N = 3 * n;
    double rot_rc = e.tab.rc + d.ropt.delta_r * 2;   // rotor.hpp:123
    double skin = std::fmax(std::pow(o.skin_frac, 1. / 3.) - 1, 0.0) * rot_rc;
    double rc = rot_rc + skin;
    if (!d.nl || d.nl->rc != rc) d.nl = make_list(&ctx, d.basis, d.periodic, rc);
    if (!d.core) d.core = std::make_unique<DevLBFGS>(&ctx, o.n);
    if (!d.rcore) d.rcore = std::make_unique<DevLBFGS>(&ctx, d.ropt.n);
    d.core->clear();
    d.slots.reserve(16);
    for (DVec* v : {&d.x, &d.ax, &d.eff, &d.g0, &d.g1, &d.g1p, &d.dg, &d.delta_g, &d.axisp, &d.delta_prev, &d.u, &d.tmp})
      v->p(N);
    double* sl = d.slots.get();
    double* x = d.x.get();
    copy(ctx, x, d_pos, N);
    copy(ctx, d.ax.get(), d_axis, N);
    d.grads = 0;

    // History of known saddles, relative to the start minimum.
    std::vector<double> sn(static_cast<std::size_t>(n_hist));
    if (o.use_history && n_hist > 0) {
      if (static_cast<int>(d.hist_d.size()) < n_hist) d.hist_d.resize(static_cast<std::size_t>(n_hist));
      for (int k = 0; k < n_hist; ++k) {
        vec(ctx, with_dots(lin(N, d.hist_d[k].p(N), d_hist + N * k, cst(1), d_in_min, cst(-1)), sl, {{nullptr, nullptr}}));
        sn[k] = std::sqrt(read_scalar(ctx, sl));
      }
    }

    int rebuilds = 0;
    list_rebuild(*d.nl, x, n);
    double curv = eff_gradient(d, *d.nl, e, t, f, n, count_frozen);
    double mag_g = read_scalar(ctx, sl);
    check_flags(ctx);
    double trust = o.min_trust, acc = 0, convex_count = 0;
    int ret = FLYG_EXIT_ITER_MAX, steps = o.max_steps;
    for (int i = 0; i < o.max_steps; ++i) {
      observe(ctx, d.obs, d.obs_user, i, x, n, d.obs_host);
      if (mag_g < o.f2norm * o.f2norm) {
        ret = FLYG_EXIT_SUCCESS;
        steps = i;
        break;
      } else if (convex_count >= o.convex_max) {
        ret = FLYG_EXIT_CONVEX;
        steps = i;
        break;
      }
      if (o.use_history && n_hist > 0 && i % o.hist_check_freq == 0) {
        double* dx = d.tmp.get();
        vec(ctx, with_dots(lin(N, dx, x, cst(1), d_in_min, cst(-1)), sl, {{nullptr, nullptr}}));
        double on = std::sqrt(read_scalar(ctx, sl));
        double ctheta = -1;
        bool hit = false;
        for (int k = 0; k < n_hist && !hit; ++k) {
          vec(ctx, dots(N, sl, {{dx, d.hist_d[k].get()}}));
          double dt = read_scalar(ctx, sl);
          ctheta = std::fmax(ctheta, dt / (sn[k] * on));
          if (ctheta > std::cos(theta_tol)) hit = true;
        }
        if (hit) {
          ret = FLYG_EXIT_COLLISION;
          steps = i;
          break;
        }
      }
      double* q = d.core->newton_step(x, d.eff.get(), N);
      vec(ctx, dots(N, sl, {{q, q}}));
      double nrm = std::sqrt(read_scalar(ctx, sl));
      double scale = std::fmin(1.0, trust / nrm);
      vec(ctx, lin(N, q, q, cst(scale)));
      acc += nrm * scale;
      vec(ctx, lin(N, x, x, cst(1), q, cst(-1)));
      if (acc > 0.5 * skin) {
        list_rebuild(*d.nl, x, n);
        acc = 0;
        ++rebuilds;
      } else {
        list_update(*d.nl, q, n);
      }
      curv = eff_gradient(d, *d.nl, e, t, f, n, count_frozen);
      convex_count = curv > 0 ? convex_count + 1 : 0;
      vec(ctx, dots(N, sl + 8, {{d.eff.get(), q}}));
      double two[2];
      read_scalar(ctx, sl, 1, two);
      mag_g = two[0];
      double proj = read_scalar(ctx, sl + 8);
      check_flags(ctx);
      if (proj < -o.proj_tol) {
        trust = std::fmax(o.min_trust, o.shrink_trust * trust);
      } else if (proj > o.proj_tol) {
        trust = std::fmin(o.max_trust, o.grow_trust * trust);
      }
    }
    copy(ctx, d_pos, x, N);
    copy(ctx, d_axis, d.ax.get(), N);
    HIPCHK(hipStreamSynchronize(ctx.stream));
    if (st) {
      st->exit_code = ret;
      st->steps = steps;
      st->grads = d.grads;
      st->rebuilds = rebuilds;
      st->curv = curv;
      st->energy = 0;
    }
    return ret;
  }

}  // namespace flyg
