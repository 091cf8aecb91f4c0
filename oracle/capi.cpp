// C entry points of the FP64 oracle, for ctypes (TEST INFRASTRUCTURE ONLY: tests/, smoke() and the
// cpu_baseline leg of bench.py).  Return codes: 0 ok, -1 RuntimeError, -2 Spline::OutOfBounds.
#include <cstring>
#include <memory>
#include <string>

#include "oracle.hpp"

using namespace orc;

namespace {
  thread_local std::string g_err;

  template <typename F>
  int guard(F&& f) {
    try {
      f();
      return 0;
    } catch (OutOfBounds const& e) {
      g_err = e.what();
      return -2;
    } catch (std::exception const& e) {
      g_err = e.what();
      return -1;
    }
  }

  struct ListH {
    Box box;
    List nl;
  };

  // Options as flat double arrays (the order of the fields of DimerOptions / RotorOptions).
  DimerOptions dimer_opts(double const* d) {
    DimerOptions opt;
    if (d) {
      opt.n = static_cast<int>(d[0]);
      opt.max_steps = static_cast<int>(d[1]);
      opt.hist_check_freq = static_cast<int>(d[2]);
      opt.f2norm = d[3];
      opt.skin_frac = d[4];
      opt.proj_tol = d[5];
      opt.max_trust = d[6];
      opt.min_trust = d[7];
      opt.grow_trust = d[8];
      opt.shrink_trust = d[9];
      opt.convex_max = static_cast<int>(d[10]);
      opt.use_history = d[11] != 0;
    }
    return opt;
  }

  RotorOptions rotor_opts(double const* r) {
    RotorOptions ropt;
    if (r) {
      ropt.n = static_cast<int>(r[0]);
      ropt.iter_max_rot = static_cast<int>(r[1]);
      ropt.delta_r = r[2];
      ropt.theta_tol = r[3];
      ropt.relax_in_convex = r[4] != 0;
    }
    return ropt;
  }

  std::vector<std::vector<double>> hist_of(double const* hist, int nhist, long len) {
    std::vector<std::vector<double>> hs;
    for (int k = 0; k < nhist; ++k) hs.emplace_back(hist + len * k, hist + len * (k + 1));
    return hs;
  }

  void put_stats(DimerStats const& st, long* stats, double* curv) {
    if (stats) {
      stats[0] = st.steps;
      stats[1] = st.grads;
      stats[2] = st.rebuilds;
    }
    if (curv) *curv = st.curv;
  }
}  // namespace

extern "C" {

char const* orc_last_error() { return g_err.c_str(); }

void* orc_box_new(double const* basis, int const* per) {
  Box* out = nullptr;
  guard([&] { out = new Box(basis, per); });
  return out;
}
void orc_box_free(void* b) { delete static_cast<Box*>(b); }
int orc_box_info(void* b, int* ortho, double* min_width) {
  auto* box = static_cast<Box*>(b);
  *ortho = box->ortho;
  *min_width = box->min_width();
  return 0;
}
int orc_box_canon(void* b, double const* x, long n, double* out) {
  auto* box = static_cast<Box*>(b);
  for (long i = 0; i < n; ++i) {
    V3 c = box->canon_image({x[3 * i], x[3 * i + 1], x[3 * i + 2]});
    for (int k = 0; k < 3; ++k) out[3 * i + k] = c[k];
  }
  return 0;
}
int orc_grid_info(void* b, double rc, int* shape, double* cell) {
  return guard([&] {
    Grid g(*static_cast<Box*>(b), rc);
    for (int k = 0; k < 3; ++k) {
      shape[k] = g.shape[k];
      cell[k] = g.cell[k];
    }
  });
}
int orc_grid_cell_idx(void* b, double rc, double const* x, long n, int* out) {
  return guard([&] {
    Grid g(*static_cast<Box*>(b), rc);
    for (long i = 0; i < n; ++i) out[i] = g.cell_idx({x[3 * i], x[3 * i + 1], x[3 * i + 2]});
  });
}

void* orc_list_new(void* b, double rc) {
  ListH* out = nullptr;
  guard([&] { out = new ListH{*static_cast<Box*>(b), List(*static_cast<Box*>(b), rc)}; });
  return out;
}
void orc_list_free(void* h) { delete static_cast<ListH*>(h); }
int orc_list_rebuild(void* h, double const* pos, long n, int threads) {
  return guard([&] { static_cast<ListH*>(h)->nl.rebuild(pos, n, threads); });
}
int orc_list_update(void* h, double const* d, long n) {
  return guard([&] { static_cast<ListH*>(h)->nl.update(d, n); });
}
long orc_list_num_ghosts(void* h) {
  auto& nl = static_cast<ListH*>(h)->nl;
  return nl.num_plus_ghosts() - nl.size();
}
long orc_list_count(void* h, double rc) {
  long tot = -1;
  guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    long t = 0;
    for (long i = 0; i < nl.size(); ++i) nl.for_neighbours(i, rc, [&](long, double, V3 const&) { ++t; });
    tot = t;
  });
  return tot;
}
// offsets[n+1]; j, shift (packed (s0+1)+3(s1+1)+9(s2+1)), dr[3*pairs]
int orc_list_export(void* h, double rc, long* offsets, int* j, int* shift, double* dr) {
  return guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    long t = 0;
    for (long i = 0; i < nl.size(); ++i) {
      offsets[i] = t;
      nl.for_neighbours_shift(i, rc, [&](long jj, int s, double, V3 const& d) {
        j[t] = static_cast<int>(jj);
        shift[t] = s;
        for (int k = 0; k < 3; ++k) dr[3 * t + k] = d[k];
        ++t;
      });
    }
    offsets[nl.size()] = t;
  });
}

void* orc_spline_new(double const* y, long n, double dx) {
  Spline* out = nullptr;
  guard([&] { out = new Spline(std::vector<double>(y, y + n), dx); });
  return out;
}
void orc_spline_free(void* s) { delete static_cast<Spline*>(s); }
namespace {
  int eval_spline(Spline const& sp, double const* x, long m, int deriv, double* out) {
    return guard([&] {
      for (long i = 0; i < m; ++i) out[i] = deriv == 0 ? sp.f(x[i]) : deriv == 1 ? sp.fp(x[i]) : sp.fpp(x[i]);
    });
  }
}  // namespace
int orc_spline_eval(void* s, double const* x, long m, int deriv, double* out) {
  return eval_spline(*static_cast<Spline*>(s), x, m, deriv, out);
}
// Coefficients (a,b,c,d) per segment; returns the segment count when out == nullptr.
long orc_spline_coeffs(void* s, double* out) {
  auto const& sp = static_cast<Spline*>(s)->spines();
  if (out)
    for (std::size_t i = 0; i < sp.size(); ++i) {
      out[4 * i] = sp[i].a;
      out[4 * i + 1] = sp[i].b;
      out[4 * i + 2] = sp[i].c;
      out[4 * i + 3] = sp[i].d;
    }
  return static_cast<long>(sp.size());
}

void* orc_data_new(char const* path) {
  DataEAM* out = nullptr;
  guard([&] { out = new DataEAM(path); });
  return out;
}
void orc_data_free(void* d) { delete static_cast<DataEAM*>(d); }
int orc_data_info(void* d, int* ntypes, double* rcut, double* masses) {
  auto* D = static_cast<DataEAM*>(d);
  *ntypes = static_cast<int>(D->num_types());
  *rcut = D->r_cut();
  if (masses)
    for (std::size_t i = 0; i < D->num_types(); ++i) masses[i] = D->mass(i);
  return 0;
}
// kind 0 = F_a, 1 = phi(a, b), 2 = V(a, b)
int orc_data_eval(void* d, int kind, int a, int b, double const* x, long m, int deriv, double* out) {
  auto* D = static_cast<DataEAM*>(d);
  Spline const& sp = kind == 0 ? D->f(static_cast<std::size_t>(a))
                     : kind == 1 ? D->phi(static_cast<std::size_t>(a), static_cast<std::size_t>(b))
                                 : D->v(static_cast<std::size_t>(a), static_cast<std::size_t>(b));
  return eval_spline(sp, x, m, deriv, out);
}

int orc_energy(void* d, void* h, std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, double* E) {
  return guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    if (nl.size() != n) throw RuntimeError("size mismatch");
    EAM pot(static_cast<DataEAM*>(d));
    *E = pot.energy(id, fzn, nl, threads);
  });
}
int orc_gradient(void* d, void* h, std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, double* g) {
  return guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    if (nl.size() != n) throw RuntimeError("size mismatch");
    EAM pot(static_cast<DataEAM*>(d));
    pot.gradient(g, id, fzn, nl, threads);
  });
}
int orc_hessian(void* d, void* h, std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, double* H) {
  return guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    EAM pot(static_cast<DataEAM*>(d));
    std::vector<double> tmp;
    pot.hessian(tmp, id, fzn, nl, threads);
    std::memcpy(H, tmp.data(), tmp.size() * sizeof(double));
  });
}
// y = H v, H from the dense analytic Hessian (lower triangle mirrored).
int orc_hvp(void* d, void* h, std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, double const* v, double* y) {
  return guard([&] {
    auto& nl = static_cast<ListH*>(h)->nl;
    EAM pot(static_cast<DataEAM*>(d));
    std::vector<double> tmp;
    pot.hessian(tmp, id, fzn, nl, threads);
    hessian_times(tmp, n, v, y);
  });
}

void* orc_steplbfgs_new(int m) { return new StepLBFGS(m); }
void orc_steplbfgs_free(void* s) { delete static_cast<StepLBFGS*>(s); }
void orc_steplbfgs_clear(void* s) { static_cast<StepLBFGS*>(s)->clear(); }
int orc_steplbfgs_step(void* s, double const* r, double const* g, long N, double* out) {
  return guard([&] {
    std::vector<double> rr(r, r + N), gg(g, g + N);
    auto& q = static_cast<StepLBFGS*>(s)->newton_step(rr, gg);
    std::memcpy(out, q.data(), static_cast<std::size_t>(N) * sizeof(double));
  });
}

// opts: n, iter_max, f2norm, skin_frac, proj_tol, max_trust, min_trust, grow_trust, shrink_trust
// Returns 0 converged, 1 not converged (the reference's `true`), <0 error.
int orc_minimise(void* b, void* d, double* r, double* g, std::uint32_t const* id, std::uint8_t const* fzn, long n,
                 int threads, double const* o, int* iters) {
  int fail = 0;
  int rc = guard([&] {
    LBFGSOptions opt;
    if (o) {
      opt.n = static_cast<int>(o[0]);
      opt.iter_max = static_cast<int>(o[1]);
      opt.f2norm = o[2];
      opt.skin_frac = o[3];
      opt.proj_tol = o[4];
      opt.max_trust = o[5];
      opt.min_trust = o[6];
      opt.grow_trust = o[7];
      opt.shrink_trust = o[8];
    }
    EAM pot(static_cast<DataEAM*>(d));
    std::vector<double> rr(r, r + 3 * n), gg;
    fail = lbfgs_minimise(opt, *static_cast<Box*>(b), pot, rr, gg, id, fzn, n, threads, iters);
    std::memcpy(r, rr.data(), rr.size() * sizeof(double));
    std::memcpy(g, gg.data(), gg.size() * sizeof(double));
  });
  return rc != 0 ? rc : fail;
}

// dopts: n, max_steps, hist_check_freq, f2norm, skin_frac, proj_tol, max_trust, min_trust, grow_trust,
//        shrink_trust, convex_max, use_history
// ropts: n, iter_max_rot, delta_r, theta_tol, relax_in_convex
// stats: steps, grads, rebuilds
int orc_find_sp(void* b, void* d, double* r, double* ax, std::uint32_t const* id, std::uint8_t const* fzn, long n,
                double const* in_min, double const* hist, int nhist, double theta_tol, int count_frozen, int threads,
                double const* dopts, double const* ropts, int* exit_code, long* stats, double* curv) {
  return guard([&] {
    DimerOptions opt = dimer_opts(dopts);
    RotorOptions ropt = rotor_opts(ropts);
    EAM pot(static_cast<DataEAM*>(d));
    std::vector<double> rr(r, r + 3 * n), aa(ax, ax + 3 * n);
    DimerStats st;
    Exit e = dimer_find_sp(opt, ropt, *static_cast<Box*>(b), pot, rr, aa, id, fzn, n, in_min, hist_of(hist, nhist, 3 * n),
                           theta_tol, count_frozen, threads, &st);
    std::memcpy(r, rr.data(), rr.size() * sizeof(double));
    std::memcpy(ax, aa.data(), aa.size() * sizeof(double));
    *exit_code = static_cast<int>(e);
    put_stats(st, stats, curv);
  });
}

// Many independent searches, one thread each (the reference's task-per-search structure,
// find.cpp:697-732 with threads = 1 inside each search, find.cpp:990).  r/ax hold k starts.
int orc_find_sp_many(void* b, void* d, double* r, double* ax, std::uint32_t const* id, std::uint8_t const* fzn,
                     long n, int k, double const* in_min, int count_frozen, int threads, double const* dopts,
                     double const* ropts, int* exit_codes, long* stats) {
  std::vector<int> rcs(static_cast<std::size_t>(k), 0);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int s = 0; s < k; ++s) {
    double c = 0;
    rcs[static_cast<std::size_t>(s)] = orc_find_sp(b, d, r + 3 * n * s, ax + 3 * n * s, id, fzn, n, in_min, nullptr, 0,
                                                   0.0, count_frozen, 1, dopts, ropts, exit_codes + s, stats + 3 * s, &c);
  }
  for (int v : rcs)
    if (v != 0) return v;
  return 0;
}

// Master::perturb with the oracle's Gaussian sampler.  nl must be a list over `r_in` with r_cut >= r_pert.
int orc_perturb(std::uint64_t const* seed, double stddev, double r_pert, void* h, long centre, double const* r_in,
                std::uint8_t const* fzn, long n, double* r_out, double* ax_out, double* sd_out) {
  return guard([&] {
    Xoshiro prng({seed[0], seed[1], seed[2], seed[3]});
    *sd_out = perturb(prng, stddev, r_pert, static_cast<ListH*>(h)->nl, centre, r_in, fzn, n, r_out, ax_out);
  });
}

int orc_xoshiro(std::uint64_t const* seed, long m, int jumps, std::uint64_t* out) {
  return guard([&] {
    Xoshiro g({seed[0], seed[1], seed[2], seed[3]});
    for (int j = 0; j < jumps; ++j) g.jump();
    for (long i = 0; i < m; ++i) out[i] = g();
  });
}

int orc_normals(std::uint64_t const* seed, long m, double mean, double sd, double* out) {
  return guard([&] {
    Xoshiro g({seed[0], seed[1], seed[2], seed[3]});
    Normal nm{mean, sd};
    for (long i = 0; i < m; ++i) out[i] = nm(g);
  });
}

// Rotor::eff_gradient on an EAM system (rotor.cpp:49-164): ax rotated in place, out = effective gradient.
int orc_rotor_eff_gradient(void* b, void* d, double const* r, double* ax, std::uint32_t const* id,
                           std::uint8_t const* fzn, long n, int count_frozen, int threads, double const* ropts,
                           double* out, double* curv, long* grads) {
  return guard([&] {
    EAM pot(static_cast<DataEAM*>(d));
    std::vector<double> aa(ax, ax + 3 * n), o;
    *curv = rotor_eff_gradient(rotor_opts(ropts), *static_cast<Box*>(b), pot, r, aa, id, fzn, n, count_frozen, threads,
                               o, grads);
    std::memcpy(ax, aa.data(), aa.size() * sizeof(double));
    std::memcpy(out, o.data(), o.size() * sizeof(double));
  });
}

// The rotor on E = 1/2 (x - c)^T A (x - c) (A: n3 x n3 row-major).
int orc_rotor_quadratic(long n3, double const* A, double const* c, double const* x, double* ax, int count_frozen,
                        double const* ropts, double* out, double* curv, long* grads) {
  return guard([&] {
    std::vector<double> aa(ax, ax + n3), o;
    *curv = rotor_quadratic(rotor_opts(ropts), n3, A, c, x, aa, count_frozen, o, grads);
    std::memcpy(ax, aa.data(), aa.size() * sizeof(double));
    std::memcpy(out, o.data(), o.size() * sizeof(double));
  });
}

int orc_dimer_quadratic(long n3, double const* A, double const* c, double* r, double* ax, double const* in_min,
                        double const* hist, int nhist, double theta_tol, int count_frozen, double const* dopts,
                        double const* ropts, int* exit_code, long* stats, double* curv) {
  return guard([&] {
    std::vector<double> rr(r, r + n3), aa(ax, ax + n3);
    DimerStats st;
    Exit e = dimer_quadratic(dimer_opts(dopts), rotor_opts(ropts), n3, A, c, rr, aa, in_min, hist_of(hist, nhist, n3),
                             theta_tol, count_frozen, &st);
    std::memcpy(r, rr.data(), rr.size() * sizeof(double));
    std::memcpy(ax, aa.data(), aa.size() * sizeof(double));
    *exit_code = static_cast<int>(e);
    put_stats(st, stats, curv);
  });
}

int orc_dimer_mueller_brown(double kz, double* r, double* ax, double const* in_min, double const* hist, int nhist,
                            double theta_tol, double const* dopts, double const* ropts, int* exit_code, long* stats,
                            double* curv) {
  return guard([&] {
    std::vector<double> rr(r, r + 3), aa(ax, ax + 3);
    DimerStats st;
    Exit e = dimer_mueller_brown(dimer_opts(dopts), rotor_opts(ropts), kz, rr, aa, in_min, hist_of(hist, nhist, 3),
                                 theta_tol, &st);
    std::memcpy(r, rr.data(), rr.size() * sizeof(double));
    std::memcpy(ax, aa.data(), aa.size() * sizeof(double));
    *exit_code = static_cast<int>(e);
    put_stats(st, stats, curv);
  });
}

int orc_mueller_brown(double const* x, double kz, double* e, double* g) {
  mueller_brown(x, kz, e, g);
  return 0;
}

}  // extern "C"
