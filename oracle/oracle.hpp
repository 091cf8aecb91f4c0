// FP64 CPU oracle for the openFLY saddle-search hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a clean-room restatement of the behaviour of the reference
// (ConorWilliams/openFLY 0.9.0) for: periodic boxes and cell grids, the ghost-atom neighbour list,
// natural cubic splines, the setfl (eam/fs) reader, EAM/Finnis-Sinclair energy / gradient / dense
// Hessian, the L-BFGS two-loop and minimiser, the dimer rotor and translator, and the seeded
// perturbation of Master.  Every function cites the reference file:line it follows.
//
// It is used ONLY by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as
// the checker or as the timed CPU baseline, never as the product path.
//
// Pinning (how each part is checked against something other than itself):
//   Box / grid / images ... brute-force minimum-image neighbour lists (tests/test_oracle_list.py)
//   List .................. brute-force O(N^2) lists, random ortho + triclinic boxes, update() invariance
//   Spline ................ exact cubic reproduction away from the ends, constant tables
//   DataEAM ............... wen.eam.fs known answers F_Fe(0), F_Fe(0.02)
//   EAM ................... wen.eam.fs known answers E0, grad(0), H(0,0); FD of energy; sum of forces
//   Hessian ............... KAT H(0,0) block; FD of the gradient
//   StepLBFGS ............. exact minimiser of quadratics in n+1 steps; first step is identity
//   LBFGS::minimise ....... converges to the analytic minimum of the perfect lattice (|g| < f2norm)
//   Rotor / Dimer ......... curvature = lowest Hessian eigenvalue (from the dense Hessian); converged
//                           saddles have exactly one negative mode and |g| < f2norm.  Step-by-step
//                           trajectories are parity unpinned (the reference has no tests for them).
//   perturb / Gaussian .... own published algorithm (Marsaglia polar on Xoshiro256**): parity unpinned
//                           against libstdc++ (implementation defined), pinned by moments only.
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

  using V3 = std::array<double, 3>;

  struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
  };

  // Spline::OutOfBounds (spline.hpp:46-51)
  struct OutOfBounds : RuntimeError {
    using RuntimeError::RuntimeError;
  };

  inline double dot3(V3 const& a, V3 const& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

  // ------------------------------------------------------------------------------------------------
  // Boxes and grids: box.hpp:63-76, orthorhombic.hpp:37-164, triclinic.hpp:42-212, hypergrid.hpp:35-100
  // ------------------------------------------------------------------------------------------------

  struct Box {
    double B[3][3];      // B[row][col]; column c is basis vector c
    bool per[3];
    bool ortho;
    double ext[3], inv_ext[3];   // orthorhombic
    double Binv[3][3];           // triclinic
    double hyper[3][3];          // triclinic: hyper[row][col], column c = unit normal of plane c

    Box(double const* basis_rowmajor, int const* periodic);

    V3 col(int c) const { return {B[0][c], B[1][c], B[2][c]}; }
    V3 canon_image(V3 x) const;
    double min_width() const;
  };

  struct Grid {
    bool ortho;
    double rc;
    int shape[3], prod[3];
    double cell[3], inv_cell[3];
    double ext[3];        // ortho extents
    double B[3][3];       // triclinic basis
    double hyper[3][3];   // triclinic normals

    Grid(Box const& box, double r_cut);
    int cell_idx(V3 const& x) const;
    bool gen_image(int sign, V3& x, int ax) const;   // sign = +1 / -1; returns true and writes image
    int ncells() const { return shape[0] * shape[1] * shape[2]; }
  };

  // ------------------------------------------------------------------------------------------------
  // Neighbour list: list.hpp:71-210, list.cpp:37-165, adjacent.hpp:35-90
  // ------------------------------------------------------------------------------------------------

  class List {
  public:
    List(Box const& box, double r_cut);

    void rebuild(double const* pos, long n, int threads = 1);
    void update(double const* delta, long n);

    template <typename F>
    void for_neighbours(long i, double r_cut, F&& f) const {
      if (r_cut > m_rc) throw RuntimeError("r_cut is bigger than the internal r_cut");
      V3 const& ri = m_r[i];
      for (int n : m_lists[i]) {
        V3 dr = {m_r[n][0] - ri[0], m_r[n][1] - ri[1], m_r[n][2] - ri[2]};
        double r_sq = (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
        if (r_sq < r_cut * r_cut) f(m_idx[n], std::sqrt(r_sq), dr);
      }
    }

    // The image shift (per axis -1/0/+1) of stored neighbour n of atom i: not part of the reference
    // API; used to compare pair sets as (i, j, shift) multisets.
    template <typename F>
    void for_neighbours_shift(long i, double r_cut, F&& f) const {
      V3 const& ri = m_r[i];
      for (int n : m_lists[i]) {
        V3 dr = {m_r[n][0] - ri[0], m_r[n][1] - ri[1], m_r[n][2] - ri[2]};
        double r_sq = (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
        if (r_sq < r_cut * r_cut) f(m_idx[n], m_shift[n], std::sqrt(r_sq), dr);
      }
    }

    long size() const { return m_n; }
    double r_cut() const { return m_rc; }
    long num_plus_ghosts() const { return m_np; }

  private:
    Box m_box;
    Grid m_grid;
    std::vector<std::vector<int>> m_adj;   // AdjacentCells
    double m_rc;
    long m_n = 0, m_np = 0;
    std::vector<V3> m_r;
    std::vector<long> m_idx;
    std::vector<int> m_shift;   // packed (s0+1) + 3 (s1+1) + 9 (s2+1)
    std::vector<int> m_next, m_head;
    std::vector<std::vector<int>> m_lists;

    void make_ghosts();
    void build_neigh_list(long i);
  };

  // ------------------------------------------------------------------------------------------------
  // Spline: spline.hpp:41-155, spline.cpp:27-103
  // ------------------------------------------------------------------------------------------------

  class Spline {
  public:
    struct Spine {
      double a, b, c, d;
    };
    Spline() = default;
    Spline(std::vector<double> y, double dx);

    double f(double x) const {
      double x0;
      Spine const& s = fetch(x, x0);
      return s.a + x0 * (s.b + x0 * (s.c + x0 * s.d));
    }
    double fp(double x) const {
      double x0;
      Spine const& s = fetch(x, x0);
      return s.b + x0 * (2.0 * s.c + x0 * 3.0 * s.d);
    }
    double fpp(double x) const {
      double x0;
      Spine const& s = fetch(x, x0);
      return 2 * s.c + 6.0 * x0 * s.d;
    }
    std::vector<Spine> const& spines() const { return m_spines; }
    double dx() const { return m_dx; }

  private:
    std::vector<Spine> m_spines;
    double m_dx = 0, m_inv_dx = 0;

    Spine const& fetch(double x, double& x0) const {
      auto i = static_cast<std::size_t>(x * m_inv_dx);
      if (i >= m_spines.size()) throw OutOfBounds("spline argument outside tabulated region");
      x0 = x - static_cast<double>(i) * m_dx;
      return m_spines[i];
    }
  };

  // ------------------------------------------------------------------------------------------------
  // DataEAM (setfl / eam.fs): data.hpp:39-113, data.cpp:48-153
  // ------------------------------------------------------------------------------------------------

  class DataEAM {
  public:
    explicit DataEAM(std::string const& path);
    double r_cut() const { return m_rcut; }
    std::size_t num_types() const { return m_n; }
    std::string const& name(std::size_t i) const { return m_names[i]; }
    double mass(std::size_t i) const { return m_mass[i]; }
    Spline const& f(std::size_t i) const { return m_f[i]; }
    Spline const& phi(std::size_t a, std::size_t b) const { return m_phi[a + m_n * b]; }
    Spline const& v(std::size_t a, std::size_t b) const {
      std::size_t hi = a > b ? a : b, lo = a > b ? b : a;
      return m_v[hi + m_n * lo];
    }

  private:
    std::size_t m_n = 0;
    double m_rcut = 0;
    std::vector<std::string> m_names;
    std::vector<double> m_mass;
    std::vector<Spline> m_f, m_phi, m_v;
  };

  // ------------------------------------------------------------------------------------------------
  // EAM: eam.hpp:42-120, eam.cpp:33-233
  // ------------------------------------------------------------------------------------------------

  class EAM {
  public:
    explicit EAM(DataEAM const* data) : m_data(data) {}
    double r_cut() const { return m_data->r_cut(); }
    double energy(std::uint32_t const* id, std::uint8_t const* fzn, List const& nl, int threads = 1) const;
    void gradient(double* grad, std::uint32_t const* id, std::uint8_t const* fzn, List const& nl, int threads = 1);
    // Dense 3N x 3N; only the lower block triangle is written (hessian.hpp:42-148).
    void hessian(std::vector<double>& H, std::uint32_t const* id, std::uint8_t const* fzn, List const& nl, int threads = 1);

  private:
    DataEAM const* m_data;
    std::vector<double> m_fprime, m_rho;
    std::vector<V3> m_mu;
  };

  // y = H_sym v where H is the lower-block-triangular Hessian written by EAM::hessian.
  void hessian_times(std::vector<double> const& H, long n, double const* v, double* y);

  // ------------------------------------------------------------------------------------------------
  // Minimisers / saddle: lbfgs_core.hpp:37-119, lbfgs_core.cpp:29-87, lbfgs.cpp:29-115,
  //                      rotor.cpp:33-164, dimer.cpp:28-171
  // ------------------------------------------------------------------------------------------------

  class StepLBFGS {
  public:
    explicit StepLBFGS(int n) : m_n(n), m_hist(static_cast<std::size_t>(n)) {}
    void clear() { m_k = 0; }
    std::vector<double>& newton_step(std::vector<double> const& r, std::vector<double> const& g);

  private:
    struct Elem {
      std::vector<double> s, y;
      double rho = 0, alpha = 0;
    };
    int m_n;
    int m_k = 0;
    std::vector<Elem> m_hist;
    std::vector<double> m_prev_x, m_prev_g, m_q, m_r;
  };

  struct System {   // the SoA<Position, TypeID, Frozen> of a caller
    long n;
    std::vector<double> r;
    std::vector<std::uint32_t> id;
    std::vector<std::uint8_t> fzn;
  };

  struct LBFGSOptions {   // lbfgs.hpp:53-88
    int n = 10;
    int iter_max = 2000;
    double f2norm = 1e-6;
    double skin_frac = 1.1;
    double proj_tol = 0;
    double max_trust = 0.5;
    double min_trust = 0.01;
    double grow_trust = 1.5;
    double shrink_trust = 0.5;
  };

  // Returns true on FAILURE (lbfgs.cpp:114).  r (in/out) positions, g (out) final gradient.
  bool lbfgs_minimise(LBFGSOptions const& opt, Box const& box, EAM& pot, std::vector<double>& r, std::vector<double>& g,
                      std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, int* iters);

  struct RotorOptions {   // rotor.hpp:47-60
    int n = 6;
    int iter_max_rot = 20;
    double delta_r = 0.001;
    double theta_tol = 1 * M_PI / 360.;
    bool relax_in_convex = true;
  };

  struct DimerOptions {   // dimer.hpp:56-98
    int n = 10;
    int max_steps = 1000;
    int hist_check_freq = 5;
    double f2norm = 1e-5;
    double skin_frac = 1.1;
    double proj_tol = 0;
    double max_trust = 0.5;
    double min_trust = 0.05;
    double grow_trust = 1.5;
    double shrink_trust = 0.5;
    int convex_max = 3;
    bool use_history = true;
  };

  enum Exit { success = 0, convex = 1, iter_max = 2, collision = 3, uninit = 4 };   // dimer.hpp:123-129

  struct DimerStats {
    int steps = 0;
    long grads = 0;
    int rebuilds = 0;
    double curv = 0;
  };

  Exit dimer_find_sp(DimerOptions const& opt, RotorOptions const& ropt, Box const& box, EAM& pot, std::vector<double>& r,
                     std::vector<double>& ax, std::uint32_t const* id, std::uint8_t const* fzn, long n,
                     double const* in_min, std::vector<std::vector<double>> const& hist_sp, double theta_tol,
                     int count_frozen, int threads, DimerStats* stats);

  // Rotor::eff_gradient (rotor.cpp:49-164) on a caller's system: a list at the rotor's cutoff is built at
  // r; ax is rotated in place; out receives the effective gradient; returns the curvature.
  double rotor_eff_gradient(RotorOptions const& ropt, Box const& box, EAM& pot, double const* r, std::vector<double>& ax,
                            std::uint32_t const* id, std::uint8_t const* fzn, long n, int count_frozen, int threads,
                            std::vector<double>& out, long* grads);

  // The same rotor and dimer on analytic surfaces (test infrastructure for pinning them):
  // E = 1/2 (x - c)^T A (x - c) over n3 coordinates, and one atom on the Mueller-Brown surface.
  void mueller_brown(double const* x, double kz, double* e, double* g);
  double rotor_quadratic(RotorOptions const& ropt, long n3, double const* A, double const* c, double const* x,
                         std::vector<double>& ax, int count_frozen, std::vector<double>& out, long* grads);
  Exit dimer_quadratic(DimerOptions const& opt, RotorOptions const& ropt, long n3, double const* A, double const* c,
                       std::vector<double>& r, std::vector<double>& ax, double const* in_min,
                       std::vector<std::vector<double>> const& hist_sp, double theta_tol, int count_frozen,
                       DimerStats* stats);
  Exit dimer_mueller_brown(DimerOptions const& opt, RotorOptions const& ropt, double kz, std::vector<double>& r,
                           std::vector<double>& ax, double const* in_min, std::vector<std::vector<double>> const& hist_sp,
                           double theta_tol, DimerStats* stats);

  // ------------------------------------------------------------------------------------------------
  // Xoshiro256** (random.hpp:42-162) and the Gaussian sampler used for seeded perturbations
  // ------------------------------------------------------------------------------------------------

  struct Xoshiro {
    std::uint64_t s[4];
    explicit Xoshiro(std::array<std::uint64_t, 4> seed);
    static std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    std::uint64_t operator()();
    void jump_impl(std::array<std::uint64_t, 4> const& J);
    void jump() { jump_impl({0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL, 0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL}); }
    void long_jump() { jump_impl({0x76e15d3efefdcbbfULL, 0xc5004e441c522fb3ULL, 0x77710069854ee241ULL, 0x39109bb02acbe635ULL}); }
    double uniform();   // [0, 1) from the top 53 bits
  };

  struct Normal {   // Marsaglia polar method (Marsaglia & Bray 1964), caching the second variate
    double mean, stddev;
    bool have = false;
    double saved = 0;
    double operator()(Xoshiro& g);
  };

  // Master::perturb (find.cpp:1147-1214) with the Gaussians above; returns the sampled stddev.
  double perturb(Xoshiro& prng, double stddev_mean, double r_pert, List const& nl, long centre, double const* r_in,
                 std::uint8_t const* fzn, long n, double* r_out, double* ax_out);

}  // namespace orc
