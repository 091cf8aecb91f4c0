// FP64 CPU oracle (TEST INFRASTRUCTURE ONLY): restatement of the reference behaviour, see oracle.hpp.
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <fstream>
#include <limits>
#include <sstream>
#include <utility>

namespace orc {

  // ================================================================================================
  // Box: box.hpp:63-76 (choice of crystal system), orthorhombic.hpp:101-139, triclinic.hpp:134-186
  // ================================================================================================

  Box::Box(double const* b, int const* periodic) {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) B[r][c] = b[3 * r + c];
    for (int k = 0; k < 3; ++k) per[k] = periodic[k] != 0;

    // box.hpp:63-76: orthorhombic if the off-diagonal norm is below 1e-4.
    double skew = 0;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        if (r != c) skew += B[r][c] * B[r][c];
    ortho = std::sqrt(skew) < 0.0001;

    if (ortho) {
      for (int k = 0; k < 3; ++k) {
        ext[k] = B[k][k];
        if (!(ext[k] > 0)) throw RuntimeError("Orthorhombic extents must be positive");
        inv_ext[k] = 1.0 / ext[k];
      }
    } else {
      // triclinic.hpp:134-158: upper triangular, non-negative, non-zero diagonal.
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < r; ++c)
          if (std::abs(B[r][c]) >= 1e-5) throw RuntimeError("Basis must be an upper triangular matrix.");
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          if (r > c) B[r][c] = 0;
          if (B[r][c] < 0) throw RuntimeError("Basis elements must be positive");
        }
      for (int k = 0; k < 3; ++k)
        if (!(B[k][k] > 1e-5)) throw RuntimeError("Diagonal elements must be non-zero");
      // Inverse of an upper-triangular 3x3 (back substitution in closed form).
      double a = B[0][0], bb = B[0][1], c = B[0][2], d = B[1][1], e = B[1][2], f = B[2][2];
      for (auto& row : Binv)
        for (double& x : row) x = 0;
      Binv[0][0] = 1 / a;
      Binv[1][1] = 1 / d;
      Binv[2][2] = 1 / f;
      Binv[0][1] = -bb / (a * d);
      Binv[1][2] = -e / (d * f);
      Binv[0][2] = (bb * e - c * d) / (a * d * f);
      // Plane normals: column i is the unit normal of the plane spanned by the other two basis vectors,
      // oriented along basis vector i (triclinic.hpp:94-115; the kernel of [P^T | 1] is their cross product).
      for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        V3 bj = col(j), bk = col(k), bi = col(i);
        V3 nrm = {bj[1] * bk[2] - bj[2] * bk[1], bj[2] * bk[0] - bj[0] * bk[2], bj[0] * bk[1] - bj[1] * bk[0]};
        double len = std::sqrt(dot3(nrm, nrm));
        for (double& x : nrm) x /= len;
        double w = dot3(nrm, bi);
        for (int r = 0; r < 3; ++r) hyper[r][i] = w < 0 ? -nrm[r] : nrm[r];
      }
    }
  }

  V3 Box::canon_image(V3 x) const {
    if (ortho) {   // orthorhombic.hpp:133-139
      for (int k = 0; k < 3; ++k) x[k] = x[k] - ext[k] * std::floor(x[k] * inv_ext[k]);
      return x;
    }
    // triclinic.hpp:176-186: canonicalise in fractional coordinates.
    V3 f;
    for (int r = 0; r < 3; ++r) f[r] = (Binv[r][0] * x[0] + Binv[r][1] * x[1]) + Binv[r][2] * x[2];
    for (int r = 0; r < 3; ++r) f[r] = f[r] - std::floor(f[r]);
    V3 out;
    for (int r = 0; r < 3; ++r) out[r] = (B[r][0] * f[0] + B[r][1] * f[1]) + B[r][2] * f[2];
    return out;
  }

  double Box::min_width() const {
    if (ortho) return std::min(std::min(ext[0], ext[1]), ext[2]);
    double m = std::numeric_limits<double>::max();   // triclinic.hpp:76-84
    for (int i = 0; i < 3; ++i) {
      V3 bi = col(i), h = {hyper[0][i], hyper[1][i], hyper[2][i]};
      m = std::min(m, dot3(bi, h));
    }
    return m;
  }

  // ================================================================================================
  // Grid: hypergrid.hpp:43-99; OrthoGrid::gen_image orthorhombic.hpp:54-74; TriGrid triclinic.hpp:47-71,94
  // ================================================================================================

  Grid::Grid(Box const& box, double r_cut) : ortho(box.ortho), rc(r_cut) {
    double e[3];
    for (int k = 0; k < 3; ++k) {
      // TriGrid feeds the column sums of the basis to HyperGrid (triclinic.hpp:94).
      e[k] = ortho ? box.ext[k] : (box.B[0][k] + box.B[1][k]) + box.B[2][k];
      ext[k] = e[k];
    }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        B[r][c] = box.B[r][c];
        hyper[r][c] = box.hyper[r][c];
      }
    if (!(r_cut > 0)) throw RuntimeError("r_cut is negative");
    for (int k = 0; k < 3; ++k)
      if (!(e[k] > r_cut)) throw RuntimeError("r_cut is too big for the box");
    if (!ortho) {
      for (int i = 0; i < 3; ++i) {
        V3 bi = box.col(i), h = {hyper[0][i], hyper[1][i], hyper[2][i]};
        if (!(std::abs(dot3(h, bi)) > r_cut)) throw RuntimeError("Box axis too small for this r_cut");
      }
    }
    int p = 1;
    for (int k = 0; k < 3; ++k) {
      shape[k] = 2 + static_cast<int>(e[k] / r_cut);
      cell[k] = e[k] / std::floor(e[k] / r_cut);
      inv_cell[k] = 1.0 / cell[k];
      prod[k] = p;
      p *= shape[k];
    }
  }

  int Grid::cell_idx(V3 const& x) const {
    int idx = 0;
    for (int k = 0; k < 3; ++k) {
      int i = static_cast<int>((x[k] + cell[k]) * inv_cell[k]);
      i = std::min(std::max(i, 0), shape[k] - 1);
      idx += i * prod[k];
    }
    return idx;
  }

  bool Grid::gen_image(int sign, V3& x, int ax) const {
    if (ortho) {
      if (sign > 0) {
        if (x[ax] < rc) {
          x[ax] += ext[ax];
          return true;
        }
      } else if (x[ax] > ext[ax] - rc) {
        x[ax] -= ext[ax];
        return true;
      }
      return false;
    }
    V3 h = {hyper[0][ax], hyper[1][ax], hyper[2][ax]};
    V3 b = {B[0][ax], B[1][ax], B[2][ax]};
    if (sign > 0) {
      if (dot3(x, h) < rc) {
        for (int r = 0; r < 3; ++r) x[r] = x[r] + b[r];
        return true;
      }
    } else {
      V3 d = {b[0] - x[0], b[1] - x[1], b[2] - x[2]};
      if (dot3(d, h) < rc) {
        for (int r = 0; r < 3; ++r) x[r] = x[r] - b[r];
        return true;
      }
    }
    return false;
  }

  // ================================================================================================
  // List: list.cpp:37-165, adjacent.hpp:40-78 (z-major order of template_for, core.hpp:450-506)
  // ================================================================================================

  List::List(Box const& box, double r_cut) : m_box(box), m_grid(box, r_cut), m_rc(r_cut) {
    int const* s = m_grid.shape;
    m_adj.resize(static_cast<std::size_t>(m_grid.ncells()));
    for (int cz = 0; cz < s[2]; ++cz)
      for (int cy = 0; cy < s[1]; ++cy)
        for (int cx = 0; cx < s[0]; ++cx) {
          int n = cx + s[0] * (cy + s[1] * cz);
          for (int z = std::max(cz - 1, 0); z < std::min(cz + 2, s[2]); ++z)
            for (int y = std::max(cy - 1, 0); y < std::min(cy + 2, s[1]); ++y)
              for (int x = std::max(cx - 1, 0); x < std::min(cx + 2, s[0]); ++x) {
                int m = x + s[0] * (y + s[1] * z);
                if (m != n) m_adj[static_cast<std::size_t>(n)].push_back(m);
              }
        }
  }

  void List::rebuild(double const* pos, long n, int threads) {
    if (threads <= 0) throw RuntimeError("not a valid number of threads for rebuild()");
    if (n != m_n) {
      m_n = n;
      std::size_t cap = static_cast<std::size_t>(n) * 27;
      m_r.assign(cap, V3{0, 0, 0});
      m_idx.assign(cap, 0);
      m_shift.assign(cap, 13);
      m_next.assign(cap, 0);
      m_lists.assign(static_cast<std::size_t>(n), {});
    }
    for (long i = 0; i < n; ++i) {   // list.cpp:68-72
      m_r[i] = m_box.canon_image({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
      m_idx[i] = i;
      m_shift[i] = 13;
    }
    make_ghosts();
    // list.cpp:76-89: linked cell list, each bucket in descending index order.
    m_head.assign(static_cast<std::size_t>(m_grid.ncells()), -1);
    for (long i = 0; i < m_np; ++i) {
      int c = m_grid.cell_idx(m_r[i]);
      m_next[i] = m_head[c];
      m_head[c] = static_cast<int>(i);
    }
#pragma omp parallel for num_threads(threads) schedule(static)
    for (long i = 0; i < n; ++i) build_neigh_list(i);
  }

  void List::make_ghosts() {   // list.cpp:100-132
    long next = m_n;
    static int const pow3[3] = {1, 3, 9};
    for (int ax = 0; ax < 3; ++ax) {
      if (!m_box.per[ax]) continue;
      long const end = next;
      for (int sign : {+1, -1}) {
        for (long j = 0; j < end; ++j) {
          V3 im = m_r[j];
          if (m_grid.gen_image(sign, im, ax)) {
            if (next >= static_cast<long>(m_r.size())) throw RuntimeError("Not enough space for ghosts");
            m_r[next] = im;
            m_idx[next] = m_idx[j];
            m_shift[next] = m_shift[j] + sign * pow3[ax];
            ++next;
          }
        }
      }
    }
    m_np = next;
  }

  void List::build_neigh_list(long i) {   // list.cpp:134-165
    auto& out = m_lists[static_cast<std::size_t>(i)];
    out.clear();
    V3 const& ri = m_r[i];
    double rc2 = m_rc * m_rc;
    auto d2 = [&](int n) {
      double dx = ri[0] - m_r[n][0], dy = ri[1] - m_r[n][1], dz = ri[2] - m_r[n][2];
      return (dx * dx + dy * dy) + dz * dz;
    };
    int c = m_grid.cell_idx(ri);
    for (int n = m_head[c]; n != -1; n = m_next[n])
      if (n != i && d2(n) < rc2) out.push_back(n);
    for (int cn : m_adj[static_cast<std::size_t>(c)])
      for (int n = m_head[cn]; n != -1; n = m_next[n])
        if (d2(n) < rc2) out.push_back(n);
  }

  void List::update(double const* x, long n) {   // list.cpp:37-51
    if (n != m_n) throw RuntimeError("Input to update has wrong number of atoms");
    for (long i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) m_r[i][k] -= x[3 * i + k];
    for (long g = n; g < m_np; ++g)
      for (int k = 0; k < 3; ++k) m_r[g][k] -= x[3 * m_idx[g] + k];
  }

  // ================================================================================================
  // Spline: spline.cpp:27-103 -- natural cubic spline (Burden & Faires, Numerical Analysis, Alg. 3.4)
  // on a uniform grid, padded with five copies of the last value.
  // ================================================================================================

  Spline::Spline(std::vector<double> y, double dx) : m_dx(dx), m_inv_dx(1 / dx) {
    double last = y.back();
    for (int i = 0; i < 5; ++i) y.push_back(last);
    std::size_t n = y.size() - 1;
    std::vector<double> a(y), b(n), d(n), alpha(n), c(n + 1), l(n + 1), mu(n + 1), z(n + 1);
    for (std::size_t i = 1; i <= n - 1; ++i) alpha[i] = 3 * (a[i + 1] - 2 * a[i] + a[i - 1]) / dx;
    l[0] = 1;
    mu[0] = 0;
    z[0] = 0;
    for (std::size_t i = 1; i <= n - 1; ++i) {
      l[i] = dx * (4 - mu[i - 1]);
      mu[i] = dx / l[i];
      z[i] = (alpha[i] - dx * z[i - 1]) / l[i];
    }
    l[n] = 1;
    z[n] = 0;
    c[n] = 0;
    for (std::size_t i = 1; i <= n; ++i) {
      std::size_t j = n - i;
      c[j] = z[j] - mu[j] * c[j + 1];
      b[j] = (a[j + 1] - a[j]) / dx - dx * (c[j + 1] + 2 * c[j]) / 3;
      d[j] = (c[j + 1] - c[j]) / (3 * dx);
    }
    m_spines.reserve(n);
    for (std::size_t i = 0; i <= n - 1; ++i) m_spines.push_back({a[i], b[i], c[i], d[i]});
  }

  // ================================================================================================
  // DataEAM: data.cpp:37-153 (setfl: 3 comment lines, species line, tabulation line, per-species
  // Z/mass + F + rho tables, then r*V pair tables for j <= i; five values per line)
  // ================================================================================================

  namespace {
    std::istringstream next_line(std::ifstream& in) {
      std::string line;
      if (!std::getline(in, line)) throw RuntimeError("File terminated too soon");
      return std::istringstream{line};
    }
    std::vector<double> read_chunked(std::ifstream& in, std::size_t n, std::size_t k) {
      std::vector<double> raw;
      raw.reserve(n);
      std::istringstream line;
      for (std::size_t i = 0; i < n; ++i) {
        if (i % k == 0) line = next_line(in);
        double t;
        line >> t;
        raw.push_back(t);
      }
      return raw;
    }
  }  // namespace

  DataEAM::DataEAM(std::string const& path) {
    std::ifstream in(path);
    if (!in.good()) throw RuntimeError("Could not open eam file " + path);
    for (int i = 0; i < 3; ++i) next_line(in);
    {
      auto line = next_line(in);
      int n;
      line >> n;
      if (n <= 0) throw RuntimeError("bad species count");
      m_n = static_cast<std::size_t>(n);
      for (std::size_t i = 0; i < m_n; ++i) {
        std::string name;
        line >> name;
        m_names.push_back(name);
      }
    }
    m_f.resize(m_n);
    m_phi.resize(m_n * m_n);
    m_v.resize(m_n * m_n);
    m_mass.resize(m_n);
    std::size_t numP, numR;
    double delP, delR;
    next_line(in) >> numP >> delP >> numR >> delR >> m_rcut;
    for (std::size_t i = 0; i < m_n; ++i) {
      std::size_t atomic;
      next_line(in) >> atomic >> m_mass[i];
      m_f[i] = Spline{read_chunked(in, numP, 5), delP};
      for (std::size_t j = 0; j < m_n; ++j) m_phi[i + m_n * j] = Spline{read_chunked(in, numR, 5), delR};
    }
    for (std::size_t i = 0; i < m_n; ++i)
      for (std::size_t j = 0; j <= i; ++j) {
        std::vector<double> raw = read_chunked(in, numR, 5);
        for (std::size_t k = 0; k < raw.size(); ++k) raw[k] /= delR * static_cast<double>(k);
        if (raw.size() < 2) throw RuntimeError("no elements!");
        raw[0] = raw[1];
        m_v[i + m_n * j] = Spline{std::move(raw), delR};
      }
  }

  // ================================================================================================
  // EAM: eam.cpp:33-233
  // ================================================================================================

  namespace {
    // Run body(i) for i in [0,n) with OpenMP; exceptions are captured and rethrown on the caller.
    template <typename F>
    void parallel_for(long n, int threads, F&& body) {
      std::atomic<int> failed{0};
      std::string msg;
      bool oob = false;
#pragma omp parallel for num_threads(threads) schedule(static)
      for (long i = 0; i < n; ++i) {
        if (failed.load(std::memory_order_relaxed)) continue;
        try {
          body(i);
        } catch (OutOfBounds const& e) {
          if (!failed.exchange(1)) {
            msg = e.what();
            oob = true;
          }
        } catch (std::exception const& e) {
          if (!failed.exchange(1)) msg = e.what();
        }
      }
      if (failed) {
        if (oob) throw OutOfBounds(msg);
        throw RuntimeError(msg);
      }
    }
  }  // namespace

  double EAM::energy(std::uint32_t const* id, std::uint8_t const*, List const& nl, int threads) const {
    long n = nl.size();
    std::vector<double> v_part(static_cast<std::size_t>(n)), f_part(static_cast<std::size_t>(n));
    double rc = r_cut();
    parallel_for(n, threads, [&](long a) {
      double v_sum = 0, rho = 0;
      nl.for_neighbours(a, rc, [&](long b, double r, V3 const&) {
        v_sum += m_data->v(id[a], id[b]).f(r);
        rho += m_data->phi(id[b], id[a]).f(r);
      });
      v_part[a] = v_sum;
      f_part[a] = m_data->f(id[a]).f(rho);
    });
    double v_sum = 0, f_sum = 0;
    for (long a = 0; a < n; ++a) {
      v_sum += v_part[a];
      f_sum += f_part[a];
    }
    return (0.5 * v_sum) + f_sum;
  }

  void EAM::gradient(double* out, std::uint32_t const* id, std::uint8_t const* fzn, List const& nl, int threads) {
    long n = nl.size();
    m_fprime.resize(static_cast<std::size_t>(n));
    double rc = r_cut();
    parallel_for(n, threads, [&](long b) {   // eam.cpp:62-86
      double rho = 0;
      nl.for_neighbours(b, rc, [&](long a, double r, V3 const&) { rho += m_data->phi(id[a], id[b]).f(r); });
      m_fprime[b] = m_data->f(id[b]).fp(rho);
    });
    parallel_for(n, threads, [&](long g) {   // eam.cpp:90-120
      V3 grad = {0, 0, 0};
      if (!fzn[g]) {
        nl.for_neighbours(g, rc, [&](long a, double r, V3 const& dr) {
          double mag = m_data->v(id[a], id[g]).fp(r) + m_fprime[g] * m_data->phi(id[a], id[g]).fp(r)
                       + m_fprime[a] * m_data->phi(id[g], id[a]).fp(r);
          for (int k = 0; k < 3; ++k) grad[k] -= (mag / r) * dr[k];
        });
      }
      for (int k = 0; k < 3; ++k) out[3 * g + k] = grad[k];
    });
  }

  void EAM::hessian(std::vector<double>& H, std::uint32_t const* id, std::uint8_t const* fzn, List const& nl, int threads) {
    long n = nl.size();
    long dim = 3 * n;
    H.assign(static_cast<std::size_t>(dim * dim), 0.0);
    m_rho.resize(static_cast<std::size_t>(n));
    m_mu.resize(static_cast<std::size_t>(n));
    double rc = r_cut();
    auto blk = [&](long i, long j, int r, int c) -> double& { return H[static_cast<std::size_t>((3 * i + r) * dim + 3 * j + c)]; };

    parallel_for(n, threads, [&](long b) {   // eam.cpp:129-142, A.13
      double rho = 0;
      V3 mu = {0, 0, 0};
      nl.for_neighbours(b, rc, [&](long a, double r, V3 const& dr) {
        rho += m_data->phi(id[a], id[b]).f(r);
        double p = m_data->phi(id[a], id[b]).fp(r);
        for (int k = 0; k < 3; ++k) mu[k] -= p * dr[k] / r;
      });
      m_rho[b] = rho;
      m_mu[b] = mu;
    });

    // eam.cpp:165-231: column block z is written only by iteration z (serial here: overlap terms
    // write to block (g, z) which belongs to column z, so a parallel loop over z is race free).
    parallel_for(n, threads, [&](long z) {
      if (fzn[z]) return;
      auto const& D = *m_data;
      V3 v = m_mu[z];
      double Fpp_z = D.f(id[z]).fpp(m_rho[z]);
      double Fp_z = D.f(id[z]).fp(m_rho[z]);
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) blk(z, z, r, c) = Fpp_z * v[r] * v[c];

      nl.for_neighbours(z, rc, [&](long a, double r, V3 const& dr) {
        double Fp_a = D.f(id[a]).fp(m_rho[a]);
        double Fpp_a = D.f(id[a]).fpp(m_rho[a]);
        double A = (D.v(id[a], id[z]).fp(r) + Fp_z * D.phi(id[a], id[z]).fp(r) + Fp_a * D.phi(id[z], id[a]).fp(r)) / r;
        double Bv = D.v(id[a], id[z]).fpp(r) + Fp_z * D.phi(id[a], id[z]).fpp(r) + Fp_a * D.phi(id[z], id[a]).fpp(r);
        double pza = D.phi(id[z], id[a]).fp(r);
        double ABFpp = (A - Bv - Fpp_a * (pza * pza)) / (r * r);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) blk(z, z, i, j) += (i == j ? A : 0.0) - ABFpp * dr[i] * dr[j];

        if (a == z) throw RuntimeError("Atom is interacting with itself");

        if (a > z && !fzn[a]) {
          double BArr = (Bv - A) / (r * r);
          double ur = Fpp_z * D.phi(id[a], id[z]).fp(r) / r;
          double ru = Fpp_a * pza / r;
          V3 const& mz = m_mu[z];
          V3 const& ma = m_mu[a];
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
              blk(a, z, i, j) += -BArr * dr[i] * dr[j] - (i == j ? A : 0.0) + ur * dr[i] * mz[j] - ru * ma[i] * dr[j];
        }

        double ddFg = Fpp_a * pza;
        nl.for_neighbours(a, rc, [&](long g, double r_ag, V3 const& dr_ag) {
          V3 s = {dr[0] + dr_ag[0], dr[1] + dr_ag[1], dr[2] + dr_ag[2]};
          if ((g > z && !fzn[g]) || (g == z && std::sqrt(dot3(s, s)) > 1e-6)) {
            double mag = ddFg / (r * r_ag) * D.phi(id[g], id[a]).fp(r_ag);
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) blk(g, z, i, j) -= mag * dr_ag[i] * dr[j];
          }
        });
      });
    });
  }

  void hessian_times(std::vector<double> const& H, long n, double const* v, double* y) {
    long dim = 3 * n;
    for (long i = 0; i < dim; ++i) y[i] = 0;
    for (long bi = 0; bi < n; ++bi)
      for (long bj = 0; bj <= bi; ++bj)
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double h = H[static_cast<std::size_t>((3 * bi + r) * dim + 3 * bj + c)];
            if (h == 0.0) continue;
            y[3 * bi + r] += h * v[3 * bj + c];
            if (bi != bj) y[3 * bj + c] += h * v[3 * bi + r];
          }
  }

  // ================================================================================================
  // StepLBFGS::newton_step_impl: lbfgs_core.cpp:29-87 (Nocedal & Wright Alg. 7.4, ring buffer)
  // ================================================================================================

  namespace {
    double vdot(std::vector<double> const& a, std::vector<double> const& b) {
      double s = 0;
      for (std::size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
      return s;
    }
    double vnorm(std::vector<double> const& a) { return std::sqrt(vdot(a, a)); }
  }  // namespace

  std::vector<double>& StepLBFGS::newton_step(std::vector<double> const& r, std::vector<double> const& g) {
    if (r.size() != g.size()) throw RuntimeError("LBFGS core inputs size mismatch");
    std::size_t N = r.size();
    int prev = m_k > 0 ? (m_k - 1) % m_n : 0;
    m_r.resize(N);
    if (m_k > 0) {
      if (N != m_prev_x.size()) throw RuntimeError("Call to newton_step() with a different number of atoms");
      Elem& e = m_hist[static_cast<std::size_t>(prev)];
      e.s.resize(N);
      e.y.resize(N);
      for (std::size_t i = 0; i < N; ++i) {
        e.s[i] = r[i] - m_prev_x[i];
        e.y[i] = g[i] - m_prev_g[i];
      }
      e.rho = 1.0 / std::abs(vdot(e.s, e.y));
    }
    m_prev_x = r;
    m_prev_g = g;
    m_q = g;
    int incur = m_k <= m_n ? 0 : m_k - m_n;
    int bound = m_k <= m_n ? m_k : m_n;
    for (int i = bound - 1; i >= 0; --i) {
      Elem& e = m_hist[static_cast<std::size_t>((i + incur) % m_n)];
      e.alpha = e.rho * vdot(e.s, m_q);
      for (std::size_t k = 0; k < N; ++k) m_q[k] -= e.alpha * e.y[k];
    }
    if (m_k > 0) {
      Elem& e = m_hist[static_cast<std::size_t>(prev)];
      double scale = 1.0 / e.rho / vdot(e.y, e.y);
      for (std::size_t k = 0; k < N; ++k) m_r[k] = m_q[k] * scale;
    } else {
      m_r = m_q;
    }
    for (int i = 0; i <= bound - 1; ++i) {
      Elem& e = m_hist[static_cast<std::size_t>((i + incur) % m_n)];
      double beta = e.rho * vdot(e.y, m_r);
      for (std::size_t k = 0; k < N; ++k) m_r[k] += (e.alpha - beta) * e.s[k];
    }
    ++m_k;
    return m_r;
  }

  // ================================================================================================
  // LBFGS::minimise: lbfgs.cpp:29-115
  // ================================================================================================

  bool lbfgs_minimise(LBFGSOptions const& opt, Box const& box, EAM& pot, std::vector<double>& r, std::vector<double>& g,
                      std::uint32_t const* id, std::uint8_t const* fzn, long n, int threads, int* iters) {
    StepLBFGS core(opt.n);
    double skin = std::max(std::pow(opt.skin_frac, 1. / 3.) - 1, 0.0) * pot.r_cut();
    List nl(box, pot.r_cut() + skin);
    g.assign(static_cast<std::size_t>(3 * n), 0.0);
    nl.rebuild(r.data(), n, threads);
    pot.gradient(g.data(), id, fzn, nl, threads);
    double trust = opt.min_trust, acc = 0;
    for (int i = 0; i < opt.iter_max; ++i) {
      if (iters) *iters = i;
      double mag_g = vdot(g, g);
      if (mag_g < opt.f2norm * opt.f2norm) return false;
      std::vector<double>& Hg = core.newton_step(r, g);
      double scale = std::min(1.0, trust / vnorm(Hg));
      for (double& x : Hg) x *= scale;
      // (Hg*Hg).colwise().sum().maxCoeff() on a flat 3N x 1 array is the squared L2 norm of the step.
      acc += std::sqrt(vdot(Hg, Hg));
      for (std::size_t k = 0; k < r.size(); ++k) r[k] -= Hg[k];
      if (acc > 0.5 * skin) {
        nl.rebuild(r.data(), n, threads);
        acc = 0;
      } else {
        nl.update(Hg.data(), n);
      }
      pot.gradient(g.data(), id, fzn, nl, threads);
      double proj = vdot(g, Hg);
      if (proj < -opt.proj_tol) {
        trust = std::max(opt.min_trust, opt.shrink_trust * trust);
      } else if (proj > opt.proj_tol) {
        trust = std::min(opt.max_trust, opt.grow_trust * trust);
      }
    }
    if (iters) *iters = opt.iter_max;
    return true;
  }

  // ================================================================================================
  // Rotor::eff_gradient: rotor.cpp:33-164
  // ================================================================================================

  namespace {
    void trans_proj(std::vector<double>& x) {   // rotor.cpp:33-47
      long n = static_cast<long>(x.size() / 3);
      V3 p = {0, 0, 0};
      for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) p[k] += x[3 * i + k];
      for (int k = 0; k < 3; ++k) p[k] *= 1.0 / static_cast<double>(n);
      for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) x[3 * i + k] -= p[k];
    }

    struct Rotor {
      RotorOptions opt;
      StepLBFGS core;
      std::vector<double> delta, delta_prev, axisp, g0, g1, g1p, delta_g, dg;
      long grads = 0;

      explicit Rotor(RotorOptions const& o) : opt(o), core(o.n) {}

      // M: gradient(g) at the model's current positions; update(d): positions -= d (List::update).
      template <typename M>
      double eff_gradient(std::vector<double>& out, std::vector<double>& ax, long n, M& model, int count_frozen) {
        std::size_t N = static_cast<std::size_t>(3 * n);
        delta.resize(N);
        delta_prev.resize(N);
        axisp.resize(N);
        g0.resize(N);
        g1.resize(N);
        g1p.resize(N);
        delta_g.resize(N);
        dg.resize(N);
        core.clear();
        double const dr = opt.delta_r;

        model.gradient(g0.data());
        ++grads;
        for (std::size_t k = 0; k < N; ++k) delta[k] = -dr * ax[k];
        model.update(delta.data());
        std::swap(delta, delta_prev);
        model.gradient(g1.data());
        ++grads;

        double curv = [&] {
          for (int i = 0;; ++i) {
            for (std::size_t k = 0; k < N; ++k) delta_g[k] = g1[k] - g0[k];
            double t = vdot(delta_g, ax);
            for (std::size_t k = 0; k < N; ++k) delta_g[k] -= t * ax[k];
            if (count_frozen == 0) trans_proj(delta_g);

            std::vector<double>& theta = core.newton_step(ax, delta_g);
            double o = vdot(theta, ax);
            for (std::size_t k = 0; k < N; ++k) theta[k] -= o * ax[k];
            double inv = 1 / vnorm(theta);
            for (double& x : theta) x *= inv;

            for (std::size_t k = 0; k < N; ++k) dg[k] = g1[k] - g0[k];
            double b_1 = vdot(dg, theta) / dr;
            double c_x0 = vdot(dg, ax) / dr;
            double theta_1 = -0.5 * std::atan(b_1 / std::abs(c_x0));

            if (std::abs(theta_1) < opt.theta_tol || (i >= opt.iter_max_rot && c_x0 < 0) || i > 10 * opt.iter_max_rot) {
              return c_x0;
            }
            double ct1 = std::cos(theta_1), st1 = std::sin(theta_1);
            for (std::size_t k = 0; k < N; ++k) axisp[k] = ax[k] * ct1 + theta[k] * st1;
            for (std::size_t k = 0; k < N; ++k) delta[k] = -dr * axisp[k];
            std::swap(delta, delta_prev);
            for (std::size_t k = 0; k < N; ++k) delta[k] = delta_prev[k] - delta[k];
            model.update(delta.data());

            model.gradient(g1p.data());
            ++grads;

            for (std::size_t k = 0; k < N; ++k) dg[k] = g1p[k] - g0[k];
            double c_x1 = vdot(dg, axisp) / dr;
            double a_1 = (c_x0 - c_x1 + b_1 * std::sin(2 * theta_1)) / (1 - std::cos(2 * theta_1));
            double theta_min = 0.5 * std::atan(b_1 / a_1);
            if (a_1 * std::cos(2 * theta_min) - a_1 + b_1 * std::sin(2 * theta_min) > 0) theta_min += M_PI / 2;

            double cm = std::cos(theta_min), sm = std::sin(theta_min);
            for (std::size_t k = 0; k < N; ++k) ax[k] = ax[k] * cm + theta[k] * sm;

            double w1 = std::sin(theta_1 - theta_min) / std::sin(theta_1);
            double w2 = std::sin(theta_min) / std::sin(theta_1);
            double w3 = 1 - std::cos(theta_min) - std::sin(theta_min) * std::tan(0.5 * theta_1);
            for (std::size_t k = 0; k < N; ++k) g1[k] = w1 * g1[k] + w2 * g1p[k] + w3 * g0[k];

            if (std::abs(theta_min) < opt.theta_tol) {
              for (std::size_t k = 0; k < N; ++k) dg[k] = g1[k] - g0[k];
              return vdot(dg, ax) / dr;
            }
          }
        }();

        for (std::size_t k = 0; k < N; ++k) delta[k] = -delta_prev[k];
        model.update(delta.data());

        double p = vdot(g0, ax);
        out.resize(N);
        if (!opt.relax_in_convex && curv > 0) {
          for (std::size_t k = 0; k < N; ++k) out[k] = -p * ax[k];
        } else {
          for (std::size_t k = 0; k < N; ++k) out[k] = g0[k] - 2 * p * ax[k];
        }
        if (count_frozen == 0) trans_proj(out);
        return curv;
      }
    };

    // EAM model of a caller's system over a neighbour list (the reference's pot + nl pair).
    struct EAMModel {
      EAM& pot;
      List& nl;
      std::uint32_t const* id;
      std::uint8_t const* fzn;
      long n;
      int threads;
      void gradient(double* g) { pot.gradient(g, id, fzn, nl, threads); }
      void update(double const* d) { nl.update(d, n); }
      void rebuild(double const* r) { nl.rebuild(r, n, threads); }
    };
  }  // namespace

  // ================================================================================================
  // Dimer::find_sp: dimer.cpp:28-171
  // ================================================================================================

  namespace {
    // Dimer::find_sp's loop over a model (M as for Rotor::eff_gradient, plus rebuild(r): new list at r).
    template <typename M>
    Exit dimer_core(DimerOptions const& opt, RotorOptions const& ropt, M& model, double skin, std::vector<double>& r,
                    std::vector<double>& ax, long n, double const* in_min,
                    std::vector<std::vector<double>> const& hist_sp, double theta_tol, int count_frozen,
                    DimerStats* stats) {
    Rotor rotor(ropt);
    StepLBFGS core(opt.n);
    std::size_t N = static_cast<std::size_t>(3 * n);
    std::vector<double> eff;
    DimerStats st;

    model.rebuild(r.data());
    double curv = rotor.eff_gradient(eff, ax, n, model, count_frozen);
    double trust = opt.min_trust, acc = 0, convex_count = 0;
    Exit ret = iter_max;

    std::vector<double> tmp1(N), tmp2(N);
    for (int i = 0; i < opt.max_steps; ++i) {
      st.steps = i;
      double mag_g = vdot(eff, eff);
      if (mag_g < opt.f2norm * opt.f2norm) {
        ret = success;
        break;
      } else if (convex_count >= opt.convex_max) {
        ret = convex;
        break;
      }
      if (opt.use_history && i % opt.hist_check_freq == 0) {
        double ctheta = -1;
        bool hit = false;
        for (auto const& sp : hist_sp) {
          for (std::size_t k = 0; k < N; ++k) {
            tmp1[k] = r[k] - in_min[k];
            tmp2[k] = sp[k] - in_min[k];
          }
          double dt = vdot(tmp1, tmp2);
          double sn = vnorm(tmp2), on = vnorm(tmp1);
          ctheta = std::max(ctheta, dt / (sn * on));
          if (ctheta > std::cos(theta_tol)) {
            hit = true;
            break;
          }
        }
        if (hit) {
          ret = collision;
          break;
        }
      }
      std::vector<double>& Hg = core.newton_step(r, eff);
      double scale = std::min(1.0, trust / vnorm(Hg));
      for (double& x : Hg) x *= scale;
      acc += std::sqrt(vdot(Hg, Hg));   // flat-array colwise().sum(): the L2 norm of the step
      for (std::size_t k = 0; k < N; ++k) r[k] -= Hg[k];
      if (acc > 0.5 * skin) {
        model.rebuild(r.data());
        acc = 0;
        ++st.rebuilds;
      } else {
        model.update(Hg.data());
      }
      curv = rotor.eff_gradient(eff, ax, n, model, count_frozen);
      convex_count = curv > 0 ? convex_count + 1 : 0;
      double proj = vdot(eff, Hg);
      if (proj < -opt.proj_tol) {
        trust = std::max(opt.min_trust, opt.shrink_trust * trust);
      } else if (proj > opt.proj_tol) {
        trust = std::min(opt.max_trust, opt.grow_trust * trust);
      }
      st.steps = i + 1;
    }
    st.grads = rotor.grads;
    st.curv = curv;
    if (stats) *stats = st;
    return ret;
    }
  }  // namespace

  Exit dimer_find_sp(DimerOptions const& opt, RotorOptions const& ropt, Box const& box, EAM& pot, std::vector<double>& r,
                     std::vector<double>& ax, std::uint32_t const* id, std::uint8_t const* fzn, long n,
                     double const* in_min, std::vector<std::vector<double>> const& hist_sp, double theta_tol,
                     int count_frozen, int threads, DimerStats* stats) {
    double rot_rcut = pot.r_cut() + ropt.delta_r * 2;   // rotor.hpp:123
    double skin = std::max(std::pow(opt.skin_frac, 1. / 3.) - 1, 0.0) * rot_rcut;
    List nl(box, rot_rcut + skin);
    EAMModel model{pot, nl, id, fzn, n, threads};
    return dimer_core(opt, ropt, model, skin, r, ax, n, in_min, hist_sp, theta_tol, count_frozen, stats);
  }

  double rotor_eff_gradient(RotorOptions const& ropt, Box const& box, EAM& pot, double const* r, std::vector<double>& ax,
                            std::uint32_t const* id, std::uint8_t const* fzn, long n, int count_frozen, int threads,
                            std::vector<double>& out, long* grads) {
    List nl(box, pot.r_cut() + ropt.delta_r * 2);   // the rotor's list cutoff, rotor.hpp:123
    nl.rebuild(r, n, threads);
    EAMModel model{pot, nl, id, fzn, n, threads};
    Rotor rotor(ropt);
    double c = rotor.eff_gradient(out, ax, n, model, count_frozen);
    if (grads) *grads = rotor.grads;
    return c;
  }

  // ------------------------------------------------------------------------------------------------
  // Analytic surfaces for pinning the rotor and dimer (no reference counterpart).
  // ------------------------------------------------------------------------------------------------

  namespace {
    // E = 1/2 (x - c)^T A (x - c), g = A (x - c); A dense 3n x 3n row-major.
    struct QuadModel {
      long n3;
      double const* A;
      double const* c;
      std::vector<double> x;
      void gradient(double* g) {
        for (long i = 0; i < n3; ++i) {
          double v = 0;
          for (long j = 0; j < n3; ++j) v += A[i * n3 + j] * (x[j] - c[j]);
          g[i] = v;
        }
      }
      void update(double const* d) {
        for (long i = 0; i < n3; ++i) x[i] -= d[i];
      }
      void rebuild(double const* r) { x.assign(r, r + n3); }
    };

    // One atom on the Mueller-Brown surface in (x, y) plus kz z^2 / 2 (K. Mueller and L. D. Brown,
    // Theor. Chim. Acta 53 (1979) 75).
    struct MBModel {
      double kz;
      std::vector<double> x = std::vector<double>(3);
      void gradient(double* g) { mueller_brown(x.data(), kz, nullptr, g); }
      void update(double const* d) {
        for (int i = 0; i < 3; ++i) x[static_cast<std::size_t>(i)] -= d[i];
      }
      void rebuild(double const* r) { x.assign(r, r + 3); }
    };
  }  // namespace

  void mueller_brown(double const* x, double kz, double* e, double* g) {
    static double const A[4] = {-200, -100, -170, 15}, a[4] = {-1, -1, -6.5, 0.7}, b[4] = {0, 0, 11, 0.6},
                        c[4] = {-10, -10, -6.5, 0.7}, x0[4] = {1, 0, -0.5, -1}, y0[4] = {0, 0.5, 1.5, 1};
    double E = 0, gx = 0, gy = 0;
    for (int k = 0; k < 4; ++k) {
      double dx = x[0] - x0[k], dy = x[1] - y0[k];
      double t = A[k] * std::exp(a[k] * dx * dx + b[k] * dx * dy + c[k] * dy * dy);
      E += t;
      gx += t * (2 * a[k] * dx + b[k] * dy);
      gy += t * (b[k] * dx + 2 * c[k] * dy);
    }
    if (e) *e = E + 0.5 * kz * x[2] * x[2];
    if (g) {
      g[0] = gx;
      g[1] = gy;
      g[2] = kz * x[2];
    }
  }

  double rotor_quadratic(RotorOptions const& ropt, long n3, double const* A, double const* c, double const* x,
                         std::vector<double>& ax, int count_frozen, std::vector<double>& out, long* grads) {
    QuadModel model{n3, A, c, std::vector<double>(x, x + n3)};
    Rotor rotor(ropt);
    double cv = rotor.eff_gradient(out, ax, n3 / 3, model, count_frozen);
    if (grads) *grads = rotor.grads;
    return cv;
  }

  Exit dimer_quadratic(DimerOptions const& opt, RotorOptions const& ropt, long n3, double const* A, double const* c,
                       std::vector<double>& r, std::vector<double>& ax, double const* in_min,
                       std::vector<std::vector<double>> const& hist_sp, double theta_tol, int count_frozen,
                       DimerStats* stats) {
    QuadModel model{n3, A, c, {}};
    return dimer_core(opt, ropt, model, 1e300, r, ax, n3 / 3, in_min, hist_sp, theta_tol, count_frozen, stats);
  }

  Exit dimer_mueller_brown(DimerOptions const& opt, RotorOptions const& ropt, double kz, std::vector<double>& r,
                           std::vector<double>& ax, double const* in_min, std::vector<std::vector<double>> const& hist_sp,
                           double theta_tol, DimerStats* stats) {
    MBModel model{kz};
    // One atom, counted as "frozen present" so the rotor does not project out its (only) translation.
    return dimer_core(opt, ropt, model, 1e300, r, ax, 1, in_min, hist_sp, theta_tol, 1, stats);
  }

  // ================================================================================================
  // Xoshiro256**: random.hpp:96-161 (Blackman & Vigna); Gaussians by the Marsaglia polar method.
  // ================================================================================================

  Xoshiro::Xoshiro(std::array<std::uint64_t, 4> seed) {
    if (seed[0] == 0 && seed[1] == 0 && seed[2] == 0 && seed[3] == 0) throw RuntimeError("bad seed");
    for (int i = 0; i < 4; ++i) s[i] = seed[static_cast<std::size_t>(i)];
  }

  std::uint64_t Xoshiro::operator()() {
    std::uint64_t const result = rotl(s[1] * 5, 7) * 9;
    std::uint64_t const t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }

  void Xoshiro::jump_impl(std::array<std::uint64_t, 4> const& J) {
    std::uint64_t t[4] = {0, 0, 0, 0};
    for (std::uint64_t jump : J)
      for (int b = 0; b < 64; ++b) {
        if (jump & (std::uint64_t{1} << b))
          for (int k = 0; k < 4; ++k) t[k] ^= s[k];
        (*this)();
      }
    for (int k = 0; k < 4; ++k) s[k] = t[k];
  }

  double Xoshiro::uniform() { return static_cast<double>((*this)() >> 11) * 0x1.0p-53; }

  double Normal::operator()(Xoshiro& g) {
    if (have) {
      have = false;
      return mean + stddev * saved;
    }
    double u, v, s;
    do {
      u = 2 * g.uniform() - 1;
      v = 2 * g.uniform() - 1;
      s = u * u + v * v;
    } while (s >= 1 || s == 0);
    double m = std::sqrt(-2 * std::log(s) / s);
    saved = v * m;
    have = true;
    return mean + stddev * (u * m);
  }

  double perturb(Xoshiro& prng, double stddev_mean, double r_pert, List const& nl, long centre, double const* r_in,
                 std::uint8_t const* fzn, long n, double* r_out, double* ax_out) {
    Normal normal{0., 1.};
    Normal prime{stddev_mean, stddev_mean / 3.0};
    double sd = -1;
    while (sd < 0) sd = prime(prng);
    Normal gauss{0, sd};
    for (long k = 0; k < 3 * n; ++k) {
      r_out[k] = r_in[k];
      ax_out[k] = 0;
    }
    // Neighbours are visited in ascending real index (the reference visits them in list order).
    std::vector<std::pair<long, double>> nb;
    nl.for_neighbours(centre, r_pert, [&](long j, double r, V3 const&) { nb.emplace_back(j, r); });
    std::sort(nb.begin(), nb.end());
    for (auto const& [j, r] : nb) {
      if (fzn[j]) continue;
      for (int k = 0; k < 3; ++k) r_out[3 * j + k] += gauss(prng) * (1. - r / r_pert);
      for (int k = 0; k < 3; ++k) ax_out[3 * j + k] += normal(prng);
    }
    if (fzn[centre]) throw RuntimeError("perturbation centred on a frozen atom");
    for (int k = 0; k < 3; ++k) r_out[3 * centre + k] += gauss(prng);
    for (int k = 0; k < 3; ++k) ax_out[3 * centre + k] += normal(prng);
    double nrm = 0;
    for (long k = 0; k < 3 * n; ++k) nrm += ax_out[k] * ax_out[k];
    nrm = std::sqrt(nrm);
    for (long k = 0; k < 3 * n; ++k) ax_out[k] /= nrm;
    return sd;
  }

}  // namespace orc
