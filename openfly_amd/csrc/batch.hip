// Batched saddle searches on gfx950, host side: the device-side counterpart of Master::find_batch
// (find.cpp:684-755: one OpenMP task per search running Master::find_sp -> Dimer::find_sp with one
// thread, find.cpp:983-1009).  The search kernel itself is in search_kernel.hpp.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <tuple>
#include <type_traits>

#include "perturb.hpp"
#include "search_kernel.hpp"


// =====================================================================================================
// Host side: resident systems, region extraction, packed launches ("flights"), re-queue, autotuning,
// the batch C-ABI (Master::find_batch) and the serving C-ABI (continuous batching over many systems).
// =====================================================================================================

namespace flyg {
  void eam_density(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* d_types, double* d_rho);

  namespace sk {
    void const* search_kernel(int T, int L, bool mix, bool multi) {
      switch (T) {
        case 256: return mix ? search_kernel_256_mixed(L, multi) : search_kernel_256_fp64(L, multi);
        case 512: return mix ? search_kernel_512_mixed(L, multi) : search_kernel_512_fp64(L, multi);
        case 1024: return mix ? search_kernel_1024_mixed(L, multi) : search_kernel_1024_fp64(L, multi);
        default: return nullptr;
      }
    }
  }  // namespace sk

  using namespace sk;

  namespace {

    // Periodic cell grid over the base positions (orthorhombic box), cells of at least `rmin` / 2: a
    // query of radius ~rmin visits the cells its sphere touches, about a third of the candidates that
    // cells of rmin (a 5 x 5 x 5 block) would give.
    struct HostGrid {
      double L[3];
      int nc[3];
      double cell[3];
      std::vector<int> start, items;
      double const* pos = nullptr;

      void build(double const* p, long n, double const* ext, double rmin) {
        pos = p;
        for (int k = 0; k < 3; ++k) {
          L[k] = ext[k];
          nc[k] = std::max(1, static_cast<int>(std::floor(L[k] / (0.5 * rmin))));
          cell[k] = L[k] / nc[k];
        }
        long nct = static_cast<long>(nc[0]) * nc[1] * nc[2];
        std::vector<int> cnt(nct + 1, 0), cid(n);
        for (long i = 0; i < n; ++i) {
          int c3[3];
          for (int k = 0; k < 3; ++k) {
            double w = p[3 * i + k] - L[k] * std::floor(p[3 * i + k] / L[k]);
            c3[k] = std::min(nc[k] - 1, std::max(0, static_cast<int>(w / cell[k])));
          }
          cid[i] = c3[0] + nc[0] * (c3[1] + nc[1] * c3[2]);
          ++cnt[cid[i] + 1];
        }
        for (long c = 0; c < nct; ++c) cnt[c + 1] += cnt[c];
        start = cnt;
        items.resize(n);
        std::vector<int> cur(cnt.begin(), cnt.end() - 1);
        for (long i = 0; i < n; ++i) items[cur[cid[i]]++] = static_cast<int>(i);
      }

      // f(j, shift code, r2) for every image of every atom within `rad` of the canonical point p.
      template <typename F>
      void query(double const* p, double rad, F&& f) const {
        int ci[3], reach[3];
        for (int k = 0; k < 3; ++k) {
          ci[k] = std::min(nc[k] - 1, std::max(0, static_cast<int>(p[k] / cell[k])));
          reach[k] = static_cast<int>(std::ceil(rad / cell[k]));
        }
        double const r2max = rad * rad;
        // Distance from p to the cell at unwrapped index cc along axis k (0 inside it), less a margin that
        // covers the rounding of the binning, so no cell holding an atom within rad is skipped.
        auto gap = [&](int k, int cc) {
          double const lo = cc * cell[k] - p[k], hi = (cc + 1) * cell[k] - p[k];
          double const g = lo > 0 ? lo : (hi < 0 ? -hi : 0.0);
          return std::max(0.0, g - 1e-9 * L[k]);
        };
        for (int oz = -reach[2]; oz <= reach[2]; ++oz) {
          double const gz = gap(2, ci[2] + oz);
          if (gz * gz > r2max) continue;
          for (int oy = -reach[1]; oy <= reach[1]; ++oy) {
            double const gy = gap(1, ci[1] + oy);
            if (gz * gz + gy * gy > r2max) continue;
            for (int ox = -reach[0]; ox <= reach[0]; ++ox) {
              double const gx = gap(0, ci[0] + ox);
              if (gz * gz + gy * gy + gx * gx > r2max) continue;
              int o[3] = {ox, oy, oz}, cw[3], sh[3];
              for (int k = 0; k < 3; ++k) {
                int cc = ci[k] + o[k];
                sh[k] = cc >= 0 ? cc / nc[k] : -((-cc + nc[k] - 1) / nc[k]);
                cw[k] = cc - sh[k] * nc[k];
              }
              if (std::abs(sh[0]) > 1 || std::abs(sh[1]) > 1 || std::abs(sh[2]) > 1) continue;
              int c = cw[0] + nc[0] * (cw[1] + nc[1] * cw[2]);
              for (int q = start[c]; q < start[c + 1]; ++q) {
                int j = items[q];
                double d0 = pos[3 * j] + sh[0] * L[0] - p[0];
                double d1 = pos[3 * j + 1] + sh[1] * L[1] - p[1];
                double d2 = pos[3 * j + 2] + sh[2] * L[2] - p[2];
                double r2 = d0 * d0 + d1 * d1 + d2 * d2;
                if (r2 <= r2max) f(j, shift_code(sh[0], sh[1], sh[2]), r2);
              }
            }
          }
        }
      }
    };

    // One base system (a relaxed minimum), resident on the device while any of its searches is queued
    // or running: canonical positions, FP64 base densities, types and the box's image shifts.
    struct SysState {
      int64_t id = 0;
      long n = 0;
      double basis[9];
      int periodic[3];
      BoxDesc box{};
      std::vector<double> base;       // canonical, strictly inside [0, L)
      std::vector<double> raw;        // the caller's minimum as given (perturbations start from it)
      std::vector<uint32_t> types;
      HostGrid grid;
      DBuf<double4> d_base;
      DBuf<double> d_rho;
      DBuf<uint8_t> d_type;
      DBuf<double3> d_S;
      int single_type = -1;           // the one species present, -1 if several
      long active = 0;                // searches queued or in flight
      bool retired = false;           // removal requested; freed once idle
      std::size_t dev_bytes() const { return d_base.cap * sizeof(double4) + d_rho.cap * 8 + d_type.cap + d_S.cap * sizeof(double3); }
      double skin_cap(double rc) const {
        return 0.5 * std::fmin(box.ext[0], std::fmin(box.ext[1], box.ext[2])) - rc;
      }
    };

    // Device bytes a system will hold, and the transient peak of building its base densities.
    std::size_t sys_bytes_estimate(long n) { return static_cast<std::size_t>(n) * (32 + 8 + 1) + 27 * 24 + 4096; }
    std::size_t sys_transient_estimate(long n) { return static_cast<std::size_t>(n) * 700; }

    // Device scratch of an admission (the full list that gives the base densities), kept by a server
    // across admissions so that admitting a system frees nothing while other searches run.
    struct AdmitScratch {
      std::unique_ptr<flyg_list_s> nl;
      DBuf<double> pos;
      DBuf<uint32_t> types;
    };

    std::unique_ptr<SysState> make_system(flyg_ctx_s* ctx, flyg_eam_s* eam, double const basis[9], int const periodic[3],
                                          double const* min_pos, uint32_t const* types, long n, double list_skin,
                                          AdmitScratch* scratch = nullptr) {
      auto S = std::make_unique<SysState>();
      S->n = n;
      std::memcpy(S->basis, basis, sizeof(S->basis));
      std::memcpy(S->periodic, periodic, sizeof(S->periodic));
      S->box = make_box(basis, periodic);
      require(S->box.ortho && periodic[0] && periodic[1] && periodic[2],
              "batched searches need a fully periodic orthorhombic box");
      require(n < MAX_ATOMS, "too many atoms");
      S->base.resize(3 * n);
      S->raw.assign(min_pos, min_pos + 3 * n);
      S->types.assign(types, types + n);
      S->single_type = static_cast<int>(types[0]);
      for (long i = 1; i < n; ++i)
        if (types[i] != types[0]) {
          S->single_type = -1;
          break;
        }
      std::vector<double4> b4(n);
      std::vector<uint8_t> t8(n);
      for (long i = 0; i < n; ++i) {
        require(types[i] < static_cast<uint32_t>(eam->tab.ntypes), "TypeID out of range");
        for (int k = 0; k < 3; ++k) {
          // Strictly inside [0, L): the canonical image of a coordinate just below 0 rounds to L, and the
          // cell grid and the image shifts must agree on which side of the box an atom is.
          double x = min_pos[3 * i + k];
          double w = x - S->box.ext[k] * std::floor(x * S->box.inv_ext[k]);
          if (w >= S->box.ext[k]) w -= S->box.ext[k];
          if (w < 0) w = 0;
          S->base[3 * i + k] = w;
        }
        b4[i] = make_double4(S->base[3 * i], S->base[3 * i + 1], S->base[3 * i + 2], 0.0);
        t8[i] = static_cast<uint8_t>(types[i]);
      }
      double rl = eam->tab.rc + list_skin;
      for (int k = 0; k < 3; ++k) require(S->box.ext[k] > 2.0 * rl, "box too small for the search region cut-off");
      S->grid.build(S->base.data(), n, S->box.ext, rl);
      hipStream_t st = ctx->stream;
      S->d_base.reserve(n);
      S->d_rho.reserve(n);
      S->d_type.reserve(n);
      S->d_S.reserve(27);
      double3 sh[27];
      shift_table(S->box, sh);
      HIPCHK(hipMemcpyAsync(S->d_base.get(), b4.data(), n * sizeof(double4), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(S->d_type.get(), t8.data(), n, hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(S->d_S.get(), sh, sizeof(sh), hipMemcpyHostToDevice, st));
      {   // global FP64 base densities from a full list at r_c
        AdmitScratch own;
        AdmitScratch& A = scratch ? *scratch : own;
        if (!A.nl) {
          A.nl = make_list(ctx, basis, periodic, eam->tab.rc);
        } else {   // same buffers, this system's box
          A.nl->box = make_box(basis, periodic);
          A.nl->grid = make_grid(A.nl->box, eam->tab.rc);
          A.nl->rc = eam->tab.rc;
          A.nl->built = false;
        }
        A.pos.reserve(3 * n);
        A.types.reserve(n);
        HIPCHK(hipMemcpyAsync(A.pos.get(), S->base.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(A.types.get(), types, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        list_rebuild(*A.nl, A.pos.get(), n);
        eam_density(*eam, *A.nl, A.types.get(), S->d_rho.get());
        unsigned fl = ctx->take_flags();   // synchronises the stream before the scratch goes away
        if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
        if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
      }
      return S;
    }

    struct Region {
      int nm = 0, nr = 0;
      std::vector<int> gidx;
      std::vector<int> rowlen;      // nr
      std::vector<unsigned> ent;    // packed (shift, local)
    };

    // Known saddles shared by the searches of one submission (Master's per-site cache, find.cpp:692-732):
    // sp_k - min over all atoms (the caller's frame) and the norms of those vectors.
    struct HistGroup {
      int nh = 0;
      long n3 = 0;
      std::vector<double> d;
      std::vector<double> norm;
    };

    struct Region;

    // A region being built in the background for the skins a search was queued with.
    struct RegionTask {
      std::mutex m;
      std::condition_variable cv;
      bool done = false;
      double skin_f = 0, skin_m = 0;
      std::shared_ptr<Region> region;
      std::exception_ptr err;
    };

    struct Pending {
      int64_t id;
      int launches = 0;
      double cos_tol = 2.0;               // cos(theta_tol) of the history test (no collision above 1)
      std::shared_ptr<HistGroup> hist;    // may be null
      // Master::perturb recipe of the start (re-perturbed after a spline-out-of-bounds failure).
      bool perturbed = false;
      Xoshiro256 prng{{0, 0, 0, 0}};
      double stddev = 0, r_pert = 0;
      int retries = 0;
      double skin0 = 0;                   // the skin and mechanism flag the start was queued with
      bool mech = false;
      SysState* sys = nullptr;
      int64_t centre;
      double radius;
      double skin_f, skin_m;        // mobile-frozen / mobile-mobile Verlet skins
      std::vector<int> mobile;      // global indices (sorted)
      std::vector<double> x, ax;    // 3*nm start state (continuous with the base positions)
      std::shared_ptr<Region> region;   // built for the skins, reset when they grow
      std::shared_ptr<RegionTask> task; // its background build (server), used while the skins are unchanged
      // State saved at a region exit (resume_len doubles) and the loop position it belongs to: the
      // relaunch in the grown region continues from it instead of from the start.
      bool resume = false;
      std::vector<double> rs;
      int r_step = 0, r_convex = 0, r_k = 0, r_rot_iters = 0, r_refreshes = 0;
      long r_grads = 0;
      double r_trust = 0, r_maxd = 0, r_maxp = 0;
      double prior_ms = 0, prior_grad_ms = 0;   // device time of the launches before this one
      int resumes = 0;                           // launches that continued from a saved state
    };

    struct Result {
      int done = 0;
      flyg_search_stats st{};
      std::vector<int> mobile;
      std::vector<double> x, ax;
      int nr = 0;
      flyg_mechanism mech{-1, 0, 0.0, 0.0, {0.0, 0.0, 0.0}};
      std::vector<double> rev, fwd;   // mechanism minima (mobile atoms), when checked
    };

    double min_image(double d, double L) { return d - L * std::round(d / L); }

    void mobile_set(SysState const& S, int64_t centre, double R, std::vector<int>& out) {
      out.clear();
      S.grid.query(&S.base[3 * centre], R, [&](int j, int, double) { out.push_back(j); });
      std::sort(out.begin(), out.end());
      out.erase(std::unique(out.begin(), out.end()), out.end());
    }

    // Queues one search of system S: mobile set around `centre`, start state continuous with the base.
    // Start positions/axis come either dense (3n arrays) or sparse (listed atoms only; the rest at base).
    // With mechanisms on, the rows must also hold the two adjacent minima (a hop moves one atom by a
    // near-neighbour distance), so the skins start at mech_f / mech_m.
    Pending make_pending(SysState& S, double rc, int64_t id, int64_t centre, double radius, double skin,
                         double const* sp, double const* sa, int64_t const* kidx, long nk, bool mech = false) {
      require(centre >= 0 && centre < S.n, "centre out of range");
      require(radius > 0 && radius < 0.5 * std::fmin(S.box.ext[0], std::fmin(S.box.ext[1], S.box.ext[2])) - 1.0,
              "active radius must be positive and below half the box");
      Pending P;
      P.id = id;
      P.sys = &S;
      P.centre = centre;
      P.radius = radius;
      mobile_set(S, centre, radius, P.mobile);
      int nm = static_cast<int>(P.mobile.size());
      P.x.resize(3 * nm);
      P.ax.assign(3 * nm, 0.0);
      for (int i = 0; i < nm; ++i)
        for (int c = 0; c < 3; ++c) P.x[3 * i + c] = S.base[3 * P.mobile[i] + c];
      auto put = [&](long g, double const* p, double const* a) {
        auto it = std::lower_bound(P.mobile.begin(), P.mobile.end(), static_cast<int>(g));
        if (it == P.mobile.end() || *it != g) return;   // frozen atoms keep their base position
        long i = it - P.mobile.begin();
        for (int c = 0; c < 3; ++c) {
          double bse = S.base[3 * g + c];
          P.x[3 * i + c] = bse + min_image(p[c] - bse, S.box.ext[c]);
          P.ax[3 * i + c] = a[c];
        }
      };
      if (kidx) {
        for (long q = 0; q < nk; ++q) {
          require(kidx[q] >= 0 && kidx[q] < S.n, "perturbed atom index out of range");
          put(kidx[q], sp + 3 * q, sa + 3 * q);
        }
      } else {
        for (int i = 0; i < nm; ++i) put(P.mobile[i], sp + 3L * P.mobile[i], sa + 3L * P.mobile[i]);
      }
      // Skins: the start state itself (kicked atoms) must sit well inside the rows.
      double d1 = 0, d2 = 0;
      for (int i = 0; i < nm; ++i) {
        double dx = P.x[3 * i] - S.base[3 * P.mobile[i]], dy = P.x[3 * i + 1] - S.base[3 * P.mobile[i] + 1],
               dz = P.x[3 * i + 2] - S.base[3 * P.mobile[i] + 2];
        double d = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (d > d1) {
          d2 = d1;
          d1 = d;
        } else if (d > d2) {
          d2 = d;
        }
      }
      double cap = S.skin_cap(rc);
      double const mech_f = mech ? 3.0 : 0.0, mech_m = mech ? 4.5 : 0.0;
      P.skin_f = std::min(cap, std::max(std::max(skin, mech_f), d1 + 1.0));
      P.skin_m = std::min(cap, std::max(std::max(std::max(2.0 * skin, mech_m), P.skin_f), d1 + d2 + 1.5));
      return P;
    }

    // A search started from Master::perturb of the caller's minimum (find.cpp:700-711): the mobile set
    // decides which atoms are frozen, the cell grid supplies the candidates near the centre, and prng is
    // the search's own generator (advanced; a re-perturbation continues it).
    Pending make_perturbed(SysState& S, double rc, int64_t id, int64_t centre, double radius, double skin, bool mech,
                           Xoshiro256 prng, double stddev, double r_pert) {
      require(centre >= 0 && centre < S.n, "centre out of range");
      std::vector<int> mob;
      mobile_set(S, centre, radius, mob);
      std::vector<long> cand;
      S.grid.query(&S.base[3 * centre], r_pert + 0.5, [&](int j, int, double) { cand.push_back(j); });
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      Kick k;
      perturb(S.box, S.raw.data(), S.n, centre, &cand,
              [&](long j) { return !std::binary_search(mob.begin(), mob.end(), static_cast<int>(j)); }, stddev, r_pert,
              prng, k);
      Pending P = make_pending(S, rc, id, centre, radius, skin, k.pos.data(), k.axis.data(), k.atoms.data(),
                               static_cast<long>(k.atoms.size()), mech);
      P.perturbed = true;
      P.prng = prng;
      P.stddev = stddev;
      P.r_pert = r_pert;
      P.skin0 = skin;
      P.mech = mech;
      return P;
    }

    std::shared_ptr<HistGroup> make_hist(SysState const& S, double const* hist_pos, int n_hist) {
      if (n_hist <= 0) return nullptr;
      require(hist_pos != nullptr, "null history");
      auto H = std::make_shared<HistGroup>();
      H->nh = n_hist;
      H->n3 = 3 * S.n;
      H->d.resize(static_cast<std::size_t>(n_hist) * H->n3);
      H->norm.resize(n_hist);
      for (int q = 0; q < n_hist; ++q) {
        double nn = 0;
        for (long i = 0; i < H->n3; ++i) {
          double v = hist_pos[q * H->n3 + i] - S.raw[i];
          H->d[q * H->n3 + i] = v;
          nn += v * v;
        }
        H->norm[q] = std::sqrt(nn);
      }
      return H;
    }

    // Region of a search: mobile atoms, then S1 (frozen atoms within rF = r_c + skin_f of a mobile atom),
    // with rows of packed (shift, local index) entries in ascending order: a mobile row lists mobile atoms
    // within rM = r_c + skin_m and S1 atoms within rF; an S1 row lists the mobile atoms within rF.  One
    // grid query per mobile atom (rM >= rF) finds its row and the S1 atoms; the S1 rows are the mobile
    // rows' S1 entries transposed (the image of i seen from j has the opposite shift), so each pair is
    // tested once and the two rows always agree.
    void build_region(SysState const& S, std::vector<int> const& mobile, double skin_f, double skin_m, double rc,
                      std::vector<int>& local_of, Region& g) {
      double const rF = rc + skin_f, rM = rc + skin_m, rq = std::max(rF, rM);
      int const nm = static_cast<int>(mobile.size());
      g.nm = nm;
      g.gidx = mobile;
      for (int i = 0; i < nm; ++i) local_of[mobile[i]] = i;
      std::vector<int> s1, hj;
      std::vector<unsigned char> hs;
      std::vector<int> hoff(nm + 1, 0);
      for (int i = 0; i < nm; ++i) {
        int const gi = mobile[i];
        S.grid.query(&S.base[3 * gi], rq, [&](int j, int sc, double r2) {
          if (j == gi && sc == 13) return;
          int const lj = local_of[j];
          if (lj >= 0) {   // mobile
            if (!(r2 < rM * rM)) return;
          } else {
            if (!(r2 < rF * rF)) return;   // frozen beyond rF of this mobile atom
            if (lj == -1) {
              local_of[j] = -2;   // collected into S1
              s1.push_back(j);
            }
          }
          hj.push_back(j);
          hs.push_back(static_cast<unsigned char>(sc));
        });
        hoff[i + 1] = static_cast<int>(hj.size());
      }
      std::sort(s1.begin(), s1.end());
      for (std::size_t k = 0; k < s1.size(); ++k) local_of[s1[k]] = nm + static_cast<int>(k);
      g.gidx.insert(g.gidx.end(), s1.begin(), s1.end());
      g.nr = static_cast<int>(g.gidx.size());
      g.rowlen.assign(g.nr, 0);
      for (std::size_t h = 0; h < hj.size(); ++h) {
        int const lj = local_of[hj[h]];
        if (lj >= nm) ++g.rowlen[lj];
      }
      for (int i = 0; i < nm; ++i) g.rowlen[i] = hoff[i + 1] - hoff[i];
      std::vector<long> roff(g.nr + 1, 0);
      for (int r = 0; r < g.nr; ++r) roff[r + 1] = roff[r] + g.rowlen[r];
      g.ent.assign(roff[g.nr], 0u);
      std::vector<long> fill(roff.begin() + nm, roff.end() - 1);
      for (int i = 0; i < nm; ++i) {
        for (int h = hoff[i]; h < hoff[i + 1]; ++h) {
          unsigned const sc = hs[h];
          int const lj = local_of[hj[h]];
          g.ent[roff[i] + (h - hoff[i])] = (sc << NB_SHIFT_BITS) | static_cast<unsigned>(lj);
          if (lj >= nm) g.ent[fill[lj - nm]++] = ((26u - sc) << NB_SHIFT_BITS) | static_cast<unsigned>(i);
        }
      }
      // Each row in ascending packed order.  Its entries are distinct (shift code, local index < nr) pairs,
      // so a bitmap of local indices per shift code present orders a long row in O(nr / 64 + length).
      std::vector<uint64_t> bits((g.nr + 63) / 64, 0);
      std::vector<unsigned> tmp;
      for (int r = 0; r < g.nr; ++r) {
        unsigned* const b = g.ent.data() + roff[r];
        unsigned* const e = g.ent.data() + roff[r + 1];
        if (e - b < 48) {
          std::sort(b, e);
          continue;
        }
        uint32_t present = 0;
        for (unsigned const* q = b; q != e; ++q) present |= 1u << (*q >> NB_SHIFT_BITS);
        tmp.assign(b, e);
        unsigned* out = b;
        for (unsigned sc = 0; sc < 27; ++sc) {
          if (!(present & (1u << sc))) continue;
          for (unsigned en : tmp)
            if ((en >> NB_SHIFT_BITS) == sc) {
              unsigned const l = en & NB_IDX_MASK;
              bits[l >> 6] |= 1ull << (l & 63);
            }
          for (std::size_t w = 0; w < bits.size(); ++w) {
            uint64_t m = bits[w];
            bits[w] = 0;
            while (m) {
              unsigned const l = static_cast<unsigned>(w * 64 + static_cast<unsigned>(__builtin_ctzll(m)));
              *out++ = (sc << NB_SHIFT_BITS) | l;
              m &= m - 1;
            }
          }
        }
      }
      for (int j : g.gidx) local_of[j] = -1;
    }

    template <typename F>
    void parallel_for(long n, F&& f) {
      unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
      if (n < 4 || hw == 1) {
        for (long i = 0; i < n; ++i) f(i, 0);
        return;
      }
      std::vector<std::thread> th;
      unsigned nt = static_cast<unsigned>(std::min<long>(hw, n));
      for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
          for (long i = t; i < n; i += nt) f(i, t);
        });
      for (auto& x : th) x.join();
    }

    // Workers that build search regions in the background: a server queues each search's region when
    // the search is submitted, so a launch finds it built instead of building it on the caller's thread
    // (which also harvests results and takes new requests).  Each worker keeps a -1-filled scratch map.
    class RegionPool {
     public:
      explicit RegionPool(unsigned threads) {
        for (unsigned t = 0; t < threads; ++t) th_.emplace_back([this] { run(); });
      }
      ~RegionPool() {
        {
          std::lock_guard<std::mutex> l(m_);
          stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
      }
      RegionPool(RegionPool const&) = delete;
      RegionPool& operator=(RegionPool const&) = delete;

      // Builds P's region for its current skins on a worker; P.task receives it.
      void submit(Pending& P, double rc) {
        auto t = std::make_shared<RegionTask>();
        t->skin_f = P.skin_f;
        t->skin_m = P.skin_m;
        P.task = t;
        SysState const* S = P.sys;
        std::vector<int> mobile = P.mobile;
        {
          std::lock_guard<std::mutex> l(m_);
          q_.push_back([t, S, mobile = std::move(mobile), rc](std::vector<int>& scratch) {
            std::shared_ptr<Region> g;
            std::exception_ptr err;
            try {
              if (static_cast<long>(scratch.size()) < S->n) scratch.assign(S->n, -1);
              g = std::make_shared<Region>();
              build_region(*S, mobile, t->skin_f, t->skin_m, rc, scratch, *g);
            } catch (...) {
              err = std::current_exception();
              std::fill(scratch.begin(), scratch.end(), -1);
            }
            {
              std::lock_guard<std::mutex> l(t->m);
              t->region = std::move(g);
              t->err = err;
              t->done = true;
            }
            t->cv.notify_all();
          });
        }
        cv_.notify_one();
      }

     private:
      void run() {
        std::vector<int> scratch;
        while (true) {
          std::function<void(std::vector<int>&)> f;
          {
            std::unique_lock<std::mutex> l(m_);
            cv_.wait(l, [&] { return stop_ || !q_.empty(); });
            if (q_.empty()) return;   // stopping and drained
            f = std::move(q_.front());
            q_.pop_front();
          }
          f(scratch);
        }
      }
      std::mutex m_;
      std::condition_variable cv_;
      std::deque<std::function<void(std::vector<int>&)>> q_;
      bool stop_ = false;
      std::vector<std::thread> th_;
    };

    // Builds the missing regions of `work`: a background build for the search's current skins is
    // waited for, the rest are built here in parallel (one -1-filled scratch map per thread).
    void ensure_regions(std::vector<Pending>& work, double rc) {
      std::vector<long> need;
      long nmax = 0;
      for (auto& P : work) {
        if (P.region || !P.task) continue;
        RegionTask& t = *P.task;
        if (t.skin_f == P.skin_f && t.skin_m == P.skin_m) {
          std::unique_lock<std::mutex> l(t.m);
          t.cv.wait(l, [&] { return t.done; });
          if (t.err) std::rethrow_exception(t.err);
          P.region = t.region;
        }
        P.task.reset();
      }
      for (long k = 0; k < static_cast<long>(work.size()); ++k)
        if (!work[k].region) {
          need.push_back(k);
          nmax = std::max(nmax, work[k].sys->n);
        }
      if (need.empty()) return;
      std::vector<std::vector<int>> scratch(16);
      parallel_for(static_cast<long>(need.size()), [&](long q, unsigned t) {
        if (static_cast<long>(scratch[t].size()) < nmax) scratch[t].assign(nmax, -1);
        Pending& P = work[need[q]];
        auto g = std::make_shared<Region>();
        build_region(*P.sys, P.mobile, P.skin_f, P.skin_m, rc, scratch[t], *g);
        P.region = std::move(g);
      });
    }

    using KernelFn = void (*)(BatchArgs);

    KernelFn kernel_for(int T, int L, bool mix, bool multi) {
      void const* k = sk::search_kernel(T, L, mix, multi);
      if (!k)
        throw Error(FLYG_INVALID_ARG, "unsupported (block_threads, lanes_per_atom); use 256, 512 or 1024 with 8 or 16");
      return reinterpret_cast<KernelFn>(const_cast<void*>(k));
    }

    DimOpt dim_opt(flyg_dimer_options const& o, flyg_rotor_options const& r, flyg_mech_options const* m = nullptr) {
      DimOpt d{};
      if (m && m->enabled) {
        d.mech = 1;
        d.nudge_frac = m->nudge_frac;
        d.basin_tol = m->basin_tol;
        d.stationary_tol = m->stationary_tol;
        d.min_iter_max = m->minimiser.iter_max;
        d.min_f2norm = m->minimiser.f2norm;
        d.min_proj_tol = m->minimiser.proj_tol;
        d.min_max_trust = m->minimiser.max_trust;
        d.min_min_trust = m->minimiser.min_trust;
        d.min_grow_trust = m->minimiser.grow_trust;
        d.min_shrink_trust = m->minimiser.shrink_trust;
      }
      d.n = o.n;
      d.max_steps = o.max_steps;
      d.convex_max = o.convex_max;
      d.rn = r.n;
      d.iter_max_rot = r.iter_max_rot;
      d.relax_in_convex = r.relax_in_convex;
      d.f2norm = o.f2norm;
      d.proj_tol = o.proj_tol;
      d.max_trust = o.max_trust;
      d.min_trust = o.min_trust;
      d.grow_trust = o.grow_trust;
      d.shrink_trust = o.shrink_trust;
      d.delta_r = r.delta_r;
      d.theta_tol = r.theta_tol;
      d.hist_check_freq = std::max(1, o.hist_check_freq);
      d.use_history = o.use_history;
      return d;
    }

    // Slots per launch for the saved states of searches that leave their region (a search that finds
    // them all taken restarts from its start in the grown region instead of continuing).
    constexpr int RS_SLOTS = 64;

    // Pinned host buffer (hipHostMalloc), grown on demand.
    struct Pinned {
      char* p = nullptr;
      std::size_t cap = 0;
      Pinned() = default;
      Pinned(Pinned const&) = delete;
      Pinned& operator=(Pinned const&) = delete;
      ~Pinned() {
        if (p) (void)free_or_defer(p, cap, true);
      }
      void reserve(std::size_t n) {
        if (n <= cap) return;
        if (p) HIPCHK(free_or_defer(p, cap, true));
        p = nullptr;
        cap = 0;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&p), n, hipHostMallocDefault));
        cap = n;
      }
    };

    // Byte layout of one launch: sub-arrays at 256-byte aligned offsets of one allocation.
    struct Layout {
      std::size_t size = 0;
      template <typename T>
      std::size_t add(long count) {
        std::size_t o = (size + 255) & ~static_cast<std::size_t>(255);
        size = o + static_cast<std::size_t>(std::max<long>(count, 1)) * sizeof(T);
        return o;
      }
    };

    // One packed launch of the persistent search kernel on its own stream: uploads in one copy, then the
    // kernel, which writes each search's results and a done flag straight into pinned host memory, so
    // searches can be harvested one by one while the kernel still runs others.
    struct Flight {
      hipStream_t st = nullptr;
      bool own_stream = false;
      hipEvent_t e_a = nullptr, e_b = nullptr, e_done = nullptr;
      bool busy = false;
      std::vector<Pending> work;
      long n3 = 0;
      DBuf<char> d_in, d_scr;
      Pinned h_in, h_out;
      std::size_t o_out = 0, o_ox = 0, o_oax = 0, o_rev = 0, o_fwd = 0, out_bytes = 0;   // offsets in h_out
      std::size_t o_rs = 0;         // saved states of region exits (RS_SLOTS slots of rs_len doubles)
      long rs_len = 0;
      int dopt_n = 0;               // translation core size the states were saved with
      int mech = 0;
      std::vector<long> v3off;      // per search: first element of its 3*nm output vectors
      std::vector<char> harvested;  // per search
      long left = 0;                // searches not yet harvested
      Flight() = default;
      Flight(Flight const&) = delete;
      Flight& operator=(Flight const&) = delete;
      void init(hipStream_t s, bool own) {
        st = s;
        own_stream = own;
        HIPCHK(hipEventCreate(&e_a));
        HIPCHK(hipEventCreate(&e_b));
        HIPCHK(hipEventCreate(&e_done));
      }
      ~Flight() {
        if (st) (void)hipStreamSynchronize(st);
        if (e_a) (void)hipEventDestroy(e_a);
        if (e_b) (void)hipEventDestroy(e_b);
        if (e_done) (void)hipEventDestroy(e_done);
        if (own_stream && st) (void)hipStreamDestroy(st);
      }
      std::size_t dev_bytes() const { return d_in.cap + d_scr.cap; }
      void release_memory() {   // idle flights only
        DBuf<char> a(std::move(d_in)), b(std::move(d_scr));
      }
      bool finished() const {
        hipError_t e = hipEventQuery(e_done);
        if (e == hipErrorNotReady) return false;
        HIPCHK(e);
        return true;
      }
      float kernel_ms() const {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e_a, e_b));
        return ms;
      }
    };

    // Device bytes a launch of `work` needs (regions must be built).
    std::size_t flight_bytes(std::vector<Pending> const& work, DimOpt const& dopt) {
      std::size_t b = 4096;
      for (auto const& P : work) {
        Region const& g = *P.region;
        int const nh = P.hist ? P.hist->nh : 0;
        b += P.rs.size() * 8;
        b += sizeof(SearchHdr) + static_cast<std::size_t>(g.nr) * (4 + 4 + 4 + 32 + 32 + 1 + 8 + 8) + g.ent.size() * 8 +
             static_cast<std::size_t>(3 * g.nm) * 8 * (4 + nh + nvec(dopt.n, dopt.rn)) + 8 * nh + sizeof(SearchOut) +
             3 * 256;
      }
      return b;
    }

    // Packs `work` (regions built) into F and launches it asynchronously on F.st.
    void flight_launch(flyg_eam_s& E, Flight& F, std::vector<Pending>&& work, int T, int L, DimOpt const& dopt, bool mix,
                       int slots) {
      require(!F.busy, "internal: flight already in use");
      F.work = std::move(work);
      long const ns = static_cast<long>(F.work.size());
      long nreg = 0, nrow = 0, nent = 0, n3 = 0, nws = 0, nhv = 0, nhn = 0, nrs = 0, rs_len = 0;
      std::vector<SearchHdr> hdr(ns);
      for (long k = 0; k < ns; ++k) {
        Pending const& P = F.work[k];
        Region const& g = *P.region;
        require(g.nm > 0, "a search has no mobile atoms");
        require(g.nr < (1 << NB_SHIFT_BITS), "search region too large");
        SearchHdr& h = hdr[k];
        h.nm = g.nm;
        h.nr = g.nr;
        h.reg_off = nreg;
        h.row_off = nrow;
        h.v3_off = n3;
        h.ws_off = nws;
        h.skin_f = P.skin_f;
        h.skin_m = P.skin_m;
        h.sbase = P.sys->d_base.get();
        h.srho = P.sys->d_rho.get();
        h.stype = P.sys->d_type.get();
        h.sS = P.sys->d_S.get();
        h.natoms = P.sys->n;
        h.centre_local = static_cast<int>(std::lower_bound(P.mobile.begin(), P.mobile.end(), static_cast<int>(P.centre)) -
                                          P.mobile.begin());
        h.n_hist = P.hist ? P.hist->nh : 0;
        h.hist_off = nhv;
        h.hn_off = nhn;
        h.cos_tol = P.cos_tol;
        h.resume = P.resume ? 1 : 0;
        if (P.resume) {
          require(static_cast<long>(P.rs.size()) == resume_len(3 * g.nm, dopt.n), "internal: saved search state size");
          h.r_step = P.r_step;
          h.r_convex = P.r_convex;
          h.r_k = P.r_k;
          h.r_rot_iters = P.r_rot_iters;
          h.r_refreshes = P.r_refreshes;
          h.r_grads = P.r_grads;
          h.r_trust = P.r_trust;
          h.r_maxd = P.r_maxd;
          h.r_maxp = P.r_maxp;
          h.rs_in_off = nrs;
          nrs += static_cast<long>(P.rs.size());
        }
        rs_len = std::max(rs_len, resume_len(3 * g.nm, dopt.n));
        nhv += 3L * g.nm * h.n_hist;
        nhn += h.n_hist;
        nreg += g.nr;
        nrow += g.nr + 1;
        nent += static_cast<long>(g.ent.size());
        n3 += 3L * g.nm;
        nws += static_cast<long>(nvec(dopt.n, dopt.rn)) * 3L * g.nm;
      }
      require(nent < (1l << 31), "too many neighbour entries in one launch");
      F.n3 = n3;
      Layout in;
      std::size_t o_hdr = in.add<SearchHdr>(ns), o_gidx = in.add<int>(nreg), o_row = in.add<int>(nrow),
                  o_ent = in.add<unsigned>(nent), o_sx = in.add<double>(n3), o_sax = in.add<double>(n3),
                  o_hv = in.add<double>(nhv), o_hn = in.add<double>(nhn), o_rs = in.add<double>(nrs);
      Layout sc;
      std::size_t o_cur = sc.add<double4>(nreg), o_base = sc.add<double4>(nreg), o_type = sc.add<uint8_t>(nreg),
                  o_rhof = sc.add<double>(nreg), o_fp = sc.add<double>(nreg), o_act = sc.add<unsigned>(nent),
                  o_alen = sc.add<int>(nreg), o_ws = sc.add<double>(nws),
                  o_q = sc.add<unsigned>(1), o_rsc = sc.add<unsigned>(1);
      std::size_t zero_end = sc.size;
      Layout ho;
      std::size_t o_flags = ho.add<int>(ns);
      F.o_out = ho.add<SearchOut>(ns);
      F.o_ox = ho.add<double>(n3);
      F.o_oax = ho.add<double>(n3);
      F.o_rev = ho.add<double>(dopt.mech ? n3 : 0);
      F.o_fwd = ho.add<double>(dopt.mech ? n3 : 0);
      F.rs_len = rs_len;
      F.o_rs = ho.add<double>(RS_SLOTS * rs_len);
      F.out_bytes = ho.size;
      // Grow with slack: freeing device or pinned memory synchronises the whole device, which would
      // stall the other streams' launches.
      auto grow = [](std::size_t need, std::size_t cap) { return need > cap ? need + need / 2 : need; };
      F.h_in.reserve(grow(in.size, F.h_in.cap));
      F.h_out.reserve(grow(F.out_bytes, F.h_out.cap));
      char* hb = F.h_in.p;
      std::memcpy(hb + o_hdr, hdr.data(), ns * sizeof(SearchHdr));
      int* hg = reinterpret_cast<int*>(hb + o_gidx);
      int* hr = reinterpret_cast<int*>(hb + o_row);
      unsigned* he = reinterpret_cast<unsigned*>(hb + o_ent);
      double* hx = reinterpret_cast<double*>(hb + o_sx);
      double* ha = reinterpret_cast<double*>(hb + o_sax);
      long e0 = 0;
      for (long k = 0; k < ns; ++k) {
        Pending const& P = F.work[k];
        Region const& g = *P.region;
        std::copy(g.gidx.begin(), g.gidx.end(), hg + hdr[k].reg_off);
        long rp = hdr[k].row_off;
        long e1 = e0;
        for (int r = 0; r < g.nr; ++r) {
          hr[rp + r] = static_cast<int>(e1);
          e1 += g.rowlen[r];
        }
        hr[rp + g.nr] = static_cast<int>(e1);
        std::copy(g.ent.begin(), g.ent.end(), he + e0);
        e0 = e1;
        std::copy(P.x.begin(), P.x.end(), hx + hdr[k].v3_off);
        std::copy(P.ax.begin(), P.ax.end(), ha + hdr[k].v3_off);
        if (hdr[k].n_hist > 0) {   // the known saddles' components on this search's mobile atoms
          HistGroup const& H = *P.hist;
          double* hv = reinterpret_cast<double*>(hb + o_hv) + hdr[k].hist_off;
          for (int q = 0; q < H.nh; ++q)
            for (int i = 0; i < g.nm; ++i)
              for (int c = 0; c < 3; ++c) hv[(static_cast<long>(q) * g.nm + i) * 3 + c] = H.d[q * H.n3 + 3L * P.mobile[i] + c];
          std::copy(H.norm.begin(), H.norm.end(), reinterpret_cast<double*>(hb + o_hn) + hdr[k].hn_off);
        }
        if (P.resume) std::copy(P.rs.begin(), P.rs.end(), reinterpret_cast<double*>(hb + o_rs) + hdr[k].rs_in_off);
      }
      F.d_in.reserve(grow(in.size, F.d_in.cap));
      F.d_scr.reserve(grow(sc.size, F.d_scr.cap));
      char* di = F.d_in.get();
      char* ds = F.d_scr.get();
      HIPCHK(hipMemcpyAsync(di, hb, in.size, hipMemcpyHostToDevice, F.st));
      HIPCHK(hipMemsetAsync(ds + o_ws, 0, zero_end - o_ws, F.st));   // workspace and queue counter

      BatchArgs A{};
      A.T = E.tab;
      A.hdr = reinterpret_cast<SearchHdr const*>(di + o_hdr);
      A.nsearch = static_cast<int>(ns);
      A.cur = reinterpret_cast<double4*>(ds + o_cur);
      A.base = reinterpret_cast<double4*>(ds + o_base);
      A.rtype = reinterpret_cast<uint8_t*>(ds + o_type);
      A.rgidx = reinterpret_cast<int const*>(di + o_gidx);
      A.rhof = reinterpret_cast<double*>(ds + o_rhof);
      A.fp = reinterpret_cast<double*>(ds + o_fp);
      A.rowptr = reinterpret_cast<int const*>(di + o_row);
      A.ent = reinterpret_cast<unsigned const*>(di + o_ent);
      A.act = reinterpret_cast<unsigned*>(ds + o_act);
      A.alen = reinterpret_cast<int*>(ds + o_alen);
      A.ws = reinterpret_cast<double*>(ds + o_ws);
      A.start_x = reinterpret_cast<double const*>(di + o_sx);
      A.start_ax = reinterpret_cast<double const*>(di + o_sax);
      A.hist = reinterpret_cast<double const*>(di + o_hv);
      A.hnorm = reinterpret_cast<double const*>(di + o_hn);
      A.rs_in = reinterpret_cast<double const*>(di + o_rs);
      A.rs_count = reinterpret_cast<unsigned*>(ds + o_rsc);
      A.rs_slots = RS_SLOTS;
      A.rs_len = rs_len;
      A.rs_out = reinterpret_cast<double*>(F.h_out.p + F.o_rs);
      char* hb_out = F.h_out.p;
      std::memset(hb_out + o_flags, 0, ns * sizeof(int));
      A.done = reinterpret_cast<int*>(hb_out + o_flags);
      A.out_x = reinterpret_cast<double*>(hb_out + F.o_ox);
      A.out_ax = reinterpret_cast<double*>(hb_out + F.o_oax);
      A.out_rev = reinterpret_cast<double*>(hb_out + F.o_rev);
      A.out_fwd = reinterpret_cast<double*>(hb_out + F.o_fwd);
      F.mech = dopt.mech;
      F.dopt_n = dopt.n;
      A.out = reinterpret_cast<SearchOut*>(hb_out + F.o_out);
      F.v3off.resize(ns);
      for (long k = 0; k < ns; ++k) F.v3off[k] = hdr[k].v3_off;
      F.harvested.assign(ns, 0);
      F.left = ns;
      A.queue = reinterpret_cast<unsigned*>(ds + o_q);
      A.o = dopt;

      unsigned grid = static_cast<unsigned>(std::max<long>(1, std::min<long>(slots, ns)));
      // One species present in every system of the launch: species indices become kernel constants.
      int t0 = F.work.empty() ? -1 : F.work[0].sys->single_type;
      for (auto const& P : F.work)
        if (P.sys->single_type != t0) t0 = -1;
      A.t0 = t0 < 0 ? 0 : t0;
      KernelFn fn = kernel_for(T, L, mix, E.tab.ntypes > 1 && t0 < 0);
      HIPCHK(hipEventRecord(F.e_a, F.st));
      hipLaunchKernelGGL(fn, dim3(grid), dim3(T), 0, F.st, A);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(F.e_b, F.st));
      HIPCHK(hipEventRecord(F.e_done, F.st));
      F.busy = true;
    }

    double wall_clock_khz() {
      static thread_local int khz = 0;
      if (!khz) {
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) khz = 100000;
      }
      return static_cast<double>(khz);
    }

    // Harvests the searches of F whose done flags are set (all of them once the kernel has finished):
    // finished searches go to `done`, searches that left their region (and may still grow it) come back
    // in `requeue` with larger skins.  The flight becomes free when the kernel has finished and every
    // search is harvested.
    template <typename Done>
    void flight_harvest(Flight& F, double rc, Done&& done, std::vector<Pending>& requeue) {
      bool const finished = F.finished();
      int const* flags = reinterpret_cast<int const*>(F.h_out.p);
      SearchOut const* outs = reinterpret_cast<SearchOut const*>(F.h_out.p + F.o_out);
      double const* ox = reinterpret_cast<double const*>(F.h_out.p + F.o_ox);
      double const* oax = reinterpret_cast<double const*>(F.h_out.p + F.o_oax);
      double const* orv = reinterpret_cast<double const*>(F.h_out.p + F.o_rev);
      double const* ofw = reinterpret_cast<double const*>(F.h_out.p + F.o_fwd);
      for (long k = 0; k < static_cast<long>(F.work.size()); ++k) {
        if (F.harvested[k]) continue;
        if (!finished && __atomic_load_n(&flags[k], __ATOMIC_ACQUIRE) == 0) continue;
        F.harvested[k] = 1;
        --F.left;
        Pending& P = F.work[k];
        SearchOut const& o = outs[k];
        int nm = P.region->nm;
        long off = F.v3off[k];
        ++P.launches;
        double cap = P.sys->skin_cap(rc);
        // Spline out of bounds: re-perturb and run again, at most three times (find.cpp:706-723).
        if (o.exit_code == EXIT_OOB && P.perturbed && P.retries < 3) {
          Pending q = make_perturbed(*P.sys, rc, P.id, P.centre, P.radius, P.skin0, P.mech, P.prng, P.stddev, P.r_pert);
          q.retries = P.retries + 1;
          q.cos_tol = P.cos_tol;
          q.hist = P.hist;
          requeue.push_back(std::move(q));
          continue;
        }
        if (o.exit_code == EXIT_REGION && (P.skin_f < cap || P.skin_m < cap)) {
          Pending q = P;
          q.skin_f = std::min(cap, std::max(q.skin_f * 1.5, o.max_disp + 1.0));
          q.skin_m = std::min(cap, std::max(std::max(q.skin_m * 1.5, q.skin_f), o.max_pair + 1.5));
          q.region.reset();
          q.task.reset();
          if (o.saved > 0 && o.saved <= RS_SLOTS) {   // continue where it left its region
            double const* src = reinterpret_cast<double const*>(F.h_out.p + F.o_rs) + (o.saved - 1) * F.rs_len;
            q.resume = true;
            q.rs.assign(src, src + resume_len(3 * nm, F.dopt_n));
            q.r_step = o.steps;
            q.r_convex = o.convex;
            q.r_k = o.lbfgs_k;
            q.r_rot_iters = o.rot_iters;
            q.r_refreshes = o.refreshes;
            q.r_grads = o.grads;
            q.r_trust = o.trust;
            q.r_maxd = o.max_disp;
            q.r_maxp = o.max_pair;
            q.prior_ms += static_cast<double>(o.ticks) / wall_clock_khz();
            q.prior_grad_ms += static_cast<double>(o.grad_ticks) / wall_clock_khz();
            ++q.resumes;
          } else {   // restart from the start state (the counters start over too)
            q.resume = false;
            q.rs.clear();
            q.prior_ms = q.prior_grad_ms = 0;
            q.resumes = 0;
          }
          requeue.push_back(std::move(q));
          continue;
        }
        Result R;
        R.done = 1;
        R.st.exit_code = o.exit_code;
        R.st.steps = o.steps;
        R.st.grads = o.grads;
        R.st.rebuilds = o.refreshes;
        R.st.curv = o.curv;
        R.st.energy = o.barrier;
        R.st.max_disp = o.max_disp;
        R.st.launches = P.launches;
        R.st.retries = P.retries;
        R.st.resumes = P.resumes;
        R.st.rot_iters = o.rot_iters;
        R.st.device_ms = P.prior_ms + static_cast<double>(o.ticks) / wall_clock_khz();
        R.st.device_grad_ms = P.prior_grad_ms + static_cast<double>(o.grad_ticks) / wall_clock_khz();
        R.mobile = P.mobile;
        R.x.assign(ox + off, ox + off + 3L * nm);
        R.ax.assign(oax + off, oax + off + 3L * nm);
        R.mech.code = o.mech_code;
        R.mech.min_iters = o.min_iters;
        R.mech.barrier = o.mech_barrier;
        R.mech.delta = o.mech_delta;
        for (int c = 0; c < 3; ++c) R.mech.drift[c] = o.drift[c];
        if (F.mech && o.mech_code >= 0 && o.mech_code != 1) {
          R.rev.assign(orv + off, orv + off + 3L * nm);
          R.fwd.assign(ofw + off, ofw + off + 3L * nm);
        }
        R.nr = P.region->nr;
        done(P, std::move(R));
      }
      if (finished && F.left == 0) {
        F.work.clear();
        F.busy = false;
      }
    }

    // Times a 20-step prefix of up to 64 searches with each candidate shape on F (synchronously).
    void autotune_shape(flyg_eam_s& E, Flight& F, std::vector<Pending> const& queue, DimOpt d, bool mix, int slots,
                        int& T_out, int& L_out) {
      std::vector<Pending> work;
      for (auto const& P : queue) {
        if (work.size() >= 64) break;
        work.push_back(P);
      }
      if (work.empty()) return;
      ensure_regions(work, E.tab.rc);
      d.max_steps = std::min(d.max_steps, 20);
      int const cand[][2] = {{256, 8}, {256, 16}, {512, 8}, {512, 16}, {1024, 8}, {1024, 16}};
      double best = 1e300;
      for (auto const& c : cand) {
        std::vector<Pending> w = work;
        flight_launch(E, F, std::move(w), c[0], c[1], d, mix, slots);
        HIPCHK(hipEventSynchronize(F.e_done));
        double ms = F.kernel_ms();
        F.work.clear();
        F.busy = false;
        if (ms < best) {
          best = ms;
          T_out = c[0];
          L_out = c[1];
        }
      }
    }

    void fill_positions(SysState const& S, Result const& R, double* pos, double* axis) {
      if (pos) {
        std::memcpy(pos, S.base.data(), 3 * S.n * sizeof(double));
        for (std::size_t i = 0; i < R.mobile.size() && 3 * i + 2 < R.x.size(); ++i)
          for (int c = 0; c < 3; ++c) pos[3 * R.mobile[i] + c] = R.x[3 * i + c];
      }
      if (axis) {
        std::memset(axis, 0, 3 * S.n * sizeof(double));
        for (std::size_t i = 0; i < R.mobile.size() && 3 * i + 2 < R.ax.size(); ++i)
          for (int c = 0; c < 3; ++c) axis[3 * R.mobile[i] + c] = R.ax[3 * i + c];
      }
    }

  }  // namespace
}  // namespace flyg

struct flyg_batch_s {
  flyg::Graveyard grave;   // buffers given up while relaunches overlap the main launch (freed after the run)
  flyg_ctx_s* ctx = nullptr;
  flyg_eam_s* eam = nullptr;
  flyg_dimer_options opt;
  flyg_rotor_options ropt;
  flyg_batch_options bopt;
  std::unique_ptr<flyg::SysState> sys;
  std::vector<flyg::Pending> pending;
  std::vector<flyg::Result> results;
  flyg_mech_options mopt{};                              // mechanisms off unless set
  int tuned_T = 0, tuned_L = 0;
  flyg::Flight flight;                                   // on the context's stream
  std::vector<std::unique_ptr<flyg::Flight>> extra;      // relaunches, on streams of their own
};

struct flyg_server_s {
  // Buffers given up while searches run (grown launch buffers, removed systems, admission scratch):
  // freed when no launch is in flight, or at once when admission needs the memory.  First member, so
  // it is destroyed last.
  flyg::Graveyard grave;
  flyg_ctx_s* ctx = nullptr;
  flyg_eam_s* eam = nullptr;
  flyg_dimer_options opt;
  flyg_rotor_options ropt;
  flyg_server_options sopt;
  flyg_mech_options mopt{};
  std::map<int64_t, std::unique_ptr<flyg::SysState>> systems;
  flyg::AdmitScratch admit;
  long admit_n = 0;   // largest system the admission scratch has held (its size is estimated from it)
  int64_t next_sys = 0;
  std::deque<flyg::Pending> queue;
  std::vector<std::unique_ptr<flyg::Flight>> flights;
  std::map<int64_t, flyg::Result> results;
  std::map<int64_t, int64_t> owner;            // search id -> system id
  int64_t next_id = 0;
  std::deque<int64_t> completed;
  std::size_t budget = 0;
  int tuned_T = 0, tuned_L = 0;
  int64_t launches = 0, searches_done = 0;
  double kernel_ms = 0;
  double host_ms = 0;   // region builds and launch packing

  // Background region builds of submitted searches (last member: its workers finish before the systems
  // they read go away).
  std::unique_ptr<flyg::RegionPool> pool;

  std::size_t used() const {
    std::size_t b = grave.bytes + flyg::sys_transient_estimate(admit_n);
    for (auto const& kv : systems) b += kv.second->dev_bytes();
    for (auto const& f : flights) b += f->dev_bytes();
    return b;
  }
};

namespace flyg {
  namespace {

    int slots_of(flyg_ctx_s const& ctx, int max_concurrent) { return max_concurrent > 0 ? max_concurrent : 2 * ctx.cus; }

    // Harvests finished flights, frees idle retired systems, and launches queued searches on free
    // flights (as many as the memory budget allows, FIFO).  Returns how many searches finished.
    long server_step(flyg_server_s& S) {
      long fin = 0;
      double const rc = S.eam->tab.rc;
      bool any_busy = false;
      for (auto const& fp : S.flights) any_busy = any_busy || fp->busy;
      if (!any_busy || S.grave.bytes > S.budget / 4) S.grave.drain();
      for (auto& fp : S.flights) {
        Flight& F = *fp;
        if (!F.busy) continue;
        std::vector<Pending> requeue;
        flight_harvest(F, rc, [&](Pending& P, Result&& R) {
          --P.sys->active;
          S.results[P.id] = std::move(R);
          S.completed.push_back(P.id);
          ++S.searches_done;
          ++fin;
        }, requeue);
        if (!F.busy) S.kernel_ms += F.kernel_ms();
        for (auto it = requeue.rbegin(); it != requeue.rend(); ++it) S.queue.push_front(std::move(*it));
      }
      for (auto it = S.systems.begin(); it != S.systems.end();) {
        if (it->second->retired && it->second->active == 0) {
          it = S.systems.erase(it);
        } else {
          ++it;
        }
      }
      if (S.queue.empty()) return fin;
      DimOpt d = dim_opt(S.opt, S.ropt, &S.mopt);
      bool mix = S.sopt.precision == FLYG_MIXED;
      int slots = slots_of(*S.ctx, S.sopt.max_concurrent);
      if (!S.tuned_T) {
        if (S.sopt.block_threads > 0 && S.sopt.lanes_per_atom > 0) {
          S.tuned_T = S.sopt.block_threads;
          S.tuned_L = S.sopt.lanes_per_atom;
        } else {
          for (auto& fp : S.flights) {
            if (fp->busy) continue;
            std::vector<Pending> q(S.queue.begin(), S.queue.end());
            autotune_shape(*S.eam, *fp, q, d, mix, slots, S.tuned_T, S.tuned_L);
            break;
          }
          if (!S.tuned_T) return fin;
        }
      }
      int const cap = S.sopt.max_batch > 0 ? S.sopt.max_batch : 4 * S.ctx->cus;
      for (auto& fp : S.flights) {
        if (S.queue.empty()) break;
        Flight& F = *fp;
        if (F.busy) continue;
        std::vector<Pending> work;
        while (!S.queue.empty() && static_cast<int>(work.size()) < cap) {
          work.push_back(std::move(S.queue.front()));
          S.queue.pop_front();
        }
        auto const h0 = std::chrono::steady_clock::now();
        ensure_regions(work, rc);
        // Memory gate: keep the prefix that fits beside everything resident (at least one search when
        // nothing else is in flight, so the queue always drains).
        std::size_t other = S.used() - F.dev_bytes();
        bool alone = true;
        for (auto const& g : S.flights)
          if (g->busy) alone = false;
        while (work.size() > 1 || (!alone && !work.empty())) {
          std::size_t need = std::max(flight_bytes(work, d), F.dev_bytes());
          if (other + need <= S.budget) break;
          S.queue.push_front(std::move(work.back()));
          work.pop_back();
        }
        if (work.empty()) {
          S.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
          break;
        }
        flight_launch(*S.eam, F, std::move(work), S.tuned_T, S.tuned_L, d, mix, slots);
        S.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
        ++S.launches;
      }
      return fin;
    }

  }  // namespace
}  // namespace flyg

using namespace flyg;

namespace {
  thread_local std::string g_berr;
  template <typename F>
  int bguard(flyg_ctx_s* ctx, F&& f) {
    try {
      if (ctx) ctx->set_device();
      f();
      return FLYG_OK;
    } catch (Error const& e) {
      (ctx ? ctx->err : g_berr) = e.what();
      return e.status;
    } catch (std::exception const& e) {
      (ctx ? ctx->err : g_berr) = e.what();
      return FLYG_INVALID_ARG;
    }
  }

  void take_options(flyg_dimer_options const* opt, flyg_rotor_options const* ropt, flyg_dimer_options& o,
                    flyg_rotor_options& r) {
    if (opt) {
      o = *opt;
    } else {
      flyg_dimer_default_options(&o);
    }
    if (ropt) {
      r = *ropt;
    } else {
      flyg_rotor_default_options(&r);
    }
    require(o.n > 0 && o.n <= 32 && r.n > 0 && r.n <= 32, "batched searches support L-BFGS histories up to 32");
  }
}  // namespace

extern "C" {

void flyg_batch_default_options(flyg_batch_options* o) { *o = flyg_batch_options{FLYG_MIXED, 0, 0, 2.0, 0}; }

int flyg_batch_create(flyg_ctx ctx, flyg_eam eam, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                      const flyg_batch_options* bopt, const double basis[9], const int periodic[3],
                      const double* min_pos, const uint32_t* types, int64_t n, flyg_batch* out) {
  return bguard(ctx, [&] {
    require(ctx && eam && basis && periodic && min_pos && types && out && n > 0, "null or empty argument");
    require(eam->ctx == ctx, "potential belongs to another context");
    auto b = std::make_unique<flyg_batch_s>();
    b->ctx = ctx;
    b->eam = eam;
    take_options(opt, ropt, b->opt, b->ropt);
    if (bopt) {
      b->bopt = *bopt;
    } else {
      flyg_batch_default_options(&b->bopt);
    }
    if (b->bopt.list_skin <= 0) b->bopt.list_skin = 2.0;
    require(b->bopt.precision == FLYG_MIXED || b->bopt.precision == FLYG_FP64, "bad precision");
    b->sys = make_system(ctx, eam, basis, periodic, min_pos, types, n, b->bopt.list_skin);
    b->flight.init(ctx->stream, false);
    ctx->retain();
    *out = b.release();
  });
}

int flyg_batch_destroy(flyg_batch b) {
  if (!b) return FLYG_OK;
  flyg_ctx_s* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete b;
  ctx->release();
  return FLYG_OK;
}

int flyg_batch_submit(flyg_batch b, int count, const int64_t* centre, const double* radius, const double* start_pos,
                      const double* start_axis, int64_t* ids) {
  return flyg_batch_submit_ex(b, count, centre, radius, start_pos, start_axis, 0.0, nullptr, 0, ids);
}

int flyg_batch_submit_ex(flyg_batch b, int count, const int64_t* centre, const double* radius, const double* start_pos,
                         const double* start_axis, double theta_tol, const double* hist_pos, int n_hist, int64_t* ids) {
  return bguard(b->ctx, [&] {
    require(count >= 0 && centre && radius && start_pos && start_axis, "null argument");
    require(n_hist >= 0, "negative history length");
    long const n = b->sys->n;
    auto H = make_hist(*b->sys, hist_pos, n_hist);
    std::vector<Pending> add;
    for (int k = 0; k < count; ++k) {
      add.push_back(make_pending(*b->sys, b->eam->tab.rc, static_cast<int64_t>(b->results.size()) + k, centre[k],
                                 radius[k], b->bopt.list_skin, start_pos + 3 * n * static_cast<long>(k),
                                 start_axis + 3 * n * static_cast<long>(k), nullptr, 0, b->mopt.enabled != 0));
      add.back().cos_tol = std::cos(theta_tol);
      add.back().hist = H;
    }
    for (int k = 0; k < count; ++k) {
      if (ids) ids[k] = add[k].id;
      b->pending.push_back(std::move(add[k]));
      b->results.emplace_back();
    }
  });
}

int flyg_batch_submit_perturbed(flyg_batch b, int count, const int64_t* centre, const double* radius, uint64_t* prng,
                                double stddev, double r_pert, double theta_tol, const double* hist_pos, int n_hist,
                                int64_t* ids) {
  return bguard(b->ctx, [&] {
    require(count >= 0 && centre && radius && prng, "null argument");
    require(n_hist >= 0, "negative history length");
    auto H = make_hist(*b->sys, hist_pos, n_hist);
    std::vector<Pending> add;
    std::vector<Xoshiro256> gens;
    for (int k = 0; k < count; ++k) {
      Xoshiro256 g{{prng[4 * k], prng[4 * k + 1], prng[4 * k + 2], prng[4 * k + 3]}};
      add.push_back(make_perturbed(*b->sys, b->eam->tab.rc, static_cast<int64_t>(b->results.size()) + k, centre[k],
                                   radius[k], b->bopt.list_skin, b->mopt.enabled != 0, g, stddev, r_pert));
      add.back().cos_tol = std::cos(theta_tol);
      add.back().hist = H;
    }
    for (int k = 0; k < count; ++k) {
      for (int c = 0; c < 4; ++c) prng[4 * k + c] = add[k].prng.s[c];
      if (ids) ids[k] = add[k].id;
      b->pending.push_back(std::move(add[k]));
      b->results.emplace_back();
    }
  });
}

int flyg_batch_run(flyg_batch b) {
  return bguard(b->ctx, [&] {
    if (b->pending.empty()) return;
    FreeLater defer(b->grave);
    DimOpt d = dim_opt(b->opt, b->ropt, &b->mopt);
    bool mix = b->bopt.precision == FLYG_MIXED;
    int slots = slots_of(*b->ctx, b->bopt.max_concurrent);
    if (b->bopt.block_threads > 0 && b->bopt.lanes_per_atom > 0) {
      b->tuned_T = b->bopt.block_threads;
      b->tuned_L = b->bopt.lanes_per_atom;
    } else if (!b->tuned_T) {
      // Batches over the same potential, system size and precision reuse the first one's tuning (the
      // shape changes no result, only the time).
      static std::mutex mu;
      static std::map<std::tuple<flyg_eam_s const*, long, bool>, std::pair<int, int>> shapes;
      auto const key = std::make_tuple(static_cast<flyg_eam_s const*>(b->eam), static_cast<long>(b->sys->n), mix);
      {
        std::lock_guard<std::mutex> lock(mu);
        auto it = shapes.find(key);
        if (it != shapes.end()) {
          b->tuned_T = it->second.first;
          b->tuned_L = it->second.second;
        }
      }
      if (!b->tuned_T) {
        autotune_shape(*b->eam, b->flight, b->pending, d, mix, slots, b->tuned_T, b->tuned_L);
        std::lock_guard<std::mutex> lock(mu);
        shapes[key] = {b->tuned_T, b->tuned_L};
      }
    }
    // All queued searches go out in one launch; a search that leaves its region is relaunched with
    // larger skins on another stream as soon as its flag shows it finished, overlapping the rest.
    constexpr int MAX_LAUNCHES = 6;
    std::vector<Pending> todo = std::move(b->pending);
    b->pending.clear();
    double const rc = b->eam->tab.rc;
    auto give_up = [&](Pending& P) {   // still outside the largest region the box allows
      auto& R = b->results[P.id];
      R.done = 1;
      R.st.exit_code = EXIT_REGION;
      R.st.launches = P.launches;
      R.st.resumes = P.resumes;
      R.mobile = P.mobile;
    };
    auto free_flight = [&]() -> Flight* {
      if (!b->flight.busy) return &b->flight;
      for (auto& f : b->extra)
        if (!f->busy) return f.get();
      if (b->extra.size() >= 3) return nullptr;
      hipStream_t st;
      HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      b->extra.push_back(std::make_unique<Flight>());
      b->extra.back()->init(st, true);
      return b->extra.back().get();
    };
    while (true) {
      bool launched = false;
      if (!todo.empty()) {
        std::vector<Pending> go;
        for (auto& P : todo) {
          if (P.launches >= MAX_LAUNCHES) {
            give_up(P);
          } else {
            go.push_back(std::move(P));
          }
        }
        todo.clear();
        if (!go.empty()) {
          Flight* F = free_flight();
          if (F) {
            ensure_regions(go, rc);
            flight_launch(*b->eam, *F, std::move(go), b->tuned_T, b->tuned_L, d, mix, slots);
            launched = true;
          } else {
            todo = std::move(go);
          }
        }
      }
      bool busy = false;
      auto harvest = [&](Flight& F) {
        if (!F.busy) return;
        flight_harvest(F, rc, [&](Pending& P, Result&& R) { b->results[P.id] = std::move(R); }, todo);
        busy = busy || F.busy;
      };
      harvest(b->flight);
      for (auto& f : b->extra) harvest(*f);
      if (!busy && todo.empty()) break;
      if (!launched) std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    b->grave.drain();   // every launch has finished
  });
}

int flyg_batch_result(flyg_batch b, int64_t id, flyg_search_stats* stats, double* pos, double* axis) {
  return bguard(b->ctx, [&] {
    require(id >= 0 && id < static_cast<int64_t>(b->results.size()), "unknown search id");
    auto const& R = b->results[id];
    require(R.done, "search has not run yet");
    if (stats) *stats = R.st;
    fill_positions(*b->sys, R, pos, axis);
  });
}

void flyg_mech_default_options(flyg_mech_options* o) {
  flyg_mech_options m{};
  m.enabled = 0;
  m.nudge_frac = 0.01;
  m.basin_tol = 0.05;
  m.stationary_tol = 0.025;
  flyg_lbfgs_default_options(&m.minimiser);
  *o = m;
}

namespace {
  // Smallest minimiser tolerance |g| the mixed form can reach: its FP32 pair forces carry rounding
  // noise of order 1e-5 eV/A over a few hundred mobile components.
  constexpr double MIXED_MIN_F2NORM = 1e-4;

  void check_mech(flyg_mech_options const* opt, bool mixed) {
    require(opt != nullptr, "null argument");
    require(opt->nudge_frac > 0 && opt->basin_tol > 0 && opt->stationary_tol > 0, "mechanism tolerances must be positive");
    require(opt->minimiser.iter_max > 0 && opt->minimiser.f2norm > 0, "bad minimiser options");
    require(!opt->enabled || !mixed || opt->minimiser.f2norm >= MIXED_MIN_F2NORM,
            "mixed precision: the mechanism relaxations need minimiser.f2norm >= 1e-4 (FP32 pair forces carry "
            "rounding noise near 1e-5 eV/A); use FP64 precision for tighter tolerances");
  }
}  // namespace

int flyg_batch_set_mechanisms(flyg_batch b, const flyg_mech_options* opt) {
  return bguard(b->ctx, [&] {
    check_mech(opt, b->bopt.precision == FLYG_MIXED);
    b->mopt = *opt;
  });
}

int flyg_batch_mechanism(flyg_batch b, int64_t id, flyg_mechanism* mech, double* rev_pos, double* fwd_pos) {
  return bguard(b->ctx, [&] {
    require(id >= 0 && id < static_cast<int64_t>(b->results.size()), "unknown search id");
    auto const& R = b->results[id];
    require(R.done, "search has not run yet");
    if (mech) *mech = R.mech;
    auto fill = [&](std::vector<double> const& v, double* out) {
      if (!out) return;
      std::memcpy(out, b->sys->base.data(), 3 * b->sys->n * sizeof(double));
      for (std::size_t i = 0; i < R.mobile.size() && 3 * i + 2 < v.size(); ++i)
        for (int c = 0; c < 3; ++c) out[3 * R.mobile[i] + c] = v[3 * i + c];
    };
    require((!rev_pos && !fwd_pos) || !R.rev.empty(), "no mechanism minima for this search");
    fill(R.rev, rev_pos);
    fill(R.fwd, fwd_pos);
  });
}

int flyg_batch_shape(flyg_batch b, int* block_threads, int* lanes_per_atom) {
  return bguard(b->ctx, [&] {
    require(block_threads && lanes_per_atom, "null argument");
    *block_threads = b->tuned_T;
    *lanes_per_atom = b->tuned_L;
  });
}

int flyg_batch_region(flyg_batch b, int64_t id, int* n_mobile, int* n_region) {
  return bguard(b->ctx, [&] {
    require(id >= 0 && id < static_cast<int64_t>(b->results.size()), "unknown search id");
    auto const& R = b->results[id];
    if (n_mobile) *n_mobile = static_cast<int>(R.mobile.size());
    if (n_region) *n_region = R.nr;
  });
}

// ------------------------------------------------------------------------------------------------
// Serving
// ------------------------------------------------------------------------------------------------

void flyg_server_default_options(flyg_server_options* o) {
  *o = flyg_server_options{FLYG_MIXED, 0, 0, 2.0, 0, 4, 0, 0};
}

int flyg_server_create(flyg_ctx ctx, flyg_eam eam, const flyg_dimer_options* opt, const flyg_rotor_options* ropt,
                       const flyg_server_options* sopt, flyg_server* out) {
  return bguard(ctx, [&] {
    require(ctx && eam && out, "null argument");
    require(eam->ctx == ctx, "potential belongs to another context");
    auto s = std::make_unique<flyg_server_s>();
    s->pool = std::make_unique<RegionPool>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
    s->ctx = ctx;
    s->eam = eam;
    take_options(opt, ropt, s->opt, s->ropt);
    if (sopt) {
      s->sopt = *sopt;
    } else {
      flyg_server_default_options(&s->sopt);
    }
    if (s->sopt.list_skin <= 0) s->sopt.list_skin = 2.0;
    if (s->sopt.streams <= 0) s->sopt.streams = 4;
    require(s->sopt.streams <= 32, "at most 32 streams");
    require(s->sopt.precision == FLYG_MIXED || s->sopt.precision == FLYG_FP64, "bad precision");
    if (s->sopt.mem_budget > 0) {
      s->budget = static_cast<std::size_t>(s->sopt.mem_budget);
    } else {
      std::size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      s->budget = fr / 10 * 8;
    }
    for (int k = 0; k < s->sopt.streams; ++k) {
      hipStream_t st;
      HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      auto f = std::make_unique<Flight>();
      f->init(st, true);
      s->flights.push_back(std::move(f));
    }
    ctx->retain();
    *out = s.release();
  });
}

int flyg_server_destroy(flyg_server s) {
  if (!s) return FLYG_OK;
  flyg_ctx_s* ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  delete s;   // flights synchronise their streams first
  ctx->release();
  return FLYG_OK;
}

int flyg_server_add_system(flyg_server s, const double basis[9], const int periodic[3], const double* min_pos,
                           const uint32_t* types, int64_t n, int64_t* sys_id) {
  return bguard(s->ctx, [&] {
    require(basis && periodic && min_pos && types && sys_id && n > 0, "null or empty argument");
    FreeLater defer(s->grave);
    auto need_for = [&] {
      return sys_bytes_estimate(n) + (n > s->admit_n ? sys_transient_estimate(n) - sys_transient_estimate(s->admit_n) : 0);
    };
    if (s->used() + need_for() > s->budget) {   // idle launch buffers, admission scratch and given-up ones go first
      for (auto& f : s->flights)
        if (!f->busy) f->release_memory();
      s->admit = AdmitScratch{};
      s->admit_n = 0;
      s->grave.drain();
    }
    std::size_t const need = need_for();
    if (s->used() + need > s->budget) {
      throw Error(FLYG_BUSY, "device memory budget full: admit the system after running searches finish");
    }
    auto sys = make_system(s->ctx, s->eam, basis, periodic, min_pos, types, n, s->sopt.list_skin, &s->admit);
    s->admit_n = std::max(s->admit_n, static_cast<long>(n));
    sys->id = s->next_sys++;
    *sys_id = sys->id;
    s->systems[sys->id] = std::move(sys);
  });
}

int flyg_server_remove_system(flyg_server s, int64_t sys_id) {
  return bguard(s->ctx, [&] {
    FreeLater defer(s->grave);
    auto it = s->systems.find(sys_id);
    require(it != s->systems.end() && !it->second->retired, "unknown system id");
    it->second->retired = true;
    if (it->second->active == 0) s->systems.erase(it);
  });
}

int flyg_server_submit(flyg_server s, int64_t sys_id, int count, const int64_t* centre, const double* radius,
                       const int64_t* kick_offsets, const int64_t* kick_atoms, const double* kick_pos,
                       const double* kick_axis, int64_t* ids) {
  return bguard(s->ctx, [&] {
    auto it = s->systems.find(sys_id);
    require(it != s->systems.end() && !it->second->retired, "unknown system id");
    require(count >= 0 && centre && radius && kick_offsets && (kick_offsets[count] == 0 || (kick_atoms && kick_pos && kick_axis)),
            "null argument");
    SysState& S = *it->second;
    std::vector<Pending> add;
    for (int k = 0; k < count; ++k) {
      long a = kick_offsets[k], b = kick_offsets[k + 1];
      require(a >= 0 && b >= a, "kick offsets must be non-decreasing");
      add.push_back(make_pending(S, s->eam->tab.rc, s->next_id + k, centre[k], radius[k], s->sopt.list_skin, kick_pos + 3 * a,
                                 kick_axis + 3 * a, kick_atoms + a, b - a, s->mopt.enabled != 0));
    }
    for (int k = 0; k < count; ++k) {
      int64_t id = s->next_id++;
      s->owner[id] = sys_id;
      if (ids) ids[k] = id;
      s->pool->submit(add[k], s->eam->tab.rc);
      s->queue.push_back(std::move(add[k]));
      ++S.active;
    }
  });
}

int flyg_server_submit_perturbed(flyg_server s, int64_t sys_id, int count, const int64_t* centre, const double* radius,
                                 uint64_t* prng, double stddev, double r_pert, int64_t* ids) {
  return bguard(s->ctx, [&] {
    auto it = s->systems.find(sys_id);
    require(it != s->systems.end() && !it->second->retired, "unknown system id");
    require(count >= 0 && centre && radius && prng, "null argument");
    SysState& S = *it->second;
    std::vector<Pending> add;
    for (int k = 0; k < count; ++k) {
      Xoshiro256 g{{prng[4 * k], prng[4 * k + 1], prng[4 * k + 2], prng[4 * k + 3]}};
      add.push_back(make_perturbed(S, s->eam->tab.rc, s->next_id + k, centre[k], radius[k], s->sopt.list_skin,
                                   s->mopt.enabled != 0, g, stddev, r_pert));
    }
    for (int k = 0; k < count; ++k) {
      for (int c = 0; c < 4; ++c) prng[4 * k + c] = add[k].prng.s[c];
      int64_t id = s->next_id++;
      s->owner[id] = sys_id;
      if (ids) ids[k] = id;
      s->pool->submit(add[k], s->eam->tab.rc);
      s->queue.push_back(std::move(add[k]));
      ++S.active;
    }
  });
}

int flyg_server_progress(flyg_server s, double wait_ms, int64_t* done_ids, int cap, int* n_done) {
  return bguard(s->ctx, [&] {
    FreeLater defer(s->grave);
    require(n_done && (cap == 0 || done_ids), "null argument");
    auto t0 = std::chrono::steady_clock::now();
    while (true) {
      server_step(*s);
      if (!s->completed.empty() || wait_ms <= 0) break;
      bool busy = false;
      for (auto const& f : s->flights) busy = busy || f->busy;
      if (!busy && s->queue.empty()) break;
      double el = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (el >= wait_ms) break;
      std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    int k = 0;
    while (k < cap && !s->completed.empty()) {
      done_ids[k++] = s->completed.front();
      s->completed.pop_front();
    }
    *n_done = k;
  });
}

int flyg_server_result(flyg_server s, int64_t id, flyg_search_stats* stats, int* n_mobile, int64_t* atoms, double* pos,
                       double* axis) {
  return bguard(s->ctx, [&] {
    auto it = s->results.find(id);
    require(it != s->results.end() && it->second.done, "search has not finished (or unknown id)");
    Result const& R = it->second;
    if (stats) *stats = R.st;
    if (n_mobile) *n_mobile = static_cast<int>(R.mobile.size());
    for (std::size_t i = 0; i < R.mobile.size(); ++i) {
      if (atoms) atoms[i] = R.mobile[i];
      for (int c = 0; c < 3; ++c) {
        if (pos && 3 * i + c < R.x.size()) pos[3 * i + c] = R.x[3 * i + c];
        if (axis && 3 * i + c < R.ax.size()) axis[3 * i + c] = R.ax[3 * i + c];
      }
    }
  });
}

int flyg_server_release(flyg_server s, int64_t id) {
  return bguard(s->ctx, [&] {
    s->results.erase(id);
    s->owner.erase(id);
  });
}

int flyg_server_set_mechanisms(flyg_server s, const flyg_mech_options* opt) {
  return bguard(s->ctx, [&] {
    check_mech(opt, s->sopt.precision == FLYG_MIXED);
    s->mopt = *opt;
  });
}

int flyg_server_mechanism(flyg_server s, int64_t id, flyg_mechanism* mech, double* rev, double* fwd) {
  return bguard(s->ctx, [&] {
    auto it = s->results.find(id);
    require(it != s->results.end() && it->second.done, "search has not finished (or unknown id)");
    Result const& R = it->second;
    if (mech) *mech = R.mech;
    require((!rev && !fwd) || !R.rev.empty(), "no mechanism minima for this search");
    if (rev) std::memcpy(rev, R.rev.data(), R.rev.size() * sizeof(double));
    if (fwd) std::memcpy(fwd, R.fwd.data(), R.fwd.size() * sizeof(double));
  });
}

int flyg_server_status(flyg_server s, flyg_server_state* st) {
  return bguard(s->ctx, [&] {
    require(st, "null argument");
    st->queued = static_cast<int64_t>(s->queue.size());
    int64_t infl = 0;
    int busy = 0;
    for (auto const& f : s->flights)
      if (f->busy) {
        infl += static_cast<int64_t>(f->work.size());
        ++busy;
      }
    st->in_flight = infl;
    st->busy_streams = busy;
    st->systems = static_cast<int64_t>(s->systems.size());
    st->device_bytes = static_cast<int64_t>(s->used());
    st->budget_bytes = static_cast<int64_t>(s->budget);
    st->launches = s->launches;
    st->searches_done = s->searches_done;
    st->kernel_ms = s->kernel_ms;
    st->block_threads = s->tuned_T;
    st->lanes_per_atom = s->tuned_L;
    st->host_ms = s->host_ms;
  });
}

}  // extern "C"
