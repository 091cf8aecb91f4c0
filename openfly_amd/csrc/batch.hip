// Batched saddle searches on gfx950: the device-side counterpart of Master::find_batch
// (find.cpp:684-755: one OpenMP task per search running Master::find_sp -> Dimer::find_sp with one
// thread, find.cpp:983-1009).
//
// One persistent kernel; each workgroup pulls searches from a queue (one atomic per search) and runs the
// whole dimer search (Rotor::eff_gradient rotor.cpp:49-164, Dimer::find_sp dimer.cpp:28-171, both L-BFGS
// cores lbfgs_core.cpp:29-87) with workgroup barriers only.  A search works on its active region:
//   M  = mobile atoms (minimum-image distance to the centre <= radius, the rest of the system is Frozen),
//   S1 = frozen atoms within r_L = r_cut + skin of a mobile atom (their density, hence F', can change).
// Neighbour rows are built once from the base positions with cut-off r_L (a Verlet skin): mobile rows
// hold every region neighbour, S1 rows only the mobile ones, and the frozen part of an S1 density is the
// base density minus the mobile part.  The rows stay complete while d1 + d2 < skin for the two largest
// mobile displacements from the base; a search that breaks this is re-run with a larger skin.  Every 3N vector of
// the reference is zero on frozen atoms, so the device keeps only the 3*|M| mobile components.
// Barrier = E_loc(saddle) - E_loc(base) in FP64, where E_loc collects every term of the EAM energy that
// depends on mobile positions (sum over M u S1 of F(rho) plus the pairs with a mobile member).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "plain.hpp"

namespace flyg {

  namespace {

    constexpr int EXIT_REGION = 5;
    constexpr int EXIT_OOB = 6;
    // Workspace vectors per search (each 3*nm doubles): 10 named vectors, then the translation core
    // (m s-vectors, m y-vectors, x_prev, g_prev, q) and the rotation core (the same with its own m).
    inline int nvec(int m, int rm) { return 16 + 2 * (m + rm); }

    struct SearchHdr {
      int nm, nr;          // mobile atoms, region atoms (M u S1)
      long reg_off;        // first region atom in the packed region arrays
      long row_off;        // first row pointer (nr + 1 entries)
      long v3_off;         // first element of this search's 3*nm start/out vectors
      long ws_off;         // first element of this search's workspace (nvec * 3*nm)
      double skin;         // Verlet skin of the rows (r_L = r_c + skin)
    };

    struct SearchOut {
      int exit_code, steps, rot_iters, pad;
      long grads;
      double curv, barrier, e_base, max_disp;
    };

    struct DimOpt {
      int n, max_steps, convex_max, rn, iter_max_rot, relax_in_convex;
      double f2norm, proj_tol, max_trust, min_trust, grow_trust, shrink_trust, delta_r, theta_tol;
    };

    struct BatchArgs {
      EAMTables T;
      double3 S[27];
      SearchHdr const* hdr;
      int nsearch;
      double4* cur;              // region positions (base for S1, current evaluation point for M)
      double4 const* base;
      uint8_t const* rtype;
      int const* rgidx;
      double const* rho_base;    // global FP64 base densities
      double* rhof;              // frozen part of S1 densities
      double* fp;                // F'(rho) per region atom
      int const* rowptr;
      unsigned const* ent;
      double* ws;
      double const* start_x;
      double const* start_ax;
      double* out_x;
      double* out_ax;
      SearchOut* out;
      unsigned* queue;
      DimOpt o;
      int multi;
    };

    // ---------------------------------------------------------------------------------------------
    // Workgroup primitives.  Element i of a 3*nm vector is always handled by thread i % T, so
    // elementwise chains need no barrier; reductions are fixed-order trees (bitwise reproducible).
    // ---------------------------------------------------------------------------------------------

    template <int T>
    struct Team {
      double* red;   // LDS, 4 * T
      int tid;

      __device__ double wave(double v) const {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
      }

      // Sums up to four values over the workgroup with a tree that does not depend on T: element i of a
      // vector belongs to virtual slot i mod 512, each slot adds in increasing i, and the 512 slots meet
      // in the xor tree 256, 128, ..., 1 (levels >= T in registers, 64..T/2 through LDS, < 64 by
      // shuffles).  v[q][k] is the slot tid + q T of value k; every thread gets the totals.
      static constexpr int VS = 512;
      static constexpr int NQ = VS / T;

      template <int K>
      __device__ void tree(double (&v)[NQ][4]) const {
#pragma unroll
        for (int bit = VS / 2; bit >= T; bit >>= 1)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
            if (((q * T) & bit) == 0 && q + bit / T < NQ)
#pragma unroll
              for (int k = 0; k < K; ++k) v[q][k] = __dadd_rn(v[q][k], v[q + bit / T][k]);
        for (int bit = T / 2; bit >= 64; bit >>= 1) {
#pragma unroll
          for (int k = 0; k < K; ++k) red[4 * tid + k] = v[0][k];
          __syncthreads();
#pragma unroll
          for (int k = 0; k < K; ++k) v[0][k] = __dadd_rn(v[0][k], red[4 * (tid ^ bit) + k]);
          __syncthreads();
        }
#pragma unroll
        for (int bit = 32; bit >= 1; bit >>= 1)
#pragma unroll
          for (int k = 0; k < K; ++k) v[0][k] = __dadd_rn(v[0][k], __shfl_xor(v[0][k], bit, 64));
      }

      template <int K>
      __device__ void sum(double* out, double (&v)[NQ][4]) const {
        tree<K>(v);
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = v[0][k];
      }

      __device__ double dot(double const* a, double const* b, int n) const {
        double v[NQ][4] = {};
        for (int i = tid, q = 0; i < n; i += T, q = (q + 1) % NQ) v[q][0] = __dadd_rn(v[q][0], a[i] * b[i]);
        double o[1];
        sum<1>(o, v);
        return o[0];
      }

      __device__ void dot2(double const* a, double const* b, double const* c, double const* d, int n, double& x,
                           double& y) const {
        double v[NQ][4] = {};
        for (int i = tid, q = 0; i < n; i += T, q = (q + 1) % NQ) {
          v[q][0] = __dadd_rn(v[q][0], a[i] * b[i]);
          v[q][1] = __dadd_rn(v[q][1], c[i] * d[i]);
        }
        double o[2];
        sum<2>(o, v);
        x = o[0];
        y = o[1];
      }

      // Workgroup sum of per-thread values that are already in a fixed per-row form (region energy).
      __device__ double sum1_rows(double x) const {
        double v[NQ][4] = {};
        v[0][0] = x;
        double o[1];
        sum<1>(o, v);
        return o[0];
      }

      // o = ca a + cb b + cc c (b, c may be null); o may alias a.
      __device__ void lin(double* o, int n, double const* a, double ca, double const* b = nullptr, double cb = 0,
                          double const* c = nullptr, double cc = 0) const {
        for (int i = tid; i < n; i += T) {
          double v = ca * a[i];
          if (b) v += cb * b[i];
          if (c) v += cc * c[i];
          o[i] = v;
        }
      }

      __device__ unsigned any_or(unsigned e) const {
        __shared__ unsigned flag;
        if (tid == 0) flag = 0;
        __syncthreads();
        if (e) atomicOr(&flag, e);
        __syncthreads();
        unsigned r = flag;
        __syncthreads();
        return r;
      }
    };

    // ---------------------------------------------------------------------------------------------
    // Per-search view
    // ---------------------------------------------------------------------------------------------

    struct SView {
      int nm, nr, n3;
      double4* cur;
      double4 const* base;
      uint8_t const* type;
      int const* gidx;
      double* rhof;
      double* fp;
      int const* rowptr;
      unsigned const* ent;
    };

    __device__ inline int vix(int nt, int a, int b) { return a > b ? a + nt * b : b + nt * a; }

    // EAM gradient of the mobile atoms at the current region positions (pass 1 rho/F' over M u S1,
    // pass 2 forces over M).  MIX: FP32 pair/spline arithmetic with FP64 accumulation.
    template <int T, int L, bool MIX>
    __device__ unsigned region_gradient(BatchArgs const& A, SView const& v, Team<T> const& tm, double* g) {
      EAMTables const& Tb = A.T;
      int const nt = Tb.ntypes;
      unsigned e = 0;
      int const lane = tm.tid % L;
      for (int r0 = tm.tid / L; r0 < v.nr; r0 += T / L) {
        double4 pi = v.cur[r0];
        int ti = A.multi ? v.type[r0] : 0;
        Acc16<L> rho_acc;
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        for (int k = b + lane, m = 0; k < c; k += L, ++m) {
          unsigned en = v.ent[k];
          int j = static_cast<int>(en & NB_IDX_MASK);
          double3 sh = A.S[en >> NB_SHIFT_BITS];
          double4 pj = v.cur[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < Tb.rc2) {
            int tj = A.multi ? v.type[j] : 0;
            if (MIX) {
              rho_acc.add(m, static_cast<double>(spl32f<0>(Tb, Tb.phi[tj + nt * ti], sqrtf(static_cast<float>(r2)), e)));
            } else {
              rho_acc.add(m, spl64<0>(Tb, Tb.phi[tj + nt * ti], sqrt(r2), e));
            }
          }
        }
        double rho = rho_acc.total();
        if (lane == 0) {
          if (r0 >= v.nm) rho += v.rhof[r0];
          v.fp[r0] = spl64<1>(Tb, Tb.F[ti], rho, e);
        }
      }
      __syncthreads();
      for (int r0 = tm.tid / L; r0 < v.nm; r0 += T / L) {
        double4 pi = v.cur[r0];
        double fpi = v.fp[r0];
        int ti = A.multi ? v.type[r0] : 0;
        Acc16<L> gx, gy, gz;
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        for (int k = b + lane, m = 0; k < c; k += L, ++m) {
          unsigned en = v.ent[k];
          int j = static_cast<int>(en & NB_IDX_MASK);
          double3 sh = A.S[en >> NB_SHIFT_BITS];
          double4 pj = v.cur[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < Tb.rc2) {
            int tj = A.multi ? v.type[j] : 0;
            double fpj = v.fp[j];
            double cf;
            if (MIX) {
              float rf = sqrtf(static_cast<float>(r2));
              float vp = spl32f<1>(Tb, Tb.V[vix(nt, ti, tj)], rf, e);
              float pa = spl32f<1>(Tb, Tb.phi[tj + nt * ti], rf, e);
              float pb = A.multi ? spl32f<1>(Tb, Tb.phi[ti + nt * tj], rf, e) : pa;
              float mag = vp + static_cast<float>(fpi) * pa + static_cast<float>(fpj) * pb;
              cf = static_cast<double>(mag / rf);
            } else {
              double r = sqrt(r2);
              double vp = spl64<1>(Tb, Tb.V[vix(nt, ti, tj)], r, e);
              double pa = spl64<1>(Tb, Tb.phi[tj + nt * ti], r, e);
              double pb = A.multi ? spl64<1>(Tb, Tb.phi[ti + nt * tj], r, e) : pa;
              cf = (vp + fpi * pa + fpj * pb) / r;
            }
            gx.add(m, -cf * dx);
            gy.add(m, -cf * dy);
            gz.add(m, -cf * dz);
          }
        }
        double tx = gx.total(), ty = gy.total(), tz = gz.total();
        if (lane == 0) {
          g[3 * r0] = tx;
          g[3 * r0 + 1] = ty;
          g[3 * r0 + 2] = tz;
        }
      }
      __syncthreads();
      return tm.any_or(e);
    }

    // FP64 local energy (all terms that depend on mobile positions).  With base = true the densities
    // of the region come from the global base pass and S1's frozen density part is recorded.
    template <int T, int L>
    __device__ double region_energy(BatchArgs const& A, SView const& v, Team<T> const& tm, bool base, unsigned& err) {
      EAMTables const& Tb = A.T;
      int const nt = Tb.ntypes;
      unsigned e = 0;
      int const lane = tm.tid % L;
      // Per-row energies go to the fp scratch (overwritten by the next gradient), then a fixed-order sum.
      for (int r0 = tm.tid / L; r0 < v.nr; r0 += T / L) {
        double4 pi = v.cur[r0];
        int ti = A.multi ? v.type[r0] : 0;
        Acc16<L> rho_acc, pair_acc;
        bool mob = r0 < v.nm;
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        for (int k = b + lane, m = 0; k < c; k += L, ++m) {
          unsigned en = v.ent[k];
          int j = static_cast<int>(en & NB_IDX_MASK);
          double3 sh = A.S[en >> NB_SHIFT_BITS];
          double4 pj = v.cur[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < Tb.rc2) {
            double r = sqrt(r2);
            int tj = A.multi ? v.type[j] : 0;
            rho_acc.add(m, spl64<0>(Tb, Tb.phi[tj + nt * ti], r, e));
            if (mob) pair_acc.add(m, (j < v.nm ? 0.5 : 1.0) * spl64<0>(Tb, Tb.V[vix(nt, ti, tj)], r, e));
          }
        }
        double rho = rho_acc.total();
        double pair = pair_acc.total();
        if (lane == 0) {
          double rt;
          if (base) {
            rt = A.rho_base[v.gidx[r0]];
            if (!mob) v.rhof[r0] = rt - rho;
          } else {
            rt = mob ? rho : rho + v.rhof[r0];
          }
          v.fp[r0] = spl64<0>(Tb, Tb.F[ti], rt, e) + pair;
        }
      }
      __syncthreads();
      double vv[Team<T>::NQ][4] = {};
      for (int i = tm.tid, q = 0; i < v.nr; i += T, q = (q + 1) % Team<T>::NQ) vv[q][0] = __dadd_rn(vv[q][0], v.fp[i]);
      double o[1];
      tm.template sum<1>(o, vv);
      err |= tm.any_or(e);
      return o[0];
    }

    // cur[i] = x + c * d for the mobile atoms (d may be null).
    template <int T>
    __device__ void place(SView const& v, Team<T> const& tm, double const* x, double c, double const* d) {
      __syncthreads();
      for (int i = tm.tid; i < v.nm; i += T) {
        double px = x[3 * i], py = x[3 * i + 1], pz = x[3 * i + 2];
        if (d) {
          px += c * d[3 * i];
          py += c * d[3 * i + 1];
          pz += c * d[3 * i + 2];
        }
        v.cur[i] = make_double4(px, py, pz, 0.0);
      }
      __syncthreads();
    }

    // L-BFGS two-loop recursion on the workgroup (lbfgs_core.cpp:29-87).
    template <int T>
    struct Core {
      double* s;      // m vectors
      double* y;      // m vectors
      double* px;
      double* pg;
      double* q;
      double* rho;    // LDS, m
      double* alpha;  // LDS, m
      int m, k, n3;

      __device__ double* newton_step(Team<T> const& tm, double const* x, double const* g) {
        int prev = k > 0 ? (k - 1) % m : 0;
        if (k > 0) {
          double* sp = s + static_cast<long>(prev) * n3;
          double* yp = y + static_cast<long>(prev) * n3;
          tm.lin(sp, n3, x, 1.0, px, -1.0);
          tm.lin(yp, n3, g, 1.0, pg, -1.0);
          double sy = tm.dot(sp, yp, n3);
          if (tm.tid == 0) rho[prev] = 1.0 / fabs(sy);
        }
        tm.lin(px, n3, x, 1.0);
        tm.lin(pg, n3, g, 1.0);
        tm.lin(q, n3, g, 1.0);
        int incur = k <= m ? 0 : k - m;
        int bound = k <= m ? k : m;
        __syncthreads();
        for (int i = bound - 1; i >= 0; --i) {
          int j = (i + incur) % m;
          double a = rho[j] * tm.dot(s + static_cast<long>(j) * n3, q, n3);
          if (tm.tid == 0) alpha[j] = a;
          tm.lin(q, n3, q, 1.0, y + static_cast<long>(j) * n3, -a);
        }
        if (k > 0) {
          double* yp = y + static_cast<long>(prev) * n3;
          double scale = 1.0 / rho[prev] / tm.dot(yp, yp, n3);
          tm.lin(q, n3, q, scale);
        }
        __syncthreads();
        for (int i = 0; i < bound; ++i) {
          int j = (i + incur) % m;
          double beta = rho[j] * tm.dot(y + static_cast<long>(j) * n3, q, n3);
          tm.lin(q, n3, q, 1.0, s + static_cast<long>(j) * n3, alpha[j] - beta);
        }
        ++k;
        __syncthreads();
        return q;
      }
    };

    template <int T, int L, bool MIX>
    __device__ void run_search(BatchArgs const& A, int sid, Team<T> const& tm, double* lds_core) {
      SearchHdr const h = A.hdr[sid];
      DimOpt const& o = A.o;
      SView v;
      v.nm = h.nm;
      v.nr = h.nr;
      v.n3 = 3 * h.nm;
      v.cur = A.cur + h.reg_off;
      v.base = A.base + h.reg_off;
      v.type = A.rtype + h.reg_off;
      v.gidx = A.rgidx + h.reg_off;
      v.rhof = A.rhof + h.reg_off;
      v.fp = A.fp + h.reg_off;
      v.rowptr = A.rowptr + h.row_off;
      v.ent = A.ent;
      int const n3 = v.n3;
      double* W = A.ws + h.ws_off;
      auto vecp = [&](int k) { return W + static_cast<long>(k) * n3; };
      double* x = vecp(0);
      double* ax = vecp(1);
      double* eff = vecp(2);
      double* g0 = vecp(3);
      double* g1 = vecp(4);
      double* g1p = vecp(5);
      double* dg = vecp(6);
      double* dgp = vecp(7);
      double* axp = vecp(8);
      double* xb = vecp(9);
      int const t0 = 10, r0v = 13 + 2 * o.n;
      Core<T> tc{vecp(t0), vecp(t0 + o.n), vecp(t0 + 2 * o.n), vecp(t0 + 2 * o.n + 1), vecp(t0 + 2 * o.n + 2),
                 lds_core, lds_core + 32, o.n, 0, n3};
      Core<T> rc{vecp(r0v), vecp(r0v + o.rn), vecp(r0v + 2 * o.rn), vecp(r0v + 2 * o.rn + 1), vecp(r0v + 2 * o.rn + 2),
                 lds_core + 64, lds_core + 96, o.rn, 0, n3};
      double const dr = o.delta_r;
      unsigned err = 0;
      long grads = 0;
      int rot_iters = 0;

      // Base state: region at base positions, frozen density parts, base local energy.
      for (int i = tm.tid; i < v.nr; i += T) v.cur[i] = v.base[i];
      __syncthreads();
      double e_base = region_energy<T, L>(A, v, tm, true, err);
      for (int i = tm.tid; i < n3; i += T) {
        int a = i / 3, c = i % 3;
        double4 b = v.base[a];
        xb[i] = c == 0 ? b.x : (c == 1 ? b.y : b.z);
        x[i] = A.start_x[h.v3_off + i];
        ax[i] = A.start_ax[h.v3_off + i];
      }
      __syncthreads();

      int ret = 2;   // iter_max
      int steps = o.max_steps;
      double curv = 0;
      double maxd = 0;

      // Rows built at the base positions with cut-off r_c + skin stay complete while every pair that is
      // missing from them stays beyond r_c: a mobile-frozen pair needs d_i < skin and a mobile-mobile
      // pair d_i + d_j < skin, i.e. d1 + d2 < skin for the two largest mobile displacements d1 >= d2.
      auto displacement_ok = [&]() {
        double m1 = 0, m2 = 0;
        for (int a = tm.tid; a < v.nm; a += T) {
          double dx = x[3 * a] - xb[3 * a], dy = x[3 * a + 1] - xb[3 * a + 1], dz = x[3 * a + 2] - xb[3 * a + 2];
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (!(d <= m1)) {   // also takes NaN, which then fails the test below
            m2 = m1;
            m1 = d;
          } else if (d > m2) {
            m2 = d;
          }
        }
        for (int off = 32; off > 0; off >>= 1) {
          double o1 = __shfl_xor(m1, off, 64), o2 = __shfl_xor(m2, off, 64);
          double hi = fmax(m1, o1), lo = fmax(fmin(m1, o1), fmax(m2, o2));
          m1 = hi;
          m2 = lo;
        }
        if ((tm.tid & 63) == 0) {
          tm.red[4 * (tm.tid >> 6)] = m1;
          tm.red[4 * (tm.tid >> 6) + 1] = m2;
        }
        __syncthreads();
        double a1 = 0, a2 = 0;
        for (int w = 0; w < T / 64; ++w) {
          double b1 = tm.red[4 * w], b2 = tm.red[4 * w + 1];
          double hi = fmax(a1, b1), lo = fmax(fmin(a1, b1), fmax(a2, b2));
          a1 = hi;
          a2 = lo;
        }
        __syncthreads();
        maxd = fmax(maxd, a1);
        return a1 + a2 < h.skin;
      };

      // Rotor::eff_gradient (rotor.cpp:49-164); returns false on a spline error.
      auto eff_gradient = [&](double& mag_g) -> bool {
        rc.k = 0;
        place(v, tm, x, 0.0, nullptr);
        err |= region_gradient<T, L, MIX>(A, v, tm, g0);
        place(v, tm, x, dr, ax);
        err |= region_gradient<T, L, MIX>(A, v, tm, g1);
        grads += 2;
        if (err) return false;
        for (int i = 0;; ++i) {
          tm.lin(dg, n3, g1, 1.0, g0, -1.0);
          double t = tm.dot(dg, ax, n3);
          tm.lin(dgp, n3, dg, 1.0, ax, -t);
          double* theta = rc.newton_step(tm, ax, dgp);
          double ot = tm.dot(theta, ax, n3);
          tm.lin(theta, n3, theta, 1.0, ax, -ot);
          double nn = tm.dot(theta, theta, n3);
          tm.lin(theta, n3, theta, 1.0 / sqrt(nn));
          double b1r, c0r;
          tm.dot2(dg, theta, dg, ax, n3, b1r, c0r);
          double b_1 = b1r / dr, c_x0 = c0r / dr;
          double theta_1 = -0.5 * atan(b_1 / fabs(c_x0));
          ++rot_iters;
          if (fabs(theta_1) < o.theta_tol || (i >= o.iter_max_rot && c_x0 < 0) || i > 10 * o.iter_max_rot) {
            curv = c_x0;
            break;
          }
          double ct = cos(theta_1), st = sin(theta_1);
          tm.lin(axp, n3, ax, ct, theta, st);
          place(v, tm, x, dr, axp);
          err |= region_gradient<T, L, MIX>(A, v, tm, g1p);
          ++grads;
          if (err) return false;
          tm.lin(dg, n3, g1p, 1.0, g0, -1.0);
          double c_x1 = tm.dot(dg, axp, n3) / dr;
          double a_1 = (c_x0 - c_x1 + b_1 * sin(2 * theta_1)) / (1 - cos(2 * theta_1));
          double theta_min = 0.5 * atan(b_1 / a_1);
          if (a_1 * cos(2 * theta_min) - a_1 + b_1 * sin(2 * theta_min) > 0) theta_min += M_PI / 2;
          double cm = cos(theta_min), sm = sin(theta_min);
          tm.lin(ax, n3, ax, cm, theta, sm);
          double w1 = sin(theta_1 - theta_min) / sin(theta_1);
          double w2 = sin(theta_min) / sin(theta_1);
          double w3 = 1 - cos(theta_min) - sin(theta_min) * tan(0.5 * theta_1);
          tm.lin(g1, n3, g1, w1, g1p, w2, g0, w3);
          if (fabs(theta_min) < o.theta_tol) {
            tm.lin(dg, n3, g1, 1.0, g0, -1.0);
            curv = tm.dot(dg, ax, n3) / dr;
            break;
          }
        }
        double p = tm.dot(g0, ax, n3);
        if (!o.relax_in_convex && curv > 0) {
          tm.lin(eff, n3, ax, -p);
        } else {
          tm.lin(eff, n3, g0, 1.0, ax, -2.0 * p);
        }
        mag_g = tm.dot(eff, eff, n3);
        return true;
      };

      // Dimer::find_sp (dimer.cpp:28-171) without a history of known saddles.
      double mag_g = 0;
      if (!displacement_ok()) {
        ret = EXIT_REGION;
        steps = 0;
      } else if (!eff_gradient(mag_g)) {
        ret = EXIT_OOB;
        steps = 0;
      } else {
        double trust = o.min_trust;
        int convex_count = 0;
        for (int i = 0; i < o.max_steps; ++i) {
          if (mag_g < o.f2norm * o.f2norm) {
            ret = 0;
            steps = i;
            break;
          } else if (convex_count >= o.convex_max) {
            ret = 1;
            steps = i;
            break;
          }
          double* q = tc.newton_step(tm, x, eff);
          double nrm = sqrt(tm.dot(q, q, n3));
          double scale = fmin(1.0, trust / nrm);
          tm.lin(q, n3, q, scale);
          tm.lin(x, n3, x, 1.0, q, -1.0);
          __syncthreads();
          if (!displacement_ok()) {
            ret = EXIT_REGION;
            steps = i;
            break;
          }
          if (!eff_gradient(mag_g)) {
            ret = EXIT_OOB;
            steps = i;
            break;
          }
          convex_count = curv > 0 ? convex_count + 1 : 0;
          double proj = tm.dot(eff, q, n3);
          if (proj < -o.proj_tol) {
            trust = fmax(o.min_trust, o.shrink_trust * trust);
          } else if (proj > o.proj_tol) {
            trust = fmin(o.max_trust, o.grow_trust * trust);
          }
        }
      }

      // Barrier in FP64 at the returned positions.
      double barrier = 0;
      if (ret != EXIT_REGION && ret != EXIT_OOB) {
        place(v, tm, x, 0.0, nullptr);
        unsigned e2 = 0;
        double e_sp = region_energy<T, L>(A, v, tm, false, e2);
        barrier = e_sp - e_base;
        if (e2) ret = EXIT_OOB;
      }
      for (int i = tm.tid; i < n3; i += T) {
        A.out_x[h.v3_off + i] = x[i];
        A.out_ax[h.v3_off + i] = ax[i];
      }
      if (tm.tid == 0) {
        SearchOut r;
        r.exit_code = ret;
        r.steps = steps;
        r.rot_iters = rot_iters;
        r.pad = 0;
        r.grads = grads;
        r.curv = curv;
        r.barrier = barrier;
        r.e_base = e_base;
        r.max_disp = maxd;
        A.out[sid] = r;
      }
      __syncthreads();
    }

    template <int T, int L, bool MIX>
    __global__ __launch_bounds__(T) void k_search(BatchArgs A) {
      __shared__ double red[4 * T];
      __shared__ double lds_core[128];
      __shared__ int s_id;
      Team<T> tm{red, static_cast<int>(threadIdx.x)};
      while (true) {
        if (threadIdx.x == 0) s_id = static_cast<int>(atomicAdd(A.queue, 1u));
        __syncthreads();
        int sid = s_id;
        __syncthreads();
        if (sid >= A.nsearch) break;
        run_search<T, L, MIX>(A, sid, tm, lds_core);
      }
    }

  }  // namespace
}  // namespace flyg

// =====================================================================================================
// Host side: region extraction, packing, launch, re-queue, autotuning and the C-ABI
// =====================================================================================================

namespace flyg {
  void eam_density(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* d_types, double* d_rho);

  namespace {

    // Periodic cell grid over the base positions (orthorhombic box), cells of at least `rmin`.
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
          nc[k] = std::max(1, static_cast<int>(std::floor(L[k] / rmin)));
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
        for (int oz = -reach[2]; oz <= reach[2]; ++oz)
          for (int oy = -reach[1]; oy <= reach[1]; ++oy)
            for (int ox = -reach[0]; ox <= reach[0]; ++ox) {
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
    };

    struct Region {
      int nm = 0, nr = 0;
      std::vector<int> gidx;
      std::vector<int> rowlen;      // nr
      std::vector<unsigned> ent;    // packed (shift, local)
      std::string err;
    };

    struct Pending {
      int64_t id;
      int launches = 0;
      int64_t centre;
      double radius;
      double skin;
      std::vector<int> mobile;      // global indices (sorted)
      std::vector<double> x, ax;    // 3*nm start state (continuous with the base positions)
    };

  }  // namespace
}  // namespace flyg

struct flyg_batch_s {
  flyg_ctx_s* ctx = nullptr;
  flyg_eam_s* eam = nullptr;
  flyg_dimer_options opt;
  flyg_rotor_options ropt;
  flyg_batch_options bopt;
  double basis[9];
  int periodic[3];
  flyg::BoxDesc box;
  long n = 0;
  std::vector<double> base;
  std::vector<uint32_t> types;
  flyg::HostGrid grid;
  flyg::DBuf<double> rho_base;
  std::vector<flyg::Pending> pending;
  struct Result {
    int done = 0;
    flyg_search_stats st{};
    std::vector<int> mobile;
    std::vector<double> x, ax;
    int nm = 0, nr = 0;
    double max_disp = 0;
  };
  std::vector<Result> results;
  int tuned_T = 0, tuned_L = 0;
  // device staging (grown on demand)
  flyg::DBuf<flyg::SearchHdr> d_hdr;
  flyg::DBuf<double4> d_cur, d_base;
  flyg::DBuf<uint8_t> d_type;
  flyg::DBuf<int> d_gidx, d_rowptr;
  flyg::DBuf<double> d_rhof, d_fp, d_ws, d_sx, d_sax, d_ox, d_oax;
  flyg::DBuf<unsigned> d_ent, d_queue;
  flyg::DBuf<flyg::SearchOut> d_out;
  double last_kernel_ms = 0;
};

namespace flyg {
  namespace {

    double min_image(double d, double L) { return d - L * std::round(d / L); }

    void mobile_set(flyg_batch_s const& B, int64_t centre, double R, std::vector<int>& out) {
      out.clear();
      double const* pc = &B.base[3 * centre];
      B.grid.query(pc, R, [&](int j, int, double) { out.push_back(j); });
      std::sort(out.begin(), out.end());
      out.erase(std::unique(out.begin(), out.end()), out.end());
    }

    Region build_region(flyg_batch_s const& B, Pending const& P, double rL, std::vector<int>& local_of) {
      Region g;
      int const nm = static_cast<int>(P.mobile.size());
      g.nm = nm;
      g.gidx = P.mobile;
      for (int i = 0; i < nm; ++i) local_of[P.mobile[i]] = i;
      std::vector<int> s1;
      for (int i = 0; i < nm; ++i)
        B.grid.query(&B.base[3 * P.mobile[i]], rL, [&](int j, int, double r2) {
          if (r2 < rL * rL && local_of[j] == -1) {
            local_of[j] = -2;   // collected into S1
            s1.push_back(j);
          }
        });
      std::sort(s1.begin(), s1.end());
      for (std::size_t k = 0; k < s1.size(); ++k) local_of[s1[k]] = nm + static_cast<int>(k);
      g.gidx.insert(g.gidx.end(), s1.begin(), s1.end());
      g.nr = static_cast<int>(g.gidx.size());
      g.rowlen.assign(g.nr, 0);
      std::vector<unsigned> row;
      for (int r = 0; r < g.nr; ++r) {
        int gi = g.gidx[r];
        bool mob = r < nm;
        row.clear();
        B.grid.query(&B.base[3 * gi], rL, [&](int j, int sc, double r2) {
          if (!(r2 < rL * rL)) return;
          if (j == gi && sc == 13) return;
          int lj = local_of[j];
          if (mob) {
            if (lj < 0) return;   // cannot happen: S1 holds every frozen atom within r_L of a mobile atom
            row.push_back((static_cast<unsigned>(sc) << NB_SHIFT_BITS) | static_cast<unsigned>(lj));
          } else if (lj >= 0 && lj < nm) {
            row.push_back((static_cast<unsigned>(sc) << NB_SHIFT_BITS) | static_cast<unsigned>(lj));
          }
        });
        std::sort(row.begin(), row.end());
        g.rowlen[r] = static_cast<int>(row.size());
        g.ent.insert(g.ent.end(), row.begin(), row.end());
      }
      for (int j : g.gidx) local_of[j] = -1;
      for (int j : s1) local_of[j] = -1;
      return g;
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

    using KernelFn = void (*)(BatchArgs);

    template <int T, int L>
    KernelFn pick(bool mix) {
      return mix ? reinterpret_cast<KernelFn>(&k_search<T, L, true>) : reinterpret_cast<KernelFn>(&k_search<T, L, false>);
    }

    KernelFn kernel_for(int T, int L, bool mix) {
      if (T == 256 && L == 4) return pick<256, 4>(mix);
      if (T == 256 && L == 8) return pick<256, 8>(mix);
      if (T == 256 && L == 16) return pick<256, 16>(mix);
      if (T == 512 && L == 4) return pick<512, 4>(mix);
      if (T == 512 && L == 8) return pick<512, 8>(mix);
      if (T == 512 && L == 16) return pick<512, 16>(mix);
      if (T == 128 && L == 8) return pick<128, 8>(mix);
      throw Error(FLYG_INVALID_ARG, "unsupported (block_threads, lanes_per_atom); use 128/256/512 and 4/8/16");
    }

    DimOpt dim_opt(flyg_dimer_options const& o, flyg_rotor_options const& r) {
      DimOpt d{};
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
      return d;
    }

    // Runs one launch over `work`; fills outs (one per work item).
    void launch(flyg_batch_s& B, std::vector<Pending*> const& work, int T, int L, DimOpt const& dopt,
                std::vector<SearchOut>& outs, std::vector<Region>& regions) {
      flyg_ctx_s& ctx = *B.ctx;
      long const ns = static_cast<long>(work.size());
      regions.assign(ns, Region{});
      std::vector<std::vector<int>> scratch(16, std::vector<int>());
      parallel_for(ns, [&](long k, unsigned t) {
        if (scratch[t].empty()) scratch[t].assign(B.n, -1);
        Pending const& P = *work[k];
        regions[k] = build_region(B, P, B.eam->tab.rc + P.skin, scratch[t]);
      });
      std::vector<SearchHdr> hdr(ns);
      long nreg = 0, nrow = 0, nent = 0, n3 = 0, nws = 0;
      for (long k = 0; k < ns; ++k) {
        Region const& g = regions[k];
        require(g.nm > 0, "a search has no mobile atoms");
        require(g.nr < (1 << NB_SHIFT_BITS), "search region too large");
        hdr[k].nm = g.nm;
        hdr[k].nr = g.nr;
        hdr[k].reg_off = nreg;
        hdr[k].row_off = nrow;
        hdr[k].v3_off = n3;
        hdr[k].ws_off = nws;
        hdr[k].skin = work[k]->skin;
        nreg += g.nr;
        nrow += g.nr + 1;
        nent += static_cast<long>(g.ent.size());
        n3 += 3L * g.nm;
        nws += static_cast<long>(nvec(dopt.n, dopt.rn)) * 3L * g.nm;
      }
      std::vector<double4> hbase(nreg);
      std::vector<uint8_t> htype(nreg);
      std::vector<int> hgidx(nreg), hrow(nrow);
      std::vector<unsigned> hent(nent);
      std::vector<double> hsx(n3), hsax(n3);
      long e0 = 0;
      for (long k = 0; k < ns; ++k) {
        Region const& g = regions[k];
        Pending const& P = *work[k];
        long ro = hdr[k].reg_off;
        for (int r = 0; r < g.nr; ++r) {
          int gi = g.gidx[r];
          hbase[ro + r] = make_double4(B.base[3 * gi], B.base[3 * gi + 1], B.base[3 * gi + 2], 0.0);
          htype[ro + r] = static_cast<uint8_t>(B.types[gi]);
          hgidx[ro + r] = gi;
        }
        long rp = hdr[k].row_off;
        for (int r = 0; r < g.nr; ++r) {
          hrow[rp + r] = static_cast<int>(e0);
          e0 += g.rowlen[r];
        }
        hrow[rp + g.nr] = static_cast<int>(e0);
        std::copy(g.ent.begin(), g.ent.end(), hent.begin() + (hrow[rp]));
        std::copy(P.x.begin(), P.x.end(), hsx.begin() + hdr[k].v3_off);
        std::copy(P.ax.begin(), P.ax.end(), hsax.begin() + hdr[k].v3_off);
      }
      require(nent < (1l << 31), "too many neighbour entries in one launch");
      hipStream_t st = ctx.stream;
      auto up = [&](auto& buf, auto const& v) {
        buf.reserve(v.size() + 1);
        if (!v.empty()) HIPCHK(hipMemcpyAsync(buf.get(), v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice, st));
      };
      up(B.d_hdr, hdr);
      up(B.d_base, hbase);
      up(B.d_type, htype);
      up(B.d_gidx, hgidx);
      up(B.d_rowptr, hrow);
      up(B.d_ent, hent);
      up(B.d_sx, hsx);
      up(B.d_sax, hsax);
      B.d_cur.reserve(nreg + 1);
      B.d_rhof.reserve(nreg + 1);
      B.d_fp.reserve(nreg + 1);
      B.d_ws.reserve(nws + 1);
      B.d_ox.reserve(n3 + 1);
      B.d_oax.reserve(n3 + 1);
      B.d_out.reserve(ns + 1);
      B.d_queue.reserve(1);
      HIPCHK(hipMemsetAsync(B.d_queue.get(), 0, sizeof(unsigned), st));
      HIPCHK(hipMemsetAsync(B.d_ws.get(), 0, (nws + 1) * sizeof(double), st));

      BatchArgs A{};
      A.T = B.eam->tab;
      shift_table(B.box, A.S);
      A.hdr = B.d_hdr.get();
      A.nsearch = static_cast<int>(ns);
      A.cur = B.d_cur.get();
      A.base = B.d_base.get();
      A.rtype = B.d_type.get();
      A.rgidx = B.d_gidx.get();
      A.rho_base = B.rho_base.get();
      A.rhof = B.d_rhof.get();
      A.fp = B.d_fp.get();
      A.rowptr = B.d_rowptr.get();
      A.ent = B.d_ent.get();
      A.ws = B.d_ws.get();
      A.start_x = B.d_sx.get();
      A.start_ax = B.d_sax.get();
      A.out_x = B.d_ox.get();
      A.out_ax = B.d_oax.get();
      A.out = B.d_out.get();
      A.queue = B.d_queue.get();
      A.o = dopt;
      A.multi = B.eam->tab.ntypes > 1;

      int slots = B.bopt.max_concurrent > 0 ? B.bopt.max_concurrent : 2 * ctx.cus;
      unsigned grid = static_cast<unsigned>(std::max<long>(1, std::min<long>(slots, ns)));
      KernelFn fn = kernel_for(T, L, B.bopt.precision == FLYG_MIXED);
      hipEvent_t e_a, e_b;
      HIPCHK(hipEventCreate(&e_a));
      HIPCHK(hipEventCreate(&e_b));
      HIPCHK(hipEventRecord(e_a, st));
      hipLaunchKernelGGL(fn, dim3(grid), dim3(T), 0, st, A);
      HIPCHK(hipGetLastError());
      HIPCHK(hipEventRecord(e_b, st));
      outs.resize(ns);
      HIPCHK(hipMemcpyAsync(outs.data(), B.d_out.get(), ns * sizeof(SearchOut), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, e_a, e_b));
      B.last_kernel_ms = ms;
      (void)hipEventDestroy(e_a);
      (void)hipEventDestroy(e_b);
    }

    void collect(flyg_batch_s& B, std::vector<Pending*> const& work, std::vector<SearchOut> const& outs,
                 std::vector<Region> const& regions, std::vector<Pending>& requeue, double skin_cap) {
      long ns = static_cast<long>(work.size());
      long n3 = 0;
      for (long k = 0; k < ns; ++k) n3 += 3L * regions[k].nm;
      std::vector<double> ox(n3), oax(n3);
      hipStream_t st = B.ctx->stream;
      if (n3) {
        HIPCHK(hipMemcpyAsync(ox.data(), B.d_ox.get(), n3 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(oax.data(), B.d_oax.get(), n3 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
      }
      long off = 0;
      for (long k = 0; k < ns; ++k) {
        Pending& P = *work[k];
        SearchOut const& o = outs[k];
        int nm = regions[k].nm;
        ++P.launches;
        if (o.exit_code == EXIT_REGION && P.skin < skin_cap) {
          Pending q = P;
          q.skin = std::min(skin_cap, std::max(P.skin * 1.5, 2.0 * o.max_disp + 0.5));
          requeue.push_back(std::move(q));
          off += 3L * nm;
          continue;
        }
        auto& R = B.results[P.id];
        R.done = 1;
        R.st.exit_code = o.exit_code;
        R.st.steps = o.steps;
        R.st.grads = o.grads;
        R.st.rebuilds = 0;
        R.st.curv = o.curv;
        R.st.energy = o.barrier;
        R.st.max_disp = o.max_disp;
        R.st.launches = P.launches;
        R.mobile = P.mobile;
        R.x.assign(ox.begin() + off, ox.begin() + off + 3L * nm);
        R.ax.assign(oax.begin() + off, oax.begin() + off + 3L * nm);
        R.nm = nm;
        R.nr = regions[k].nr;
        R.max_disp = o.max_disp;
        off += 3L * nm;
      }
    }

    void autotune(flyg_batch_s& B) {
      if (B.bopt.block_threads > 0 && B.bopt.lanes_per_atom > 0) {
        B.tuned_T = B.bopt.block_threads;
        B.tuned_L = B.bopt.lanes_per_atom;
        return;
      }
      if (B.tuned_T) return;
      // Time a short prefix (20 dimer steps) of up to 64 queued searches with each candidate shape.
      std::vector<Pending*> work;
      for (auto& P : B.pending) {
        if (work.size() >= 64) break;
        work.push_back(&P);
      }
      if (work.empty()) return;
      DimOpt d = dim_opt(B.opt, B.ropt);
      d.max_steps = std::min(d.max_steps, 20);
      int const cand[][2] = {{256, 4}, {256, 8}, {256, 16}, {512, 8}, {512, 16}, {128, 8}};
      double best = 1e300;
      std::vector<SearchOut> outs;
      std::vector<Region> regions;
      for (auto const& c : cand) {
        launch(B, work, c[0], c[1], d, outs, regions);
        if (B.last_kernel_ms < best) {
          best = B.last_kernel_ms;
          B.tuned_T = c[0];
          B.tuned_L = c[1];
        }
      }
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
    if (opt) {
      b->opt = *opt;
    } else {
      flyg_dimer_default_options(&b->opt);
    }
    if (ropt) {
      b->ropt = *ropt;
    } else {
      flyg_rotor_default_options(&b->ropt);
    }
    if (bopt) {
      b->bopt = *bopt;
    } else {
      flyg_batch_default_options(&b->bopt);
    }
    if (b->bopt.list_skin <= 0) b->bopt.list_skin = 2.0;
    require(b->opt.n > 0 && b->opt.n <= 32 && b->ropt.n > 0 && b->ropt.n <= 32,
            "batched searches support L-BFGS histories up to 32");
    require(b->bopt.precision == FLYG_MIXED || b->bopt.precision == FLYG_FP64, "bad precision");
    std::memcpy(b->basis, basis, sizeof(b->basis));
    std::memcpy(b->periodic, periodic, sizeof(b->periodic));
    b->box = make_box(basis, periodic);
    require(b->box.ortho && periodic[0] && periodic[1] && periodic[2],
            "batched searches need a fully periodic orthorhombic box");
    b->n = n;
    b->base.resize(3 * n);
    b->types.assign(types, types + n);
    for (long i = 0; i < n; ++i) {
      require(types[i] < static_cast<uint32_t>(eam->tab.ntypes), "TypeID out of range");
      for (int k = 0; k < 3; ++k) {
        // Strictly inside [0, L): the canonical image of a coordinate just below 0 rounds to L, and the
        // cell grid and the image shifts must agree on which side of the box an atom is.
        double x = min_pos[3 * i + k];
        double w = x - b->box.ext[k] * std::floor(x * b->box.inv_ext[k]);
        if (w >= b->box.ext[k]) w -= b->box.ext[k];
        if (w < 0) w = 0;
        b->base[3 * i + k] = w;
      }
    }
    double rl = eam->tab.rc + b->bopt.list_skin;
    for (int k = 0; k < 3; ++k) require(b->box.ext[k] > 2.0 * rl, "box too small for the search region cut-off");
    b->grid.build(b->base.data(), n, b->box.ext, rl);
    // Global FP64 base densities.
    auto nl = make_list(ctx, basis, periodic, eam->tab.rc);
    DBuf<double> dpos;
    dpos.reserve(3 * n);
    DBuf<uint32_t> dt;
    dt.reserve(n);
    HIPCHK(hipMemcpyAsync(dpos.get(), b->base.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dt.get(), types, n * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
    list_rebuild(*nl, dpos.get(), n);
    b->rho_base.reserve(n);
    eam_density(*eam, *nl, dt.get(), b->rho_base.get());
    unsigned fl = ctx->take_flags();
    if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
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
  return bguard(b->ctx, [&] {
    require(count >= 0 && centre && radius && start_pos && start_axis, "null argument");
    long const n = b->n;
    for (int k = 0; k < count; ++k) {
      require(centre[k] >= 0 && centre[k] < n, "centre out of range");
      require(radius[k] > 0 && radius[k] < 0.5 * std::fmin(b->box.ext[0], std::fmin(b->box.ext[1], b->box.ext[2])) - 1.0,
              "active radius must be positive and below half the box");
      Pending P;
      P.id = static_cast<int64_t>(b->results.size());
      P.centre = centre[k];
      P.radius = radius[k];
      P.skin = b->bopt.list_skin;
      mobile_set(*b, centre[k], radius[k], P.mobile);
      int nm = static_cast<int>(P.mobile.size());
      P.x.resize(3 * nm);
      P.ax.resize(3 * nm);
      double const* sp = start_pos + 3 * n * static_cast<long>(k);
      double const* sa = start_axis + 3 * n * static_cast<long>(k);
      for (int i = 0; i < nm; ++i) {
        int g = P.mobile[i];
        for (int c = 0; c < 3; ++c) {
          double bse = b->base[3 * g + c];
          P.x[3 * i + c] = bse + min_image(sp[3 * g + c] - bse, b->box.ext[c]);
          P.ax[3 * i + c] = sa[3 * g + c];
        }
      }
      b->results.emplace_back();
      if (ids) ids[k] = P.id;
      b->pending.push_back(std::move(P));
    }
  });
}

int flyg_batch_run(flyg_batch b) {
  return bguard(b->ctx, [&] {
    if (b->pending.empty()) return;
    autotune(*b);
    DimOpt d = dim_opt(b->opt, b->ropt);
    double cap = 0.5 * std::fmin(b->box.ext[0], std::fmin(b->box.ext[1], b->box.ext[2])) - b->eam->tab.rc;
    std::vector<Pending> todo = std::move(b->pending);
    b->pending.clear();
    for (int round = 0; round < 6 && !todo.empty(); ++round) {
      std::vector<Pending*> work;
      for (auto& P : todo) work.push_back(&P);
      std::vector<SearchOut> outs;
      std::vector<Region> regions;
      launch(*b, work, b->tuned_T, b->tuned_L, d, outs, regions);
      std::vector<Pending> requeue;
      collect(*b, work, outs, regions, requeue, cap);
      todo = std::move(requeue);
    }
    for (auto& P : todo) {   // still outside the largest region the box allows
      auto& R = b->results[P.id];
      R.done = 1;
      R.st.exit_code = EXIT_REGION;
      R.st.launches = P.launches;
      R.mobile = P.mobile;
    }
  });
}

int flyg_batch_result(flyg_batch b, int64_t id, flyg_search_stats* stats, double* pos, double* axis) {
  return bguard(b->ctx, [&] {
    require(id >= 0 && id < static_cast<int64_t>(b->results.size()), "unknown search id");
    auto const& R = b->results[id];
    require(R.done, "search has not run yet");
    if (stats) *stats = R.st;
    if (pos) {
      std::memcpy(pos, b->base.data(), 3 * b->n * sizeof(double));
      for (std::size_t i = 0; i < R.mobile.size() && 3 * i + 2 < R.x.size(); ++i)
        for (int c = 0; c < 3; ++c) pos[3 * R.mobile[i] + c] = R.x[3 * i + c];
    }
    if (axis) {
      std::memset(axis, 0, 3 * b->n * sizeof(double));
      for (std::size_t i = 0; i < R.mobile.size() && 3 * i + 2 < R.ax.size(); ++i)
        for (int c = 0; c < 3; ++c) axis[3 * R.mobile[i] + c] = R.ax[3 * i + c];
    }
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

}  // extern "C"
