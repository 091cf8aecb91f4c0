// Batched saddle searches on gfx950: the device-side counterpart of Master::find_batch
// (find.cpp:684-755: one OpenMP task per search running Master::find_sp -> Dimer::find_sp with one
// thread, find.cpp:983-1009).
//
// One persistent kernel; each workgroup pulls searches from a queue (one atomic per search) and runs the
// whole dimer search (Rotor::eff_gradient rotor.cpp:49-164, Dimer::find_sp dimer.cpp:28-171, both L-BFGS
// cores lbfgs_core.cpp:29-87) with workgroup barriers only.  A search works on its active region:
//   M  = mobile atoms (minimum-image distance to the centre <= radius, the rest of the system is Frozen),
//   S1 = frozen atoms within r_cut + skin_f of a mobile atom (their density, hence F', can change).
// Neighbour rows are built once from the base positions with Verlet skins: mobile rows
// hold every region neighbour, S1 rows only the mobile ones, and the frozen part of an S1 density is the
// base density minus the mobile part.  Mobile-frozen pairs are listed within r_cut + skin_f and
// mobile-mobile pairs within r_cut + skin_m, so the rows stay complete while d1 < skin_f and d1 + d2 <
// skin_m for the two largest mobile displacements from the base; a search that breaks this is re-run
// with larger skins.  Every 3N vector of
// the reference is zero on frozen atoms, so the device keeps only the 3*|M| mobile components.
// Barrier = E_loc(saddle) - E_loc(base) in FP64, where E_loc collects every term of the EAM energy that
// depends on mobile positions (sum over M u S1 of F(rho) plus the pairs with a mobile member).
//
// This header holds the device side (search kernel k_search and its helpers).  Each of the
// search_kernel_<T>_<precision>.hip translation units instantiates the variants of one workgroup size
// and precision (both lane counts and species modes), so the variants compile in parallel; batch.hip holds the host side.
#pragma once

#include <cmath>
#include <type_traits>

#include "plain.hpp"

namespace flyg {

  namespace sk {

    constexpr int EXIT_REGION = 5;
    constexpr int EXIT_OOB = 6;
    // Workspace vectors per search (each 3*nm doubles): 10 named vectors, then the translation core
    // (m s-vectors, m y-vectors, x_prev, g_prev, q), the rotation core (the same with its own m), and
    // the positions at the last active-list refresh.
    __host__ __device__ inline int nvec(int m, int rm) { return 17 + 2 * (m + rm); }

    // Saved dimer state of a search that left its region (x, ax, the translation core s, y, x_prev,
    // g_prev, q, then its m rho values): what a relaunch in a grown region continues from.
    __host__ __device__ inline long resume_len(int n3, int m) { return static_cast<long>(2 * m + 5) * n3 + m; }

    // Active lists: the part of each candidate row within r_c + ACT_SKIN_F (pairs with a frozen atom)
    // or r_c + ACT_SKIN_M (mobile pairs) at the last refresh, rebuilt on the device when the mobile
    // atoms have moved too far since (a Verlet list inside the candidate rows).
    constexpr double ACT_SKIN_F = 0.5;
    constexpr double ACT_SKIN_M = 1.0;

    struct SearchHdr {
      int nm, nr;          // mobile atoms, region atoms (M u S1)
      long reg_off;        // first region atom in the packed region arrays
      long row_off;        // first row pointer (nr + 1 entries)
      long v3_off;         // first element of this search's 3*nm start/out vectors
      long ws_off;         // first element of this search's workspace (nvec * 3*nm)
      double skin_f;       // Verlet skin of mobile-frozen pairs (and of S1)
      double skin_m;       // Verlet skin of mobile-mobile pairs (>= skin_f)
      // The search's system, resident on the device for as long as any of its searches runs.
      double4 const* sbase;   // canonical base positions, by global atom index
      double const* srho;     // FP64 base densities
      uint8_t const* stype;   // type ids
      double3 const* sS;      // the 27 image shifts of the box, by shift code
      long natoms;            // atoms in the system (centroids run over all of them)
      int centre_local;       // mobile index of the search's centre atom
      // Known saddles of the history test (Dimer::find_sp's hist_sp, dimer.cpp:102-131): n_hist vectors
      // sp_k - min over the mobile atoms (3*nm each, from hist_off) and their full-system norms (hn_off).
      int n_hist;
      long hist_off, hn_off;
      double cos_tol;         // cos(theta_tol): a known saddle closer in angle than theta_tol is a collision
      // Resumption after a region exit: the host grew the skins and the search continues from the state
      // it saved (SearchOut::saved) at rs_in_off of BatchArgs::rs_in, re-entering its step loop at step
      // r_step right after that step's translation, with the counters and trust it had.
      int resume;
      int r_step, r_convex, r_k, r_rot_iters, r_refreshes;
      long r_grads;
      double r_trust, r_maxd, r_maxp;
      long rs_in_off;
    };

    struct SearchOut {
      int exit_code, steps, rot_iters, refreshes;   // refreshes: active-list rebuilds
      long grads;
      double curv, barrier, e_base, max_disp;
      double max_pair;   // largest d1 + d2 seen
      int mech_code, min_iters;      // mechanism check (0 = a min->sp->min pathway), minimiser iterations
      double mech_barrier, mech_delta;   // E(sp) - E(rev), E(fwd) - E(rev)
      double drift[3];                   // centroid drift of rev from the initial minimum (all atoms)
      long ticks;        // device wall-clock ticks the workgroup spent on this search
      long grad_ticks;   // ... of which inside region gradients
      // Region exit inside the step loop: 1 + the slot of BatchArgs::rs_out holding the saved state
      // (0: not saved, the search restarts), with the translation core's pair count, the convex-step
      // count and the trust radius at the exit.
      int saved, lbfgs_k, convex;
      double trust;
    };

    struct DimOpt {
      int n, max_steps, convex_max, rn, iter_max_rot, relax_in_convex, hist_check_freq, use_history;
      double f2norm, proj_tol, max_trust, min_trust, grow_trust, shrink_trust, delta_r, theta_tol;
      // Mechanisms (Master::do_adj_min / saddle_2_mech): 0 = off.
      int mech, min_iter_max;
      double nudge_frac, basin_tol, stationary_tol;
      double min_f2norm, min_proj_tol, min_max_trust, min_min_trust, min_grow_trust, min_shrink_trust;
    };

    struct BatchArgs {
      EAMTables T;
      SearchHdr const* hdr;
      int nsearch;
      double4* cur;              // region positions (base for S1, current evaluation point for M)
      double4* base;             // region base positions (gathered from the system at search start)
      uint8_t* rtype;            // region type ids (gathered)
      int const* rgidx;          // region atom -> global atom of the system
      double* rhof;              // frozen part of S1 densities
      double* fp;                // F'(rho) per region atom
      int const* rowptr;
      unsigned const* ent;
      unsigned* act;             // active entries (same layout as ent)
      int* alen;                 // active entries per row
      double* ws;
      double const* start_x;
      double const* start_ax;
      double* out_x;
      double* out_ax;
      double const* hist;        // known-saddle vectors of every search (SearchHdr::hist_off)
      double const* hnorm;       // their norms (SearchHdr::hn_off)
      double* out_rev;           // mechanism minima (when mechanisms are on)
      double* out_fwd;
      SearchOut* out;
      double const* rs_in;       // saved states of resumed searches (SearchHdr::rs_in_off)
      double* rs_out;            // rs_slots slots of rs_len doubles (host-visible memory)
      unsigned* rs_count;        // slots taken
      int rs_slots;
      long rs_len;
      unsigned* queue;
      int* done;                 // per-search completion flags (host-visible memory)
      DimOpt o;
      int t0;   // the one species present when the kernel is instantiated with MULTI = false
    };

    // ---------------------------------------------------------------------------------------------
    // Workgroup primitives.  Element i of a 3*nm vector is always handled by thread i % T, so
    // elementwise chains need no barrier; reductions are fixed-order trees (bitwise reproducible).
    // Every device function of the search kernel is force-inlined: an out-of-line call spills the
    // caller's live registers around it and pins the kernel argument block in per-lane scratch.
    // ---------------------------------------------------------------------------------------------

    template <int T>
    struct Team {
      double* red;   // LDS, RED_LDS doubles: two 64-double reduction buffers, then top-2 partials
      int tid;

      __device__ __forceinline__ double wave(double v) const {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
      }

      // Sums up to four values over the workgroup in an order that does not depend on T: element i of a
      // vector belongs to virtual slot i mod 1024, each slot adds its elements in increasing i; the
      // slots form 16 virtual wavefronts of 64 (slot s in wavefront s / 64, lane s mod 64), each reduced
      // by the xor shuffle tree 32, 16, ..., 1; the 16 wavefront sums are then added in order 0..15.
      // v[q][k] is the slot tid + q T of value k; every thread gets the totals.  Two barriers.
      static constexpr int VS = 1024;
      static constexpr int NQ = VS / T;
      static constexpr int VW = VS / 64;

      // The partials alternate between two LDS buffers (red[0..63], red[64..127]): a thread can only
      // write a buffer again after every thread has passed the barrier of the reduction in between,
      // i.e. after every read of the previous use, so one barrier per reduction suffices.
      mutable int par = 0;

      template <int K>
      __device__ __forceinline__ void tree(double (&v)[NQ][4]) const {
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int k = 0; k < K; ++k)
#pragma unroll
            for (int bit = 32; bit >= 1; bit >>= 1) v[q][k] = __dadd_rn(v[q][k], __shfl_xor(v[q][k], bit, 64));
        double* buf = red + 64 * par;
        par ^= 1;
        if ((tid & 63) == 0) {
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int k = 0; k < K; ++k) buf[4 * ((tid >> 6) + q * (T / 64)) + k] = v[q][k];
        }
        __syncthreads();
        double t[K];
#pragma unroll
        for (int k = 0; k < K; ++k) t[k] = buf[k];
        for (int w = 1; w < VW; ++w)
#pragma unroll
          for (int k = 0; k < K; ++k) t[k] = __dadd_rn(t[k], buf[4 * w + k]);
#pragma unroll
        for (int k = 0; k < K; ++k) v[0][k] = t[k];
      }

      template <int K>
      __device__ __forceinline__ void sum(double* out, double (&v)[NQ][4]) const {
        tree<K>(v);
#pragma unroll
        for (int k = 0; k < K; ++k) out[k] = v[0][k];
      }

      __device__ __forceinline__ double dot(double const* a, double const* b, int n) const {
        double v[NQ][4] = {};
        for (int i0 = 0; i0 < n; i0 += VS)
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            int i = i0 + tid + q * T;
            if (i < n) v[q][0] = __dadd_rn(v[q][0], a[i] * b[i]);
          }
        double o[1];
        sum<1>(o, v);
        return o[0];
      }

      __device__ __forceinline__ void dot2(double const* a, double const* b, double const* c, double const* d, int n, double& x,
                           double& y) const {
        double v[NQ][4] = {};
        for (int i0 = 0; i0 < n; i0 += VS)
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            int i = i0 + tid + q * T;
            if (i < n) {
              v[q][0] = __dadd_rn(v[q][0], a[i] * b[i]);
              v[q][1] = __dadd_rn(v[q][1], c[i] * d[i]);
            }
          }
        double o[2];
        sum<2>(o, v);
        x = o[0];
        y = o[1];
      }

      // Workgroup sum of per-thread values that are already in a fixed per-row form (region energy).
      __device__ __forceinline__ double sum1_rows(double x) const {
        double v[NQ][4] = {};
        v[0][0] = x;
        double o[1];
        sum<1>(o, v);
        return o[0];
      }

      // o = ca a + cb b + cc c (b, c may be null); o may alias a.
      __device__ __forceinline__ void lin(double* o, int n, double const* a, double ca, double const* b = nullptr, double cb = 0,
                          double const* c = nullptr, double cc = 0) const {
        for (int i = tid; i < n; i += T) {
          double v = ca * a[i];
          if (b) v += cb * b[i];
          if (c) v += cc * c[i];
          o[i] = v;
        }
      }

      __device__ __forceinline__ void sync() const { __syncthreads(); }

      __device__ __forceinline__ unsigned any_or(unsigned e) const {
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
    // One wavefront's vector engine.  The searches' vector algebra (L-BFGS two-loops, rotor, trust
    // steps: 3*nm elements, a few hundred to a few thousand) runs on wave 0 of the workgroup while the
    // other waves wait at the next barrier: element i belongs to lane i % 64, each lane adds its elements
    // in increasing i, and the 64 lane sums meet in a fixed tree (quad butterflies, two rotations within
    // each row of 16, then the four row sums in order).  A dot product needs no barrier and no LDS round
    // trip, and its bits do not depend on the workgroup shape.
    // ---------------------------------------------------------------------------------------------

    struct Wave {
      int lane;

      template <int CTRL>
      __device__ __forceinline__ static double dpp(double v) {
        return __builtin_amdgcn_update_dpp(0.0, v, CTRL, 0xF, 0xF, false);
      }
      __device__ __forceinline__ static double lane_value(double v, int l) {
        long long b = __double_as_longlong(v);
        int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
        int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
        return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
      }
      // Sum of the 64 lanes' values; every lane gets the same bits.
      __device__ __forceinline__ static double total(double v) {
        v = __dadd_rn(v, dpp<0xB1>(v));    // quad_perm [1,0,3,2]
        v = __dadd_rn(v, dpp<0x4E>(v));    // quad_perm [2,3,0,1]
        v = __dadd_rn(v, dpp<0x124>(v));   // row_ror:4
        v = __dadd_rn(v, dpp<0x128>(v));   // row_ror:8: lane 16 r holds row r's sum
        return __dadd_rn(__dadd_rn(lane_value(v, 0), lane_value(v, 16)),
                         __dadd_rn(lane_value(v, 32), lane_value(v, 48)));
      }

      __device__ __forceinline__ double dot(double const* a, double const* b, int n) const {
        double s = 0;
        for (int i = lane; i < n; i += 64) s = __dadd_rn(s, a[i] * b[i]);
        return total(s);
      }

      __device__ __forceinline__ void dot2(double const* a, double const* b, double const* c, double const* d, int n,
                                           double& x, double& y) const {
        double s = 0, t = 0;
        for (int i = lane; i < n; i += 64) {
          s = __dadd_rn(s, a[i] * b[i]);
          t = __dadd_rn(t, c[i] * d[i]);
        }
        x = total(s);
        y = total(t);
      }

      // A wavefront executes in lockstep; only the compiler must keep LDS accesses in program order.
      __device__ __forceinline__ void sync() const { __builtin_amdgcn_wave_barrier(); }

      // o = ca a + cb b + cc c (b, c may be null); o may alias a.
      __device__ __forceinline__ void lin(double* o, int n, double const* a, double ca, double const* b = nullptr,
                                          double cb = 0, double const* c = nullptr, double cc = 0) const {
        for (int i = lane; i < n; i += 64) {
          double v = ca * a[i];
          if (b) v += cb * b[i];
          if (c) v += cc * c[i];
          o[i] = v;
        }
      }
    };

    // ---------------------------------------------------------------------------------------------
    // Per-search view
    // ---------------------------------------------------------------------------------------------

    // Spline descriptors of the potential, staged in LDS at kernel start.  The kernels index them by
    // atom types; indexing the by-value kernel argument instead would place the whole argument block
    // in per-lane scratch.
    struct SplineSet {
      SplineDesc F[MAX_TYPES];
      SplineDesc phi[MAX_TYPES * MAX_TYPES];
      SplineDesc V[MAX_TYPES * MAX_TYPES];
    };

    struct SView {
      int nm, nr, n3;
      SplineSet const* spl; // LDS copy of the spline descriptors
      double3 const* S;     // image shifts (LDS copy)
      double const* srho;   // system base densities
      double4* cur;
      double4 const* base;
      uint8_t const* type;
      int const* gidx;
      double* rhof;
      double* fp;
      int const* rowptr;
      unsigned const* ent;
      unsigned* act;
      int* alen;
      double act_f, act_m;   // active-list skins (mobile-frozen, mobile-mobile)
      unsigned* errw;        // LDS: spline-error bits of the search's region gradients
    };

    __device__ inline int vix(int nt, int a, int b) { return a > b ? a + nt * b : b + nt * a; }

    // ---------------------------------------------------------------------------------------------
    // Pipelined neighbour walks.  A thread handles U = 4 entries per step: all entry loads, then all
    // position loads, then all table loads are issued before any arithmetic, and an entry that is
    // past the row end or beyond r_c takes a dummy in-range table slot and is not accumulated, so the
    // sums equal those of the plain one-entry-at-a-time loop bit for bit.
    // ---------------------------------------------------------------------------------------------

    constexpr int U = 4;

    // Geometry of U entries of one row: entry k_q = k0 + q * stride (valid while < c).
    template <bool MIX>
    struct Quad {
      using R = typename Seg<MIX>::R;
      double dx[U], dy[U], dz[U], w[U];
      int j[U];
      bool in[U];
      R r[U], rinv[U];
    };

    template <bool MIX>
    __device__ inline void load_quad(SView const& v, unsigned const* ent, double rc2, double4 const& pi, int k0,
                                     int stride, int b, int c, Quad<MIX>& Q) {
      unsigned en[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        int k = k0 + q * stride;
        en[q] = ent[k < c ? k : b];
      }
      double4 pj[U];
      double3 sh[U];
#pragma unroll
      for (int q = 0; q < U; ++q) {
        Q.j[q] = static_cast<int>(en[q] & NB_IDX_MASK);
        pj[q] = v.cur[Q.j[q]];
        sh[q] = v.S[en[q] >> NB_SHIFT_BITS];
      }
#pragma unroll
      for (int q = 0; q < U; ++q) {
        Q.dx[q] = (pj[q].x + sh[q].x) - pi.x;
        Q.dy[q] = (pj[q].y + sh[q].y) - pi.y;
        Q.dz[q] = (pj[q].z + sh[q].z) - pi.z;
        Q.w[q] = pj[q].w;
        double r2 = Q.dx[q] * Q.dx[q] + Q.dy[q] * Q.dy[q] + Q.dz[q] * Q.dz[q];
        Q.in[q] = (k0 + q * stride < c) && (r2 < rc2);
        if constexpr (MIX) {
          float r2f = Q.in[q] ? static_cast<float>(r2) : 1.0f;
          Q.rinv[q] = __frsqrt_rn(r2f);
          Q.r[q] = r2f * Q.rinv[q];
        } else {
          Q.r[q] = Q.in[q] ? sqrt(r2) : 1.0;
          Q.rinv[q] = 1.0;
        }
      }
    }

    constexpr int RED_LDS = 256;   // the workgroup reductions use 160 doubles of it at most

    // Region rows staged in LDS per workgroup size (96 KB at 1024 threads: one workgroup per CU anyway).
    template <int T>
    __host__ __device__ constexpr int lds_rows() { return T >= 1024 ? 3072 : (T >= 512 ? 1792 : 1024); }
    // LDS doubles for the rotor's workspace vectors (the rest of the 160 KB after the reduction buffer
    // and the region rows).
    template <int T>
    __host__ __device__ constexpr int lds_ws() { return T >= 1024 ? 6800 : (T >= 512 ? 11750 : 14700); }

    // Rebuilds the active lists from the candidate rows at the current region positions (stable:
    // an active row is the ordered subsequence of its candidate row, whatever the workgroup shape).
    template <int T, int L>
    __device__ __forceinline__ void refresh_active(BatchArgs const& A, SView const& v, Team<T> const& tm) {
      double const cf = A.T.rc + v.act_f, cm = A.T.rc + v.act_m;
      double const cf2 = cf * cf, cm2 = cm * cm;
      int const lane = tm.tid % L;
      int const gbase = (tm.tid & 63) - lane;   // first lane of this row group in the wavefront
      for (int r0 = tm.tid / L; r0 < v.nm; r0 += T / L) {
        double4 pi = v.cur[r0];
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        int cnt = 0;
        for (int k0 = b; k0 < c; k0 += L) {
          int k = k0 + lane;
          bool keep = false;
          unsigned en = 0;
          if (k < c) {
            en = v.ent[k];
            int j = static_cast<int>(en & NB_IDX_MASK);
            double3 sh = v.S[en >> NB_SHIFT_BITS];
            double4 pj = v.cur[j];
            double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
            double r2 = dx * dx + dy * dy + dz * dz;
            keep = r2 < (j < v.nm ? cm2 : cf2);
          }
          unsigned long long m = __ballot(keep);
          unsigned gm = static_cast<unsigned>((m >> gbase) & ((1ull << L) - 1ull));
          if (keep) v.act[b + cnt + __popc(gm & ((1u << lane) - 1u))] = en;
          cnt += __popc(gm);
        }
        if (lane == 0) v.alen[r0] = cnt;
      }
      for (int r0 = v.nm + tm.tid; r0 < v.nr; r0 += T) {
        double4 pi = v.cur[r0];
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        int cnt = 0;
        for (int k = b; k < c; ++k) {
          unsigned en = v.ent[k];
          int j = static_cast<int>(en & NB_IDX_MASK);
          double3 sh = v.S[en >> NB_SHIFT_BITS];
          double4 pj = v.cur[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          if (dx * dx + dy * dy + dz * dz < cf2) v.act[b + cnt++] = en;
        }
        v.alen[r0] = cnt;
      }
      __syncthreads();
    }

    // EAM gradient of the mobile atoms at the current region positions (pass 1 rho/F' over M u S1,
    // pass 2 forces over M).  MIX: FP32 pair/spline arithmetic with FP64 accumulation.
    template <int T, int L, bool MIX, bool MULTI>
    __device__ __forceinline__ unsigned region_gradient(BatchArgs const& A, SView const& v, Team<T> const& tm, double* g) {
      using R = typename Seg<MIX>::R;
      constexpr int NS = 16 / L;
      static_assert(U % NS == 0, "a step must cover whole slot cycles");
      EAMTables const& Tb = A.T;
      int const nt = Tb.ntypes;
      SplineSet const& SS = *v.spl;
      // One species: its three descriptors, read once.
      int const t0 = MULTI ? 0 : A.t0;
      SplineDesc const phi0 = SS.phi[t0 + nt * t0], V0 = SS.V[t0 + nt * t0], F0 = SS.F[t0];
      auto PHI = [&](int tj, int ti) -> SplineDesc const& {
        if constexpr (MULTI) return SS.phi[tj + nt * ti]; else return phi0;
      };
      auto VV = [&](int ti, int tj) -> SplineDesc const& {
        if constexpr (MULTI) return SS.V[vix(nt, ti, tj)]; else return V0;
      };
      auto FF = [&](int ti) -> SplineDesc const& {
        if constexpr (MULTI) return SS.F[ti]; else return F0;
      };
      unsigned e = 0;
      int const lane = tm.tid % L;
      // Densities of U entries (FP64 values, zero where not used).
      auto densities = [&](Quad<MIX> const& Q, int ti, double (&out)[U]) {
        Seg<MIX> sg[U];
        decltype(coef_of<MIX>(Tb, phi0, sg[0])) cf[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
          int tj = MULTI ? v.type[Q.j[q]] : t0;
          SplineDesc const& sd = PHI(tj, ti);
          sg[q] = seg_of<MIX>(sd, Q.r[q], Q.in[q], e);
          cf[q] = coef_of<MIX>(Tb, sd, sg[q]);
        }
#pragma unroll
        for (int q = 0; q < U; ++q) out[q] = sg[q].use ? static_cast<double>(poly0(cf[q], sg[q].t)) : 0.0;
      };
      // Mobile rows (long: every region neighbour): L lanes per row, 16-slot sums.
      for (int r0 = tm.tid / L; r0 < v.nm; r0 += T / L) {
        double4 pi = v.cur[r0];
        int ti = MULTI ? v.type[r0] : t0;
        Acc16<L> rho_acc;
        int b = v.rowptr[r0], c = b + v.alen[r0];
        for (int k0 = b + lane; k0 < c; k0 += U * L) {
          Quad<MIX> Q;
          load_quad<MIX>(v, v.act, Tb.rc2, pi, k0, L, b, c, Q);
          double d[U];
          densities(Q, ti, d);
#pragma unroll
          for (int q = 0; q < U; ++q)
            if (Q.in[q]) rho_acc.add_slot(q % NS, d[q]);
        }
        double rho = rho_acc.total();
        if (lane == 0) v.cur[r0].w = spl64<1>(Tb, FF(ti), rho, e);
      }
      // Frozen-shell rows (short: mobile neighbours only): one thread per row, entries summed in order,
      // then the frozen part of the density.  Neither order depends on the workgroup shape.
      for (int r0 = v.nm + tm.tid; r0 < v.nr; r0 += T) {
        double4 pi = v.cur[r0];
        int ti = MULTI ? v.type[r0] : t0;
        int b = v.rowptr[r0], c = b + v.alen[r0];
        double rho = 0.0;
        for (int k0 = b; k0 < c; k0 += U) {
          Quad<MIX> Q;
          load_quad<MIX>(v, v.act, Tb.rc2, pi, k0, 1, b, c, Q);
          double d[U];
          densities(Q, ti, d);
#pragma unroll
          for (int q = 0; q < U; ++q)
            if (Q.in[q]) rho = __dadd_rn(rho, d[q]);
        }
        rho = __dadd_rn(rho, v.rhof[r0]);
        v.cur[r0].w = spl64<1>(Tb, FF(ti), rho, e);
      }
      __syncthreads();
      for (int r0 = tm.tid / L; r0 < v.nm; r0 += T / L) {
        double4 pi = v.cur[r0];
        double fpi = pi.w;
        int ti = MULTI ? v.type[r0] : t0;
        Acc16<L> gx, gy, gz;
        int b = v.rowptr[r0], c = b + v.alen[r0];
        for (int k0 = b + lane; k0 < c; k0 += U * L) {
          Quad<MIX> Q;
          load_quad<MIX>(v, v.act, Tb.rc2, pi, k0, L, b, c, Q);
          Seg<MIX> sv[U], sa[U], sb[U];
          decltype(coef_of<MIX>(Tb, phi0, sv[0])) cv[U], ca[U], cb[U];
#pragma unroll
          for (int q = 0; q < U; ++q) {
            int tj = MULTI ? v.type[Q.j[q]] : t0;
            SplineDesc const& dv = VV(ti, tj);
            SplineDesc const& da = PHI(tj, ti);
            sv[q] = seg_of<MIX>(dv, Q.r[q], Q.in[q], e);
            sa[q] = seg_of<MIX>(da, Q.r[q], Q.in[q], e);
            cv[q] = coef_of<MIX>(Tb, dv, sv[q]);
            ca[q] = coef_of<MIX>(Tb, da, sa[q]);
            if (MULTI) {
              SplineDesc const& db = PHI(ti, tj);
              sb[q] = seg_of<MIX>(db, Q.r[q], Q.in[q], e);
              cb[q] = coef_of<MIX>(Tb, db, sb[q]);
            }
          }
#pragma unroll
          for (int q = 0; q < U; ++q) {
            R vp = sv[q].use ? poly1(cv[q], sv[q].t) : static_cast<R>(0);
            R pa = sa[q].use ? poly1(ca[q], sa[q].t) : static_cast<R>(0);
            R pb = MULTI ? (sb[q].use ? poly1(cb[q], sb[q].t) : static_cast<R>(0)) : pa;
            double cf;
            if constexpr (MIX) {
              float mag = vp + static_cast<float>(fpi) * pa + static_cast<float>(Q.w[q]) * pb;
              cf = static_cast<double>(mag * Q.rinv[q]);
            } else {
              cf = (vp + fpi * pa + Q.w[q] * pb) / Q.r[q];
            }
            if (Q.in[q]) {
              gx.add_slot(q % NS, -cf * Q.dx[q]);
              gy.add_slot(q % NS, -cf * Q.dy[q]);
              gz.add_slot(q % NS, -cf * Q.dz[q]);
            }
          }
        }
        double tx = gx.total(), ty = gy.total(), tz = gz.total();
        if (lane == 0) {
          g[3 * r0] = tx;
          g[3 * r0 + 1] = ty;
          g[3 * r0 + 2] = tz;
        }
      }
      // Spline errors gather in one LDS word (sticky for the search): every thread reads it after the barrier.
      if (e) atomicOr(v.errw, e);
      __syncthreads();
      return *v.errw;
    }

    // FP64 local energy (all terms that depend on mobile positions).  With base = true the densities
    // of the region come from the global base pass and S1's frozen density part is recorded.
    template <int T, int L, bool MULTI>
    __device__ __forceinline__ double region_energy(BatchArgs const& A, SView const& v, Team<T> const& tm, bool base, unsigned& err) {
      EAMTables const& Tb = A.T;
      int const nt = Tb.ntypes;
      SplineSet const& SS = *v.spl;
      int const t0 = MULTI ? 0 : A.t0;
      unsigned e = 0;
      int const lane = tm.tid % L;
      // Per-row energies go to the fp scratch (overwritten by the next gradient), then a fixed-order sum.
      for (int r0 = tm.tid / L; r0 < v.nr; r0 += T / L) {
        double4 pi = v.cur[r0];
        int ti = MULTI ? v.type[r0] : t0;
        Acc16<L> rho_acc, pair_acc;
        bool mob = r0 < v.nm;
        int b = v.rowptr[r0], c = v.rowptr[r0 + 1];
        for (int k = b + lane, m = 0; k < c; k += L, ++m) {
          unsigned en = v.ent[k];
          int j = static_cast<int>(en & NB_IDX_MASK);
          double3 sh = v.S[en >> NB_SHIFT_BITS];
          double4 pj = v.cur[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < Tb.rc2) {
            double r = sqrt(r2);
            int tj = MULTI ? v.type[j] : t0;
            rho_acc.add(m, spl64<0>(Tb, SS.phi[tj + nt * ti], r, e));
            if (mob) pair_acc.add(m, (j < v.nm ? 0.5 : 1.0) * spl64<0>(Tb, SS.V[vix(nt, ti, tj)], r, e));
          }
        }
        double rho = rho_acc.total();
        double pair = pair_acc.total();
        if (lane == 0) {
          double rt;
          if (base) {
            rt = v.srho[v.gidx[r0]];
            if (!mob) v.rhof[r0] = rt - rho;
          } else {
            rt = mob ? rho : rho + v.rhof[r0];
          }
          v.fp[r0] = spl64<0>(Tb, SS.F[ti], rt, e) + pair;
        }
      }
      __syncthreads();
      double vv[Team<T>::NQ][4] = {};
      for (int i0 = 0; i0 < v.nr; i0 += Team<T>::VS)
#pragma unroll
        for (int q = 0; q < Team<T>::NQ; ++q) {
          int i = i0 + tm.tid + q * T;
          if (i < v.nr) vv[q][0] = __dadd_rn(vv[q][0], v.fp[i]);
        }
      double o[1];
      tm.template sum<1>(o, vv);
      err |= tm.any_or(e);
      return o[0];
    }

    // cur[i] = x + c * d for the mobile atoms (d may be null).
    template <int T>
    __device__ __forceinline__ void place(SView const& v, Team<T> const& tm, double const* x, double c, double const* d) {
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

    // L-BFGS two-loop recursion (lbfgs_core.cpp:29-87) on a team (the workgroup or one wavefront).
    template <class TM>
    struct Core {
      double* s;      // m vectors
      double* y;      // m vectors
      double* px;
      double* pg;
      double* q;
      double* rho;    // LDS, m
      double* alpha;  // LDS, m
      int m, k, n3;

      // The scalars rho and alpha live in LDS; every lane of the team stores the same value.
      __device__ __forceinline__ double* newton_step(TM const& tm, double const* x, double const* g) {
        if constexpr (std::is_same_v<TM, Wave>) {
          return newton_step_fused(tm, x, g);
        } else {
          return newton_step_passes(tm, x, g);
        }
      }

      // On one wavefront: the same recursion with its element passes fused.  A lane owns the same
      // elements in every pass, so one pass can finish an update of q and start the next dot product
      // with q's new elements: 2 bound + 1 passes instead of 4 bound + 8.  Every element takes the same
      // expressions as in the separate passes (lin's v = 1 q; v += c u, dot's __dadd_rn over a * b in
      // element order, the same 64-lane total), so the results are the same bits.
      // fdot (optional): *fdot_out = dot(result, fdot), taken in the last pass.
      __device__ __forceinline__ double* newton_step_fused(Wave const& w, double const* x, double const* g,
                                                           double const* fdot = nullptr, double* fdot_out = nullptr) {
        int const prev = k > 0 ? (k - 1) % m : 0;
        int const incur = k <= m ? 0 : k - m;
        int const bound = k <= m ? k : m;
        double* const sp = s + static_cast<long>(prev) * n3;
        double* const yp = y + static_cast<long>(prev) * n3;
        // Pass 1: the newest (s, y) pair, the copies of x and g, s.y, y.y and the first loop's first dot.
        int jd = bound > 0 ? (bound - 1 + incur) % m : 0;
        double const* dv = s + static_cast<long>(jd) * n3;
        double sy = 0, yy = 0, d = 0, fd = 0;
        for (int i = w.lane; i < n3; i += 64) {
          double const xi = x[i], gi = g[i];
          if (k > 0) {
            double vs = 1.0 * xi;
            vs += -1.0 * px[i];
            double vy = 1.0 * gi;
            vy += -1.0 * pg[i];
            sp[i] = vs;
            yp[i] = vy;
            sy = __dadd_rn(sy, vs * vy);
            yy = __dadd_rn(yy, vy * vy);
          }
          px[i] = 1.0 * xi;
          pg[i] = 1.0 * gi;
          double const qi = 1.0 * gi;
          q[i] = qi;
          if (bound > 0) d = __dadd_rn(d, dv[i] * qi);
          if (bound == 0 && fdot) fd = __dadd_rn(fd, qi * fdot[i]);
        }
        double yy_t = 0;
        if (k > 0) {
          rho[prev] = 1.0 / fabs(Wave::total(sy));
          yy_t = Wave::total(yy);
        }
        w.sync();
        if (bound == 0) {
          if (fdot) *fdot_out = Wave::total(fd);
          ++k;
          w.sync();
          return q;
        }
        double dot = Wave::total(d);
        // A pass: q += c u (when u), then q = sc q (when scaled), then the dot of q with dv (when dv).
        auto pass = [&](double const* u, double c, bool scaled, double sc, double const* dvec) {
          double acc = 0;
          for (int i = w.lane; i < n3; i += 64) {
            double qi = q[i];
            if (u) {
              double v = 1.0 * qi;
              v += c * u[i];
              qi = v;
            }
            if (scaled) qi = sc * qi;
            q[i] = qi;
            if (dvec) acc = __dadd_rn(acc, dvec[i] * qi);
          }
          return dvec ? Wave::total(acc) : 0.0;
        };
        // First loop, newest pair first: alpha_j = rho_j s_j.q, then q -= alpha_j y_j.
        double const* pu = nullptr;
        double pc = 0;
        for (int i = bound - 1; i >= 0; --i) {
          int const j = (i + incur) % m;
          if (i < bound - 1) dot = pass(pu, pc, false, 0.0, s + static_cast<long>(j) * n3);
          double const a = rho[j] * dot;
          alpha[j] = a;
          pu = y + static_cast<long>(j) * n3;
          pc = -a;
        }
        // H0 scaling, then the second loop, oldest pair first: beta = rho_j y_j.q, q += (alpha_j - beta) s_j.
        bool scaled = k > 0;
        double const scale = k > 0 ? 1.0 / rho[prev] / yy_t : 1.0;
        for (int i = 0; i < bound; ++i) {
          int const j = (i + incur) % m;
          double const beta = rho[j] * pass(pu, pc, scaled, scale, y + static_cast<long>(j) * n3);
          scaled = false;
          pu = s + static_cast<long>(j) * n3;
          pc = alpha[j] - beta;
        }
        if (fdot) {
          *fdot_out = pass(pu, pc, false, 0.0, fdot);
        } else {
          pass(pu, pc, false, 0.0, nullptr);
        }
        ++k;
        w.sync();
        return q;
      }

      __device__ __forceinline__ double* newton_step_passes(TM const& tm, double const* x, double const* g) {
        int prev = k > 0 ? (k - 1) % m : 0;
        if (k > 0) {
          double* sp = s + static_cast<long>(prev) * n3;
          double* yp = y + static_cast<long>(prev) * n3;
          tm.lin(sp, n3, x, 1.0, px, -1.0);
          tm.lin(yp, n3, g, 1.0, pg, -1.0);
          double sy = tm.dot(sp, yp, n3);
          rho[prev] = 1.0 / fabs(sy);
        }
        tm.lin(px, n3, x, 1.0);
        tm.lin(pg, n3, g, 1.0);
        tm.lin(q, n3, g, 1.0);
        int incur = k <= m ? 0 : k - m;
        int bound = k <= m ? k : m;
        tm.sync();
        for (int i = bound - 1; i >= 0; --i) {
          int j = (i + incur) % m;
          double a = rho[j] * tm.dot(s + static_cast<long>(j) * n3, q, n3);
          alpha[j] = a;
          tm.lin(q, n3, q, 1.0, y + static_cast<long>(j) * n3, -a);
        }
        if (k > 0) {
          double* yp = y + static_cast<long>(prev) * n3;
          double scale = 1.0 / rho[prev] / tm.dot(yp, yp, n3);
          tm.lin(q, n3, q, scale);
        }
        tm.sync();
        for (int i = 0; i < bound; ++i) {
          int j = (i + incur) % m;
          double beta = rho[j] * tm.dot(y + static_cast<long>(j) * n3, q, n3);
          tm.lin(q, n3, q, 1.0, s + static_cast<long>(j) * n3, alpha[j] - beta);
        }
        ++k;
        tm.sync();
        return q;
      }
    };

    template <int T, int L, bool MIX, bool MULTI>
    __device__ __forceinline__ void run_search(BatchArgs const& A, int sid, Team<T> const& tm, double* lds_core, double3* lds_S,
                               double4* lds_area, SplineSet const* lds_spl, unsigned* lds_err,
                               double* bc) {
      long const t_begin = static_cast<long>(wall_clock64());
      SearchHdr const h = A.hdr[sid];
      DimOpt const& o = A.o;
      SView v;
      v.nm = h.nm;
      v.nr = h.nr;
      v.n3 = 3 * h.nm;
      v.S = lds_S;
      v.spl = lds_spl;
      v.srho = h.srho;
      // Region positions (w = F' between the two passes of a gradient) live in LDS when they fit.
      bool const pos_lds = h.nr <= lds_rows<T>();
      v.cur = pos_lds ? lds_area : A.cur + h.reg_off;
      // The workspace part of the area starts after the positions; what rows a small region leaves
      // unused goes to the workspace.
      int const area = 4 * lds_rows<T>() + 4 * (lds_ws<T>() / 4);
      double* const lds_ws_buf = reinterpret_cast<double*>(lds_area) + (pos_lds ? 4 * h.nr : 0);
      int const ws_cap = area - (pos_lds ? 4 * h.nr : 0);
      v.base = A.base + h.reg_off;
      v.type = A.rtype + h.reg_off;
      v.gidx = A.rgidx + h.reg_off;
      v.rhof = A.rhof + h.reg_off;
      v.fp = A.fp + h.reg_off;
      v.rowptr = A.rowptr + h.row_off;
      v.ent = A.ent;
      v.act = A.act;
      v.alen = A.alen + h.reg_off;
      v.act_f = fmin(ACT_SKIN_F, h.skin_f);
      v.act_m = fmin(ACT_SKIN_M, h.skin_m);
      v.errw = lds_err;
      // The vector algebra runs on wave 0 (w0); bc[] carries its results to the other waves.
      bool const w0 = tm.tid < 64;
      Wave const wv{tm.tid & 63};
      int const n3 = v.n3;
      // Workspace: the vectors the rotor touches every iteration come first, most used first (the
      // named ones, then the rotation core's q, x_prev, g_prev, s and y histories), then eff and xb (once
      // per step); as many of them as fit live in LDS, the rest in global memory.  Where a vector lives
      // changes no result.
      double* W = A.ws + h.ws_off;
      int const hot = 13 + 2 * o.rn;
      int const r0v = 8;   // rotation core: q, x_prev, g_prev, then rn s and rn y vectors
      // The cores address their s and y histories as blocks (s + j n3), so a block is in LDS whole or not
      // at all: a count of LDS vectors that ends inside a block stops at the block's start.
      int fit = n3 > 0 ? min(hot, ws_cap / n3) : hot;
      if (fit > r0v + 3 && fit < r0v + 3 + o.rn) fit = r0v + 3;
      if (fit > r0v + 3 + o.rn && fit < r0v + 3 + 2 * o.rn) fit = r0v + 3 + o.rn;
      // Then, when all of those are in LDS, the translation core (its s, y, x_prev, g_prev and q as one
      // block, which the saved state also addresses as a block) and xr, as far as they fit.
      int const tend = hot + 2 * o.n + 3;   // xr
      if (fit == hot && n3 > 0) {
        if (static_cast<long>(tend + 1) * n3 <= ws_cap) {
          fit = tend + 1;
        } else if (static_cast<long>(tend) * n3 <= ws_cap) {
          fit = tend;
        }
      }
      auto vecp = [&](int k) { return (k < fit ? lds_ws_buf : W) + static_cast<long>(k) * n3; };
      double* x = vecp(0);
      double* ax = vecp(1);
      double* g0 = vecp(2);
      double* g1 = vecp(3);
      double* g1p = vecp(4);
      double* dg = vecp(5);
      double* dgp = vecp(6);
      double* axp = vecp(7);
      double* eff = vecp(r0v + 3 + 2 * o.rn);
      double* xb = vecp(r0v + 4 + 2 * o.rn);
      double* xr = vecp(nvec(o.n, o.rn) - 1);   // positions at the last active-list refresh
      int const t0 = 13 + 2 * o.rn;
      Core<Wave> tc{vecp(t0), vecp(t0 + o.n), vecp(t0 + 2 * o.n), vecp(t0 + 2 * o.n + 1), vecp(t0 + 2 * o.n + 2),
                    lds_core, lds_core + 32, o.n, 0, n3};
      Core<Wave> rc{vecp(r0v + 3), vecp(r0v + 3 + o.rn), vecp(r0v + 1), vecp(r0v + 2), vecp(r0v), lds_core + 64,
                    lds_core + 96, o.rn, 0, n3};
      // The mechanism relaxations keep the workgroup-wide algebra (same memory as the translation core).
      Core<Team<T>> tcm{tc.s, tc.y, tc.px, tc.pg, tc.q, tc.rho, tc.alpha, o.n, 0, n3};
      double const dr = o.delta_r;
      unsigned err = 0;
      long grads = 0;
      int rot_iters = 0;

      // Gather the region from the resident system, then the base state: region at base positions,
      // frozen density parts, base local energy.
      if (tm.tid < 27) lds_S[tm.tid] = h.sS[tm.tid];
      if (tm.tid == 0) *lds_err = 0;
      for (int i = tm.tid; i < v.nr; i += T) {
        int gi = v.gidx[i];
        double4 b = h.sbase[gi];
        A.base[h.reg_off + i] = b;
        A.rtype[h.reg_off + i] = h.stype[gi];
        v.cur[i] = b;
      }
      __syncthreads();
      double e_base = region_energy<T, L, MULTI>(A, v, tm, true, err);
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
      double maxd = 0, maxp = 0;
      int refreshes = 0;

      // Saved state of a region exit (resume_len layout): x, ax, the translation core (s, y, x_prev,
      // g_prev, q: 2m + 3 consecutive workspace vectors) and its rho values.  Every thread takes part.
      double* const tblk = vecp(t0);
      long const nblk = static_cast<long>(2 * o.n + 3) * n3;
      auto state_elem = [&](long i) -> double* {
        if (i < n3) return x + i;
        if (i < 2L * n3) return ax + (i - n3);
        if (i < 2L * n3 + nblk) return tblk + (i - 2L * n3);
        return tc.rho + (i - 2L * n3 - nblk);
      };
      long const rlen = resume_len(n3, o.n);
      if (h.resume) {
        double const* src = A.rs_in + h.rs_in_off;
        for (long i = tm.tid; i < rlen; i += T) *state_elem(i) = src[i];
        __syncthreads();
        tc.k = h.r_k;
        grads = h.r_grads;
        rot_iters = h.r_rot_iters;
        refreshes = h.r_refreshes;
        maxd = h.r_maxd;
        maxp = h.r_maxp;
      }
      int saved = 0;
      // A region exit inside the step loop saves the state to a free slot of rs_out, if there is one.
      auto save_state = [&]() {
        if (tm.tid == 0) bc[6] = static_cast<double>(atomicAdd(A.rs_count, 1u));
        __syncthreads();
        int const slot = static_cast<int>(bc[6]);
        __syncthreads();
        if (slot >= A.rs_slots || rlen > A.rs_len) return;
        double* dst = A.rs_out + static_cast<long>(slot) * A.rs_len;
        for (long i = tm.tid; i < rlen; i += T) dst[i] = *state_elem(i);
        saved = slot + 1;
        __syncthreads();
      };

      // Rows built at the base positions stay complete while every pair that is missing from them stays
      // beyond r_c: a mobile-frozen pair needs d_i < skin_f and a mobile-mobile pair d_i + d_j < skin_m,
      // i.e. d1 < skin_f and d1 + d2 < skin_m for the two largest mobile displacements d1 >= d2 (plus
      // the dimer separation, as the second image sits delta_r further along the axis).
      long gticks = 0;
      auto timed_gradient = [&](double* g) -> unsigned {
        long t = static_cast<long>(wall_clock64());
        unsigned r = region_gradient<T, L, MIX, MULTI>(A, v, tm, g);
        gticks += static_cast<long>(wall_clock64()) - t;
        return r;
      };
      // Two largest mobile displacements |x_i - ref_i| (a1 >= a2) over the workgroup.
      auto top2 = [&](double const* ref, double& a1, double& a2) {
        double m1 = 0, m2 = 0;
        for (int a = tm.tid; a < v.nm; a += T) {
          double dx = x[3 * a] - ref[3 * a], dy = x[3 * a + 1] - ref[3 * a + 1], dz = x[3 * a + 2] - ref[3 * a + 2];
          double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (!(d <= m1)) {   // also takes NaN, which then fails the tests of the callers
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
        double* buf = tm.red + 128;   // clear of the two reduction buffers
        if ((tm.tid & 63) == 0) {
          buf[2 * (tm.tid >> 6)] = m1;
          buf[2 * (tm.tid >> 6) + 1] = m2;
        }
        __syncthreads();
        a1 = 0;
        a2 = 0;
        for (int w = 0; w < T / 64; ++w) {
          double b1 = buf[2 * w], b2 = buf[2 * w + 1];
          double hi = fmax(a1, b1), lo = fmax(fmin(a1, b1), fmax(a2, b2));
          a1 = hi;
          a2 = lo;
        }
        __syncthreads();
      };
      auto displacement_ok = [&]() {
        double a1, a2;
        top2(xb, a1, a2);
        maxd = fmax(maxd, a1);
        maxp = fmax(maxp, a1 + a2);
        return a1 + dr < h.skin_f && a1 + a2 + 2.0 * dr < h.skin_m;
      };
      // Active lists built at xr stay complete for evaluations at x and x + dr * axis while the same
      // condition holds with the active skins and displacements from xr.
      bool active_built = false;
      auto ensure_active = [&]() {
        if (active_built) {
          double a1, a2;
          top2(xr, a1, a2);
          if (a1 + dr < v.act_f && a1 + a2 + 2.0 * dr < v.act_m) return;
        }
        refresh_active<T, L>(A, v, tm);
        tm.lin(xr, n3, x, 1.0);
        active_built = true;
        ++refreshes;
      };

      // Trial end of the dimer for the rotor, placed by wave 0: cur = x + dr * d over the mobile atoms (the
      // arithmetic of place()).
      auto place_w = [&](double const* d) {
        for (int a = wv.lane; a < v.nm; a += 64) {
          double px = x[3 * a], py = x[3 * a + 1], pz = x[3 * a + 2];
          px += dr * d[3 * a];
          py += dr * d[3 * a + 1];
          pz += dr * d[3 * a + 2];
          v.cur[a] = make_double4(px, py, pz, 0.0);
        }
      };

      // Rotor::eff_gradient (rotor.cpp:49-164); returns false on a spline error.  Gradients run on the
      // whole workgroup, the rotor's algebra on wave 0 between them (two barriers per rotor iteration).
      // With qproj, bc[7] also receives g_eff . qproj for the trust update.
      auto eff_gradient = [&](double& mag_g, double const* qproj) -> bool {
        place(v, tm, x, 0.0, nullptr);
        ensure_active();
        err |= timed_gradient(g0);
        place(v, tm, x, dr, ax);
        err |= timed_gradient(g1);
        grads += 2;
        if (err) return false;
        rc.k = 0;
        double* theta = nullptr;
        double b_1 = 0, c_x0 = 0, theta_1 = 0;
        for (int i = 0;; ++i) {
          if (w0) {
            // The passes of lin(dg = g1 - g0), dot(dg, ax), lin(dgp), newton_step, dot(theta, ax),
            // lin(theta -= ot ax), dot(theta, theta), lin(theta *= 1 / |theta|) and dot2(dg, theta, dg, ax),
            // fused where one pass can finish one step and start the next (same expressions per element, so
            // the same bits); c0r is dot(dg, ax) again, which t already is.
            double t;
            {
              double acc = 0;
              for (int e = wv.lane; e < n3; e += 64) {
                double v = 1.0 * g1[e];
                v += -1.0 * g0[e];
                dg[e] = v;
                acc = __dadd_rn(acc, v * ax[e]);
              }
              t = Wave::total(acc);
            }
            wv.lin(dgp, n3, dg, 1.0, ax, -t);
            double ot;
            theta = rc.newton_step_fused(wv, ax, dgp, ax, &ot);
            double nn;
            {
              double acc = 0;
              for (int e = wv.lane; e < n3; e += 64) {
                double v = 1.0 * theta[e];
                v += -ot * ax[e];
                theta[e] = v;
                acc = __dadd_rn(acc, v * v);
              }
              nn = Wave::total(acc);
            }
            double b1r;
            double const c0r = t;
            {
              double const cn = 1.0 / sqrt(nn);
              double acc = 0;
              for (int e = wv.lane; e < n3; e += 64) {
                double const v = cn * theta[e];
                theta[e] = v;
                acc = __dadd_rn(acc, dg[e] * v);
              }
              b1r = Wave::total(acc);
            }
            b_1 = b1r / dr;
            c_x0 = c0r / dr;
            theta_1 = -0.5 * atan(b_1 / fabs(c_x0));
            bool const stop =
                fabs(theta_1) < o.theta_tol || (i >= o.iter_max_rot && c_x0 < 0) || i > 10 * o.iter_max_rot;
            if (!stop) {
              wv.lin(axp, n3, ax, cos(theta_1), theta, sin(theta_1));
              place_w(axp);
            }
            if (wv.lane == 0) {
              bc[0] = stop ? 1.0 : 0.0;
              bc[1] = c_x0;
            }
          }
          __syncthreads();
          ++rot_iters;
          if (bc[0] != 0.0) {
            curv = bc[1];
            break;
          }
          err |= timed_gradient(g1p);
          ++grads;
          if (err) return false;
          if (w0) {
            double c_x1;
            {   // lin(dg = g1p - g0) and dot(dg, axp) in one pass
              double acc = 0;
              for (int e = wv.lane; e < n3; e += 64) {
                double v = 1.0 * g1p[e];
                v += -1.0 * g0[e];
                dg[e] = v;
                acc = __dadd_rn(acc, v * axp[e]);
              }
              c_x1 = Wave::total(acc) / dr;
            }
            double a_1 = (c_x0 - c_x1 + b_1 * sin(2 * theta_1)) / (1 - cos(2 * theta_1));
            double theta_min = 0.5 * atan(b_1 / a_1);
            if (a_1 * cos(2 * theta_min) - a_1 + b_1 * sin(2 * theta_min) > 0) theta_min += M_PI / 2;
            double cm = cos(theta_min), sm = sin(theta_min);
            double w1 = sin(theta_1 - theta_min) / sin(theta_1);
            double w2 = sin(theta_min) / sin(theta_1);
            double w3 = 1 - cos(theta_min) - sin(theta_min) * tan(0.5 * theta_1);
            bool const stop = fabs(theta_min) < o.theta_tol;
            double c = 0;
            {   // lin(ax), lin(g1) and, when stopping, lin(dg = g1 - g0) with dot(dg, ax), in one pass
              double acc = 0;
              for (int e = wv.lane; e < n3; e += 64) {
                double va = cm * ax[e];
                va += sm * theta[e];
                ax[e] = va;
                double vg = w1 * g1[e];
                vg += w2 * g1p[e];
                vg += w3 * g0[e];
                g1[e] = vg;
                if (stop) {
                  double vd = 1.0 * vg;
                  vd += -1.0 * g0[e];
                  dg[e] = vd;
                  acc = __dadd_rn(acc, vd * va);
                }
              }
              if (stop) c = Wave::total(acc) / dr;
            }
            if (wv.lane == 0) {
              bc[2] = stop ? 1.0 : 0.0;
              bc[3] = c;
            }
          }
          __syncthreads();
          if (bc[2] != 0.0) {
            curv = bc[3];
            break;
          }
        }
        if (w0) {
          double p = wv.dot(g0, ax, n3);
          // lin(eff), dot(eff, eff) and dot(eff, qproj) in one pass (the same expressions per element).
          bool const only_ax = !o.relax_in_convex && curv > 0;
          double const cax = only_ax ? -p : -2.0 * p;
          double amg = 0, apj = 0;
          for (int e = wv.lane; e < n3; e += 64) {
            double vv;
            if (only_ax) {
              vv = cax * ax[e];
            } else {
              vv = 1.0 * g0[e];
              vv += cax * ax[e];
            }
            eff[e] = vv;
            amg = __dadd_rn(amg, vv * vv);
            if (qproj) apj = __dadd_rn(apj, vv * qproj[e]);
          }
          double mg = Wave::total(amg);
          double pj = qproj ? Wave::total(apj) : 0.0;
          if (wv.lane == 0) {
            bc[4] = mg;
            bc[7] = pj;
          }
        }
        __syncthreads();
        mag_g = bc[4];
        return true;
      };

      // Dimer::find_sp (dimer.cpp:28-171).  A resumed search enters the loop at its saved step, right
      // after that step's translation (the displacement check and the effective gradient come next).
      double mag_g = 0;
      double trust = h.resume ? h.r_trust : o.min_trust;
      int convex_count = h.resume ? h.r_convex : 0;
      bool resumed = h.resume != 0;
      if (!resumed && !displacement_ok()) {
        ret = EXIT_REGION;
        steps = 0;
      } else if (!resumed && !eff_gradient(mag_g, nullptr)) {
        ret = EXIT_OOB;
        steps = 0;
      } else {
        for (int i = resumed ? h.r_step : 0; i < o.max_steps; ++i) {
          double* q = resumed ? tc.q : nullptr;
          if (!resumed) {
            if (mag_g < o.f2norm * o.f2norm) {
              ret = 0;
              steps = i;
              break;
            } else if (convex_count >= o.convex_max) {
              ret = 1;
              steps = i;
              break;
            }
            if (w0) {
              // History test every hist_check_freq steps (dimer.cpp:102-131): the angle between x - min and
              // sp_k - min below theta_tol for a known saddle k ends the search as a collision.  Frozen atoms
              // sit at the minimum, so x - min is zero there and the dot product runs over the mobile atoms.
              bool hit = false;
              if (o.use_history && h.n_hist > 0 && i % o.hist_check_freq == 0) {
                wv.lin(dg, n3, x, 1.0, xb, -1.0);
                double const on = sqrt(wv.dot(dg, dg, n3));
                for (int k = 0; k < h.n_hist && !hit; ++k) {
                  double const dt = wv.dot(dg, A.hist + h.hist_off + static_cast<long>(k) * n3, n3);
                  hit = dt / (A.hnorm[h.hn_off + k] * on) > h.cos_tol;
                }
              }
              if (!hit) {
                double qq;
                q = tc.newton_step_fused(wv, x, eff, tc.q, &qq);   // dot(q, q) in its last pass
                double nrm = sqrt(qq);
                double scale = fmin(1.0, trust / nrm);
                for (int e = wv.lane; e < n3; e += 64) {   // lin(q *= scale) and lin(x -= q) in one pass
                  double const vq = scale * q[e];
                  q[e] = vq;
                  double vx = 1.0 * x[e];
                  vx += -1.0 * vq;
                  x[e] = vx;
                }
              }
              if (wv.lane == 0) bc[5] = hit ? 1.0 : 0.0;
            }
            __syncthreads();
            if (bc[5] != 0.0) {
              ret = 3;
              steps = i;
              break;
            }
          }   // !resumed
          resumed = false;
          if (!displacement_ok()) {
            ret = EXIT_REGION;
            steps = i;
            save_state();
            break;
          }
          if (!eff_gradient(mag_g, q)) {
            ret = EXIT_OOB;
            steps = i;
            break;
          }
          convex_count = curv > 0 ? convex_count + 1 : 0;
          double const proj = bc[7];
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
        double e_sp = region_energy<T, L, MULTI>(A, v, tm, false, e2);
        barrier = e_sp - e_base;
        if (e2) ret = EXIT_OOB;
      }
      // Master::do_adj_min + saddle_2_mech (find.cpp:757-837, 1014-1090): relax from the saddle nudged
      // by +-nudge_frac |sp - min| along the axis (LBFGS::minimise, lbfgs.cpp:29-115), call the minimum
      // nearer the initial one rev, remove the centroid drift of rev, and check the pathway.
      int mech_code = -1, min_iters = 0;
      double mech_barrier = 0, mech_delta = 0, drift[3] = {0, 0, 0};
      double* xsp = axp;    // free after the search
      double* xfw = g1p;
      if (o.mech && ret == 0) {
        auto minimise = [&](int& iters) -> int {   // 0 converged, 1 iter_max, EXIT_REGION, EXIT_OOB
          tcm.k = 0;
          place(v, tm, x, 0.0, nullptr);
          if (!displacement_ok()) return EXIT_REGION;
          ensure_active();
          err |= timed_gradient(g0);
          ++grads;
          if (err) return EXIT_OOB;
          double mag = tm.dot(g0, g0, n3);
          double trust = o.min_min_trust;
          for (int i = 0; i < o.min_iter_max; ++i) {
            iters = i;
            if (mag < o.min_f2norm * o.min_f2norm) return 0;
            double* q = tcm.newton_step(tm, x, g0);
            double nrm = sqrt(tm.dot(q, q, n3));
            double scale = fmin(1.0, trust / nrm);
            tm.lin(q, n3, q, scale);
            tm.lin(x, n3, x, 1.0, q, -1.0);
            __syncthreads();
            if (!displacement_ok()) return EXIT_REGION;
            place(v, tm, x, 0.0, nullptr);
            ensure_active();
            err |= timed_gradient(g0);
            ++grads;
            if (err) return EXIT_OOB;
            double proj;
            tm.dot2(g0, q, g0, g0, n3, proj, mag);
            if (proj < -o.min_proj_tol) {
              trust = fmax(o.min_min_trust, o.min_shrink_trust * trust);
            } else if (proj > o.min_proj_tol) {
              trust = fmin(o.min_max_trust, o.min_grow_trust * trust);
            }
          }
          iters = o.min_iter_max;
          return 1;
        };
        tm.lin(xsp, n3, x, 1.0);
        tm.lin(dg, n3, x, 1.0, xb, -1.0);
        double disp = sqrt(tm.dot(dg, dg, n3));
        int it_f = 0, it_r = 0;
        tm.lin(x, n3, xsp, 1.0, ax, disp * o.nudge_frac);
        int rf = minimise(it_f);
        tm.lin(xfw, n3, x, 1.0);
        int rr = rf == 0 ? (tm.lin(x, n3, xsp, 1.0, ax, -disp * o.nudge_frac), minimise(it_r)) : rf;
        min_iters = it_f + it_r;
        if (rf == EXIT_REGION || rr == EXIT_REGION) {
          ret = EXIT_REGION;   // re-run with larger skins (the minima lie outside the rows)
        } else if (rf == EXIT_OOB || rr == EXIT_OOB) {
          ret = EXIT_OOB;
        } else if (rf != 0 || rr != 0) {
          mech_code = 1;       // a minimisation failed
        } else {
          double* xrv = x;
          double df, dr2;
          tm.lin(dg, n3, xfw, 1.0, xb, -1.0);
          tm.lin(dgp, n3, xrv, 1.0, xb, -1.0);
          tm.dot2(dg, dg, dgp, dgp, n3, df, dr2);
          if (dr2 > df) {      // forward went to the initial minimum: swap
            double* t = xrv;
            xrv = xfw;
            xfw = t;
          }
          // Drift of the centroid over all natoms (frozen atoms do not move).
          double sdx = 0, sdy = 0, sdz = 0;
          {
            double vv[Team<T>::NQ][4] = {};
            for (int i0 = 0; i0 < v.nm; i0 += Team<T>::VS)
#pragma unroll
              for (int q = 0; q < Team<T>::NQ; ++q) {
                int a = i0 + tm.tid + q * T;
                if (a < v.nm) {
                  vv[q][0] = __dadd_rn(vv[q][0], xrv[3 * a] - xb[3 * a]);
                  vv[q][1] = __dadd_rn(vv[q][1], xrv[3 * a + 1] - xb[3 * a + 1]);
                  vv[q][2] = __dadd_rn(vv[q][2], xrv[3 * a + 2] - xb[3 * a + 2]);
                }
              }
            double o3[3];
            tm.template sum<3>(o3, vv);
            sdx = o3[0];
            sdy = o3[1];
            sdz = o3[2];
          }
          double const Nd = static_cast<double>(h.natoms);
          drift[0] = sdx / Nd;
          drift[1] = sdy / Nd;
          drift[2] = sdz / Nd;
          // |rev - drift - in| over all atoms: mobile terms + the frozen atoms shifted by -drift.
          for (int i = tm.tid; i < n3; i += T) {
            int const c = i % 3;
            dg[i] = xrv[i] - (c == 0 ? drift[0] : (c == 1 ? drift[1] : drift[2])) - xb[i];
          }
          double i2r2 = tm.dot(dg, dg, n3) + (Nd - v.nm) * (drift[0] * drift[0] + drift[1] * drift[1] + drift[2] * drift[2]);
          double r2f2, r2w2, w2f2;
          tm.lin(dg, n3, xfw, 1.0, xrv, -1.0);
          tm.lin(dgp, n3, xrv, 1.0, xsp, -1.0);
          tm.dot2(dg, dg, dgp, dgp, n3, r2f2, r2w2);
          tm.lin(dgp, n3, xfw, 1.0, xsp, -1.0);
          w2f2 = tm.dot(dgp, dgp, n3);
          // Atom that moves most from rev to fwd (first of equals in atom order) must be the centre.
          double best = 0;
          int arg = 0;
          for (int a = tm.tid; a < v.nm; a += T) {
            double ddx = xfw[3 * a] - xrv[3 * a], ddy = xfw[3 * a + 1] - xrv[3 * a + 1], ddz = xfw[3 * a + 2] - xrv[3 * a + 2];
            double r = ddx * ddx + ddy * ddy + ddz * ddz;
            if (r > best) {
              best = r;
              arg = a;
            }
          }
          for (int off = 32; off > 0; off >>= 1) {
            double ob = __shfl_xor(best, off, 64);
            int oa = __shfl_xor(arg, off, 64);
            if (ob > best || (ob == best && oa < arg)) {
              best = ob;
              arg = oa;
            }
          }
          __shared__ double mb[16];
          __shared__ int ma[16];
          if ((tm.tid & 63) == 0) {
            mb[tm.tid >> 6] = best;
            ma[tm.tid >> 6] = arg;
          }
          __syncthreads();
          best = mb[0];
          arg = ma[0];
          for (int w = 1; w < T / 64; ++w)
            if (mb[w] > best || (mb[w] == best && ma[w] < arg)) {
              best = mb[w];
              arg = ma[w];
            }
          __syncthreads();
          // Energies (local FP64: the frozen terms cancel in the differences).
          unsigned e3 = 0;
          place(v, tm, xrv, 0.0, nullptr);
          double e_rev = region_energy<T, L, MULTI>(A, v, tm, false, e3);
          place(v, tm, xfw, 0.0, nullptr);
          double e_fwd = region_energy<T, L, MULTI>(A, v, tm, false, e3);
          double e_sp = barrier + e_base;
          mech_barrier = e_sp - e_rev;
          mech_delta = e_fwd - e_rev;
          double const bt = o.basin_tol, st = o.stationary_tol;
          if (e3) {
            ret = EXIT_OOB;
          } else if (i2r2 > bt * bt) {
            mech_code = 2;     // starts away from the initial minimum: unlinked
          } else if (r2f2 < bt * bt) {
            mech_code = 3;     // converged back to the initial minimum
          } else if (r2w2 < st * st) {
            mech_code = 4;     // min -> sp displacement below the stationary tolerance
          } else if (w2f2 < st * st) {
            mech_code = 5;     // sp -> min displacement below the stationary tolerance
          } else if (arg != h.centre_local) {
            mech_code = 6;     // the mechanism is centred elsewhere
          } else if (!(e_sp > e_rev)) {
            mech_code = 7;     // saddle not above its minimum
          } else {
            mech_code = 0;
          }
          for (int i = tm.tid; i < n3; i += T) {
            A.out_rev[h.v3_off + i] = xrv[i];
            A.out_fwd[h.v3_off + i] = xfw[i];
          }
        }
        tm.lin(x, n3, xsp, 1.0);   // the saddle goes out as the search's positions
        __syncthreads();
      }
      for (int i = tm.tid; i < n3; i += T) {
        A.out_x[h.v3_off + i] = x[i];
        A.out_ax[h.v3_off + i] = ax[i];
      }
      if (tm.tid == 0) {
        SearchOut r;
        r.mech_code = mech_code;
        r.min_iters = min_iters;
        r.mech_barrier = mech_barrier;
        r.mech_delta = mech_delta;
        r.drift[0] = drift[0];
        r.drift[1] = drift[1];
        r.drift[2] = drift[2];
        r.exit_code = ret;
        r.steps = steps;
        r.rot_iters = rot_iters;
        r.refreshes = refreshes;
        r.grads = grads;
        r.curv = curv;
        r.barrier = barrier;
        r.e_base = e_base;
        r.max_disp = maxd;
        r.max_pair = maxp;
        r.ticks = static_cast<long>(wall_clock64()) - t_begin;
        r.grad_ticks = gticks;
        r.saved = saved;
        r.lbfgs_k = tc.k;
        r.convex = convex_count;
        r.trust = trust;
        A.out[sid] = r;
      }
      // Results (host-visible memory) first, then the flag: the host may read a search as soon as its
      // flag is set, while the kernel still runs other searches.
      __threadfence_system();
      __syncthreads();
      if (tm.tid == 0) {
        __hip_atomic_store(&A.done[sid], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      __syncthreads();
    }

    template <int T, int L, bool MIX, bool MULTI>
    __global__ __launch_bounds__(T) void k_search(BatchArgs A) {
      __shared__ double red[RED_LDS];
      __shared__ double lds_core[128];
      __shared__ int s_id;
      __shared__ double3 lds_S[27];
      // One LDS area for a search's region positions (when its rows fit) and, after them, its workspace.
      __shared__ double4 lds_area[lds_rows<T>() + lds_ws<T>() / 4];
      __shared__ SplineSet lds_spl;
      __shared__ unsigned lds_err;
      __shared__ double lds_bc[8];
      Team<T> tm{red, static_cast<int>(threadIdx.x)};
      if (threadIdx.x == 0) {   // constant indices only
#pragma unroll
        for (int i = 0; i < MAX_TYPES; ++i) lds_spl.F[i] = A.T.F[i];
#pragma unroll
        for (int i = 0; i < MAX_TYPES * MAX_TYPES; ++i) {
          lds_spl.phi[i] = A.T.phi[i];
          lds_spl.V[i] = A.T.V[i];
        }
      }
      __syncthreads();
      while (true) {
        if (threadIdx.x == 0) s_id = static_cast<int>(atomicAdd(A.queue, 1u));
        __syncthreads();
        int sid = s_id;
        __syncthreads();
        if (sid >= A.nsearch) break;
        run_search<T, L, MIX, MULTI>(A, sid, tm, lds_core, lds_S, lds_area, &lds_spl, &lds_err, lds_bc);
      }
    }

  }  // namespace sk
}  // namespace flyg

namespace flyg {
  namespace sk {
    // Kernel entry of k_search<T, L, MIX, MULTI> (a host-side launch handle), defined by the
    // translation unit that instantiates that variant (search_kernel_<T>_<precision>.hip); null for an
    // unsupported shape.
    void const* search_kernel(int T, int L, bool mix, bool multi);
#define FLYG_SK_DECL(T, P) void const* search_kernel_##T##_##P(int L, bool multi);
    FLYG_SK_DECL(256, fp64)
    FLYG_SK_DECL(256, mixed)
    FLYG_SK_DECL(512, fp64)
    FLYG_SK_DECL(512, mixed)
    FLYG_SK_DECL(1024, fp64)
    FLYG_SK_DECL(1024, mixed)
#undef FLYG_SK_DECL
  }  // namespace sk
}  // namespace flyg
