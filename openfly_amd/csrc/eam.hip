// EAM / Finnis-Sinclair energy, gradient and Hessian-vector product on gfx950
// (replaces potential::EAM::energy / gradient / hessian, eam.cpp:33-233).
//
// Work mapping: L lanes of a wavefront cooperate on one atom's neighbour row (ELL layout, 32-bit
// packed entries), each lane striding the row by L; partial sums are combined with cross-lane
// shuffles in a fixed order, so results are bitwise reproducible run to run.  No atomics touch
// floating-point data.  Positions are double4 (x, y, z, w): a neighbour gather is one 32-byte load,
// and pass 1 writes (x, y, z, F'(rho)) so pass 2 gets the neighbour's position and F' in that load (in
// the mixed form w holds F' as FP32 and the species, see pack_fp_type).
//
//   pass 1 (density):  rho_b = sum_a phi_{t_a->t_b}(r_ab);  F'_b = F_{t_b}'(rho_b)     (eam.cpp:62-86)
//   pass 2 (force):    dE/dr_g = -sum_a [V'(r) + F'_g phi'_{a->g}(r) + F'_a phi'_{g->a}(r)] (r_a - r_g)/r
//                      frozen atoms get exactly zero                                   (eam.cpp:90-120)
//   energy:            E = 1/2 sum_a sum_b V(r_ab) + sum_a F(rho_a), frozen included   (eam.cpp:33-51)
//   HVP (matrix-free directional derivative of the gradient along v, v masked to non-frozen atoms):
//      pass 1: drho_b = sum_a phi'_{a->b}(r) u.(v_a - v_b);  dF'_b = F''(rho_b) drho_b
//      pass 2: (Hv)_g = -sum_a [dm u + m (dv - u s)/r],  s = u.dv,  dv = v_a - v_g, u = (r_a - r_g)/r,
//              dm = (V'' + F'_g phi''_{a->g} + F'_a phi''_{g->a}) s + dF'_g phi'_{a->g} + dF'_a phi'_{g->a}
//      which is H v for the analytic Hessian of eam.cpp:122-233 (pinned against it in the tests).
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>

#include "eam_host.hpp"

namespace flyg {

  namespace {

    constexpr int NT = 256;

    // Image shift of a packed entry: sum_a s_a * basis column a, s_a in {-1, 0, 1}.  Exact for
    // orthorhombic boxes (each component is 0 or +-L), computed in registers (no table lookup).
    struct Shifts {
      double b[3][3];   // b[row][col]
      __device__ double3 of(unsigned code) const {
        int s0 = static_cast<int>(code % 3u) - 1;
        int s1 = static_cast<int>((code / 3u) % 3u) - 1;
        int s2 = static_cast<int>(code / 9u) - 1;
        return make_double3(s0 * b[0][0] + s1 * b[0][1] + s2 * b[0][2], s0 * b[1][0] + s1 * b[1][1] + s2 * b[1][2],
                            s0 * b[2][0] + s1 * b[2][1] + s2 * b[2][2]);
      }
    };

    __device__ inline int vidx(EAMTables const& T, int a, int b) { return a > b ? a + T.ntypes * b : b + T.ntypes * a; }

    // Mixed-form pass-1 output: the force pass uses F' only in FP32, so the w slot of (x, y, z, w) carries
    // F' as a float in its low 32 bits and the species in its high 32 bits, and a neighbour's species
    // arrives with its position (one gather less per pair).
    __device__ inline double pack_fp_type(float fp, int t) {
      return __longlong_as_double((static_cast<long long>(t) << 32) |
                                  static_cast<long long>(static_cast<unsigned>(__float_as_int(fp))));
    }
    __device__ inline float packed_fp(double w) {
      return __int_as_float(static_cast<int>(static_cast<unsigned>(__double_as_longlong(w) & 0xffffffffll)));
    }
    __device__ inline int packed_type(double w) { return static_cast<int>(__double_as_longlong(w) >> 32); }

    // One neighbour entry: image position difference and squared distance.
    struct PairGeo {
      double dx, dy, dz, r2, w;
      unsigned j;
    };

    __device__ inline PairGeo geo(Shifts const& S, double4 const* __restrict__ pos, unsigned en, double4 const& pi) {
      PairGeo p;
      p.j = en & NB_IDX_MASK;
      unsigned const code = en >> NB_SHIFT_BITS;
      double4 pj = pos[p.j];
      if (code != 13u) {   // image of a periodic neighbour (rows of interior atoms never take this branch)
        double3 sh = S.of(code);
        pj.x += sh.x;
        pj.y += sh.y;
        pj.z += sh.z;
      }
      p.dx = pj.x - pi.x;
      p.dy = pj.y - pi.y;
      p.dz = pj.z - pi.z;
      p.r2 = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz;
      p.w = pj.w;
      return p;
    }

    // Walks lane `lane`'s share of a row (entries k = lane + m L): NS = 16 / L entries per step, entry m
    // of the step going to row-sum slot m (compile time), so every lane-count gives the same sums; for
    // L = 16 two entries per step share slot 0 in order.  All loads of a step are issued first.
    template <int L, typename F>
    __device__ inline void walk_row(unsigned const* __restrict__ row, int c, int lane, Shifts const& S,
                                    double4 const* __restrict__ pos, double4 const& pi, F&& one) {
      constexpr int NS = 16 / L;
      constexpr int U = NS == 1 ? 2 : NS;
      int m = 0;
      for (; lane + (m + U - 1) * L < c; m += U) {
        PairGeo p[U];
#pragma unroll
        for (int q = 0; q < U; ++q) p[q] = geo(S, pos, row[lane + (m + q) * L], pi);
#pragma unroll
        for (int q = 0; q < U; ++q) one(q % NS, p[q]);
      }
#pragma unroll
      for (int q = 0; q < U - 1; ++q)
        if (lane + (m + q) * L < c) one(q % NS, geo(S, pos, row[lane + (m + q) * L], pi));
    }

    // Density pass.  MIX (and not ENERGY): r and the density spline in FP32, sums in FP64.
    template <int L, bool MULTI, bool ENERGY, bool MIX>
    __global__ __launch_bounds__(NT) void k_density(EAMTables T, Shifts S, PairArgs A, double4* __restrict__ posf,
                                                    double* __restrict__ eatom, unsigned* __restrict__ err) {
      __shared__ SplineDesc s_phi[MAX_TYPES * MAX_TYPES], s_V[MAX_TYPES * MAX_TYPES];
      if (MULTI) {   // per-lane species pairs: table descriptors from LDS, not from the argument block
        int t = threadIdx.x;
        if (t < T.ntypes * T.ntypes) {
          s_phi[t] = T.phi[t];
          s_V[t] = T.V[t];
        }
        __syncthreads();
      }
      auto PHI = [&](int k) -> SplineDesc const& {
        if constexpr (MULTI) {
          return s_phi[k];
        } else {
          return T.phi[k];
        }
      };
      auto VV = [&](int k) -> SplineDesc const& {
        if constexpr (MULTI) {
          return s_V[k];
        } else {
          return T.V[k];
        }
      };
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(NT) + threadIdx.x;
      long i = gid / L;
      int lane = static_cast<int>(gid % L);
      bool act = i < A.n;
      Acc16<L> rho, vs;
      unsigned e = 0;
      double4 pi = make_double4(0, 0, 0, 0);
      int ti = 0;
      if (act) {
        pi = A.pos[i];
        ti = MULTI ? checked_type(A.types[i], T.ntypes, e) : A.t0;
        int const c = min(A.cnt[i], A.K);
        unsigned const* row = A.nbr + i * static_cast<long>(A.K);
        walk_row<L>(row, c, lane, S, A.pos, pi, [&](int q, PairGeo const& p) {
          if (p.r2 < T.rc2) {
            int tj = MULTI ? checked_type(A.type_in_w ? static_cast<uint32_t>(p.w) : A.types[p.j], T.ntypes, e) : A.t0;
            if (MIX && !ENERGY) {
              float r2f = static_cast<float>(p.r2);
              float rf = r2f * __frsqrt_rn(r2f);
              rho.add_slot(q, static_cast<double>(spl32f<0>(T, PHI(tj + T.ntypes * ti), rf, e)));
            } else {
              double r = sqrt(p.r2);
              rho.add_slot(q, spl64<0>(T, PHI(tj + T.ntypes * ti), r, e));
              if (ENERGY) vs.add_slot(q, spl64<0>(T, VV(vidx(T, ti, tj)), r, e));
            }
          }
        });
      }
      double rt = rho.total();
      double vt = ENERGY ? vs.total() : 0.0;
      if (act && lane == 0) {
        double f, fp;
        spl64_ffp(T, T.F[ti], rt, f, fp, e);
        double const w = MIX && !ENERGY ? pack_fp_type(static_cast<float>(fp), ti) : fp;
        posf[i] = make_double4(pi.x, pi.y, pi.z, w);
        if (ENERGY) eatom[i] = 0.5 * vt + f;
      }
      if (e) atomicOr(err, e);
    }

    template <int L, bool MULTI, bool MIX>
    __global__ __launch_bounds__(NT) void k_force(EAMTables T, Shifts S, PairArgs A, double4 const* __restrict__ posf,
                                                  double* __restrict__ grad, unsigned* __restrict__ err) {
      __shared__ SplineDesc s_phi[MAX_TYPES * MAX_TYPES], s_V[MAX_TYPES * MAX_TYPES];
      if (MULTI) {   // per-lane species pairs: table descriptors from LDS, not from the argument block
        int t = threadIdx.x;
        if (t < T.ntypes * T.ntypes) {
          s_phi[t] = T.phi[t];
          s_V[t] = T.V[t];
        }
        __syncthreads();
      }
      auto PHI = [&](int k) -> SplineDesc const& {
        if constexpr (MULTI) {
          return s_phi[k];
        } else {
          return T.phi[k];
        }
      };
      auto VV = [&](int k) -> SplineDesc const& {
        if constexpr (MULTI) {
          return s_V[k];
        } else {
          return T.V[k];
        }
      };
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(NT) + threadIdx.x;
      long i = gid / L;   // row: the atom, or its slot when A.order is set
      int lane = static_cast<int>(gid % L);
      bool inr = i < A.n;
      long const ia = inr && A.order ? A.order[i] : i;   // the atom, for its frozen flag and gradient
      unsigned e = 0;
      // The row's own data is loaded beside the frozen flag (not after it): the flag waits on `order`.
      double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
      int ti = MULTI ? 0 : A.t0;
      int c = 0;
      if (inr) {
        pi = posf[i];
        c = min(A.cnt[i], A.K);
        // Mixed form: the species the density pass packed beside F' (checked there).
        if (MULTI) ti = MIX ? packed_type(pi.w) : checked_type(A.types[i], T.ntypes, e);
      }
      bool act = inr && !A.frozen[ia];
      Acc16<L> gx, gy, gz;
      if (act) {
        unsigned const* row = A.nbr + i * static_cast<long>(A.K);
        walk_row<L>(row, c, lane, S, posf, pi, [&](int q, PairGeo const& p) {
          if (p.r2 < T.rc2) {
            int tj = MULTI ? (MIX ? packed_type(p.w) : checked_type(A.types[p.j], T.ntypes, e)) : A.t0;
            double cf;
            if (MIX) {
              float r2f = static_cast<float>(p.r2);
              float rinv = __frsqrt_rn(r2f);
              float rf = r2f * rinv;
              float vp = spl32f<1>(T, VV(vidx(T, ti, tj)), rf, e);
              float pa = spl32f<1>(T, PHI(tj + T.ntypes * ti), rf, e);
              // phi_{g->a} is phi_{a->g} for a same-species pair: no second table load.
              float pb = MULTI && ti != tj ? spl32f<1>(T, PHI(ti + T.ntypes * tj), rf, e) : pa;
              float mag = vp + packed_fp(pi.w) * pa + packed_fp(p.w) * pb;
              cf = static_cast<double>(mag * rinv);
            } else {
              double r = sqrt(p.r2);
              double vp = spl64<1>(T, VV(vidx(T, ti, tj)), r, e);
              double pa = spl64<1>(T, PHI(tj + T.ntypes * ti), r, e);
              double pb = MULTI && ti != tj ? spl64<1>(T, PHI(ti + T.ntypes * tj), r, e) : pa;
              cf = (vp + pi.w * pa + p.w * pb) / r;
            }
            gx.add_slot(q, -cf * p.dx);
            gy.add_slot(q, -cf * p.dy);
            gz.add_slot(q, -cf * p.dz);
          }
        });
      }
      double tx = gx.total(), ty = gy.total(), tz = gz.total();
      if (L >= 3 && inr && lane < 3) {
        double v = lane == 0 ? tx : (lane == 1 ? ty : tz);
        grad[3 * ia + lane] = act ? v : 0.0;
      }
      if (L < 3 && inr && lane == 0) {
        grad[3 * ia] = act ? tx : 0.0;
        grad[3 * ia + 1] = act ? ty : 0.0;
        grad[3 * ia + 2] = act ? tz : 0.0;
      }
      if (e) atomicOr(err, e);
    }

    template <int L, bool MULTI>
    __global__ __launch_bounds__(NT) void k_hvp_density(EAMTables T, Shifts S, PairArgs A, double const* __restrict__ v,
                                                        double2* __restrict__ dfp, unsigned* __restrict__ err) {
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(NT) + threadIdx.x;
      long i = gid / L;
      int lane = static_cast<int>(gid % L);
      bool act = i < A.n;
      double rho = 0, drho = 0;
      unsigned e = 0;
      int ti = 0;
      if (act) {
        double4 pi = A.pos[i];
        ti = MULTI ? checked_type(A.types[i], T.ntypes, e) : A.t0;
        bool fi = A.frozen[i];
        double vix = fi ? 0 : v[3 * i], viy = fi ? 0 : v[3 * i + 1], viz = fi ? 0 : v[3 * i + 2];
        int const c = min(A.cnt[i], A.K);
        unsigned const* row = A.nbr + i * static_cast<long>(A.K);
        for (int k = lane; k < c; k += L) {
          unsigned en = row[k];
          unsigned j = en & NB_IDX_MASK;
          double3 sh = S.of(en >> NB_SHIFT_BITS);
          double4 pj = A.pos[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < T.rc2) {
            double r = sqrt(r2);
            int tj = MULTI ? checked_type(A.types[j], T.ntypes, e) : A.t0;
            double f, fp;
            spl64_ffp(T, T.phi[tj + T.ntypes * ti], r, f, fp, e);
            rho += f;
            bool fj = A.frozen[j];
            double vjx = fj ? 0 : v[3 * j], vjy = fj ? 0 : v[3 * j + 1], vjz = fj ? 0 : v[3 * j + 2];
            double s = (dx * (vjx - vix) + dy * (vjy - viy) + dz * (vjz - viz)) / r;
            drho += fp * s;
          }
        }
      }
      rho = group_sum<L>(rho);
      drho = group_sum<L>(drho);
      if (act && lane == 0) {
        double fp = spl64<1>(T, T.F[ti], rho, e);
        double fpp = spl64<2>(T, T.F[ti], rho, e);
        dfp[i] = make_double2(fp, fpp * drho);
      }
      if (e) atomicOr(err, e);
    }

    template <int L, bool MULTI>
    __global__ __launch_bounds__(NT) void k_hvp_force(EAMTables T, Shifts S, PairArgs A, double const* __restrict__ v,
                                                      double2 const* __restrict__ dfp, double* __restrict__ hv,
                                                      unsigned* __restrict__ err) {
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(NT) + threadIdx.x;
      long i = gid / L;
      int lane = static_cast<int>(gid % L);
      bool inr = i < A.n;
      bool act = inr && !A.frozen[i];
      double hx = 0, hy = 0, hz = 0;
      unsigned e = 0;
      if (act) {
        double4 pi = A.pos[i];
        double2 di = dfp[i];
        int ti = MULTI ? checked_type(A.types[i], T.ntypes, e) : A.t0;
        double vix = v[3 * i], viy = v[3 * i + 1], viz = v[3 * i + 2];
        int const c = min(A.cnt[i], A.K);
        unsigned const* row = A.nbr + i * static_cast<long>(A.K);
        for (int k = lane; k < c; k += L) {
          unsigned en = row[k];
          unsigned j = en & NB_IDX_MASK;
          double3 sh = S.of(en >> NB_SHIFT_BITS);
          double4 pj = A.pos[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < T.rc2) {
            double r = sqrt(r2);
            int tj = MULTI ? checked_type(A.types[j], T.ntypes, e) : A.t0;
            double2 dj = dfp[j];
            bool fj = A.frozen[j];
            double vjx = fj ? 0 : v[3 * j], vjy = fj ? 0 : v[3 * j + 1], vjz = fj ? 0 : v[3 * j + 2];
            double ux = dx / r, uy = dy / r, uz = dz / r;
            double dvx = vjx - vix, dvy = vjy - viy, dvz = vjz - viz;
            double s = ux * dvx + uy * dvy + uz * dvz;
            double vp = spl64<1>(T, T.V[vidx(T, ti, tj)], r, e);
            double vpp = spl64<2>(T, T.V[vidx(T, ti, tj)], r, e);
            double pa = spl64<1>(T, T.phi[tj + T.ntypes * ti], r, e);   // phi'_{a->g}
            double paa = spl64<2>(T, T.phi[tj + T.ntypes * ti], r, e);
            double pb = pa, pbb = paa;
            if (MULTI) {
              pb = spl64<1>(T, T.phi[ti + T.ntypes * tj], r, e);          // phi'_{g->a}
              pbb = spl64<2>(T, T.phi[ti + T.ntypes * tj], r, e);
            }
            double m = vp + di.x * pa + dj.x * pb;
            double dm = (vpp + di.x * paa + dj.x * pbb) * s + di.y * pa + dj.y * pb;
            double q = m / r;
            hx -= dm * ux + q * (dvx - ux * s);
            hy -= dm * uy + q * (dvy - uy * s);
            hz -= dm * uz + q * (dvz - uz * s);
          }
        }
      }
      hx = group_sum<L>(hx);
      hy = group_sum<L>(hy);
      hz = group_sum<L>(hz);
      if (inr && lane == 0) {
        hv[3 * i] = act ? hx : 0.0;
        hv[3 * i + 1] = act ? hy : 0.0;
        hv[3 * i + 2] = act ? hz : 0.0;
      }
      if (e) atomicOr(err, e);
    }

    // FP64 densities only (base state of batched searches).
    template <int L, bool MULTI>
    __global__ __launch_bounds__(NT) void k_rho(EAMTables T, Shifts S, PairArgs A, double* __restrict__ rho_out,
                                                unsigned* __restrict__ err) {
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(NT) + threadIdx.x;
      long i = gid / L;
      int lane = static_cast<int>(gid % L);
      bool act = i < A.n;
      double rho = 0;
      unsigned e = 0;
      if (act) {
        double4 pi = A.pos[i];
        int ti = MULTI ? checked_type(A.types[i], T.ntypes, e) : A.t0;
        int const c = min(A.cnt[i], A.K);
        unsigned const* row = A.nbr + i * static_cast<long>(A.K);
        for (int k = lane; k < c; k += L) {
          unsigned en = row[k];
          unsigned j = en & NB_IDX_MASK;
          double3 sh = S.of(en >> NB_SHIFT_BITS);
          double4 pj = A.pos[j];
          double dx = (pj.x + sh.x) - pi.x, dy = (pj.y + sh.y) - pi.y, dz = (pj.z + sh.z) - pi.z;
          double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < T.rc2) {
            int tj = MULTI ? checked_type(A.types[j], T.ntypes, e) : A.t0;
            rho += spl64<0>(T, T.phi[tj + T.ntypes * ti], sqrt(r2), e);
          }
        }
      }
      rho = group_sum<L>(rho);
      if (act && lane == 0) rho_out[i] = rho;
      if (e) atomicOr(err, e);
    }

    __global__ void k_table_eval(EAMTables T, SplineDesc s, double const* __restrict__ x, long m, int deriv,
                                 double* __restrict__ out, unsigned* __restrict__ err) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i >= m) return;
      unsigned e = 0;
      double v = deriv == 0 ? spl64<0>(T, s, x[i], e) : deriv == 1 ? spl64<1>(T, s, x[i], e) : spl64<2>(T, s, x[i], e);
      out[i] = v;
      if (e) atomicOr(err, e);
    }

    Shifts shifts_of(BoxDesc const& b) {
      Shifts s;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) s.b[r][c] = b.B[r][c];
      return s;
    }

    // t0 >= 0 selects the single-species kernels (constant table indices): always with one species
    // tabulated, and with several when the staged host types hold one species only.
    PairArgs pair_args(flyg_eam_s const& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen) {
      int t0 = E.tab.ntypes == 1 ? 0 : (types == E.types.get() ? E.staged_single : -1);
      return PairArgs{nl.pos.get(), nl.nbr.get(), nl.ncnt.get(), nl.K, nl.n, types, frozen, t0};
    }

    __global__ void k_types_by_slot(uint32_t const* __restrict__ types, int const* __restrict__ order, long n,
                                    uint32_t* __restrict__ out, double4* __restrict__ gpos) {
      long k = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (k < n) {
        uint32_t const t = types[order[k]];
        out[k] = t;
        gpos[k].w = static_cast<double>(t);
      }
    }

    // The gradient passes run in the list's cell-order layout: a workgroup's atoms are neighbours in
    // space and so are the rows they gather, which raises the L1/L2 hit rate of the gathers.  Rows hold
    // the same entries in the same order, so the sums (and the bits) are those of the atom layout.  With
    // several species the types are gathered into slot order, and into the w slot of the slot-ordered
    // positions (the density pass reads a neighbour's species from there): once per list build for staged
    // host types (the library sees every change of them), on every call for device types.  Type stamps
    // are unique across handles, so a list shared by several potentials or type arrays is re-gathered
    // whenever another one wrote its w slots last.
    PairArgs slot_args(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen) {
      PairArgs A = pair_args(E, nl, types, frozen);
      A.pos = nl.gpos.get();
      A.nbr = nl.grow.get();
      A.cnt = nl.gcnt.get();
      A.order = nl.order.get();
      if (A.t0 < 0) {
        bool const staged = types == E.types.get();
        bool const cached = staged && E.gtype_build == nl.build_id && E.gtype_stamp == E.types_stamp &&
                            nl.wtype_build == nl.build_id && nl.wtype_stamp == E.types_stamp;
        if (!cached) {
          E.gtype_build = -1;   // the buffer may be reallocated: its contents are no longer known
          nl.wtype_build = -1;
          E.gtype.reserve(nl.n);
          k_types_by_slot<<<static_cast<unsigned>((nl.n + 255) / 256), 256, 0, E.ctx->stream>>>(
              types, nl.order.get(), nl.n, E.gtype.get(), nl.gpos.get());
          HIPCHK(hipGetLastError());
          if (staged) {
            E.gtype_build = nl.build_id;
            E.gtype_stamp = E.types_stamp;
            nl.wtype_build = nl.build_id;
            nl.wtype_stamp = E.types_stamp;
          }
        }
        A.types = E.gtype.get();
        A.type_in_w = true;
      }
      return A;
    }

    // Lanes per atom for the full-system kernels (a row of ~60 entries split 8 ways keeps ~8 pairs
    // in flight per lane and the grid at >= 8 wavefronts per SIMD for >= 30k atoms).
    constexpr int LANES = 8;

    unsigned grid_for(long n, int L) { return static_cast<unsigned>((n * L + NT - 1) / NT); }

  }  // namespace

  void eam_setup_tables(flyg_eam_s& E) {
    SetflData const& D = E.data;
    int const n = D.ntypes;
    std::vector<double4> all;
    EAMTables T{};
    T.ntypes = n;
    T.rc = D.rcut;
    T.rc2 = D.rcut * D.rcut;
    T.rcf = static_cast<float>(D.rcut);
    auto add = [&](std::vector<double> const& y, double dx) {
      std::vector<double4> c = spline_coefficients(y, dx);
      SplineDesc s{};
      s.off = static_cast<int>(all.size());
      s.count = static_cast<int>(c.size());
      s.dx = dx;
      s.inv_dx = 1 / dx;
      s.dxf = static_cast<float>(dx);
      s.inv_dxf = static_cast<float>(1 / dx);
      all.insert(all.end(), c.begin(), c.end());
      return s;
    };
    for (int a = 0; a < n; ++a) T.F[a] = add(D.F[a], D.drho);
    for (int k = 0; k < n * n; ++k) T.phi[k] = add(D.phi[k], D.dr);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b <= a; ++b) T.V[a + n * b] = add(D.V[a + n * b], D.dr);
    std::vector<float4> all32(all.size());
    for (std::size_t i = 0; i < all.size(); ++i)
      all32[i] = make_float4(static_cast<float>(all[i].x), static_cast<float>(all[i].y), static_cast<float>(all[i].z),
                             static_cast<float>(all[i].w));
    E.c64.reserve(all.size());
    E.c32.reserve(all.size());
    HIPCHK(hipMemcpy(E.c64.get(), all.data(), all.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(E.c32.get(), all32.data(), all32.size() * sizeof(float4), hipMemcpyHostToDevice));
    T.c64 = E.c64.get();
    T.c32 = E.c32.get();
    E.tab = T;
  }

  namespace {
    // energy = true: FP64 densities and pair energies (energies are always FP64); otherwise the density
    // pass of a gradient in the given precision.
    template <int L>
    void launch_density(flyg_eam_s& E, Shifts const& S, PairArgs const& A, bool energy, int precision, double4* pf) {
      hipStream_t st = E.ctx->stream;
      unsigned* err = E.ctx->dflags.get();
      bool multi = A.t0 < 0;
      bool mix = precision == FLYG_MIXED;
      unsigned g = grid_for(A.n, L);
      if (multi) {
        if (energy)
          k_density<L, true, true, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, E.eatom.get(), err);
        else if (mix)
          k_density<L, true, false, true><<<g, NT, 0, st>>>(E.tab, S, A, pf, nullptr, err);
        else
          k_density<L, true, false, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, nullptr, err);
      } else {
        if (energy)
          k_density<L, false, true, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, E.eatom.get(), err);
        else if (mix)
          k_density<L, false, false, true><<<g, NT, 0, st>>>(E.tab, S, A, pf, nullptr, err);
        else
          k_density<L, false, false, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, nullptr, err);
      }
    }

    template <int L>
    void launch_force(flyg_eam_s& E, Shifts const& S, PairArgs const& A, double* grad, int precision,
                      double4 const* pf) {
      hipStream_t st = E.ctx->stream;
      unsigned* err = E.ctx->dflags.get();
      bool multi = A.t0 < 0;
      bool mix = precision == FLYG_MIXED;
      unsigned g = grid_for(A.n, L);
      if (multi) {
        if (mix)
          k_force<L, true, true><<<g, NT, 0, st>>>(E.tab, S, A, pf, grad, err);
        else
          k_force<L, true, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, grad, err);
      } else {
        if (mix)
          k_force<L, false, true><<<g, NT, 0, st>>>(E.tab, S, A, pf, grad, err);
        else
          k_force<L, false, false><<<g, NT, 0, st>>>(E.tab, S, A, pf, grad, err);
      }
    }
  }  // namespace

  double eam_energy(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen) {
    require(nl.built, "List used before rebuild");
    require(E.tab.rc <= nl.rc, "r_cut is bigger than the internal r_cut of the list");
    E.posf.reserve(nl.n);
    E.eatom.reserve(nl.n);
    launch_density<LANES>(E, shifts_of(nl.box), pair_args(E, nl, types, frozen), true, FLYG_FP64, E.posf.get());
    HIPCHK(hipGetLastError());
    dev_sum(*E.ctx, E.eatom.get(), nl.n, E.ctx->dscal.get());
    HIPCHK(hipMemcpyAsync(E.ctx->hscal, E.ctx->dscal.get(), sizeof(double), hipMemcpyDeviceToHost, E.ctx->stream));
    unsigned fl = E.ctx->take_flags();
    if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
    if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
    return E.ctx->hscal[0];
  }

  namespace {
    // Density then force pass; pf holds the pass-1 output (positions + F') of every atom.
    // Lanes per atom row of the two passes, packed as density | force << 8 (either count gives the same
    // sums, Acc16, so each pass takes its own).
    constexpr int pack_lanes(int ld, int lf) { return ld | (lf << 8); }

    void density_with(flyg_eam_s& E, Shifts const& S, PairArgs const& A, int precision, int L, double4* pf) {
      switch (L) {
        case 4:
          launch_density<4>(E, S, A, false, precision, pf);
          break;
        case 16:
          launch_density<16>(E, S, A, false, precision, pf);
          break;
        default:
          launch_density<8>(E, S, A, false, precision, pf);
      }
    }

    void force_with(flyg_eam_s& E, Shifts const& S, PairArgs const& A, double* d_grad, int precision, int L,
                    double4 const* pf) {
      switch (L) {
        case 4:
          launch_force<4>(E, S, A, d_grad, precision, pf);
          break;
        case 16:
          launch_force<16>(E, S, A, d_grad, precision, pf);
          break;
        default:
          launch_force<8>(E, S, A, d_grad, precision, pf);
      }
    }

    void gradient_with(flyg_eam_s& E, Shifts const& S, PairArgs const& A, double* d_grad, int precision, int lanes,
                       double4* pf) {
      density_with(E, S, A, precision, lanes & 0xff, pf);
      force_with(E, S, A, d_grad, precision, lanes >> 8, pf);
    }

    // Lanes per atom row of each pass: the row sums do not depend on it (Acc16), so each pass's count is
    // picked by timing that pass alone (the force pass reads the density pass's output of the last run).
    int tuned_lanes(flyg_eam_s& E, Shifts const& S, PairArgs const& A, double* d_grad, int precision) {
      if (E.lanes > 0) return pack_lanes(E.lanes, E.lanes);
      auto key = std::make_pair(A.n, precision);
      auto it = E.tuned.find(key);
      if (it != E.tuned.end()) return it->second;
      hipStream_t st = E.ctx->stream;
      hipEvent_t a, b;
      HIPCHK(hipEventCreate(&a));
      HIPCHK(hipEventCreate(&b));
      auto best_of = [&](auto&& run) {
        int best = 8;
        float best_ms = 1e30f;
        for (int L : {4, 8, 16}) {
          run(L);   // warm
          HIPCHK(hipEventRecord(a, st));
          for (int r = 0; r < 3; ++r) run(L);
          HIPCHK(hipEventRecord(b, st));
          HIPCHK(hipEventSynchronize(b));
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, a, b));
          if (ms < best_ms) {
            best_ms = ms;
            best = L;
          }
        }
        return best;
      };
      int const ld = best_of([&](int L) { density_with(E, S, A, precision, L, E.posf.get()); });
      density_with(E, S, A, precision, ld, E.posf.get());
      int const lf = best_of([&](int L) { force_with(E, S, A, d_grad, precision, L, E.posf.get()); });
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      int const lanes = pack_lanes(ld, lf);
      E.tuned[key] = lanes;
      return lanes;
    }
  }  // namespace

  void eam_gradient(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen, double* d_grad,
                    int precision) {
    require(nl.built, "List used before rebuild");
    require(E.tab.rc <= nl.rc, "r_cut is bigger than the internal r_cut of the list");
    E.posf.reserve(nl.n);
    Shifts S = shifts_of(nl.box);
    if (E.timing) {   // the pair covers the species gather too (and, on a first call, the lane tuning)
      if (E.events_used == E.events.size()) {
        hipEvent_t a, b;
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
        E.events.emplace_back(a, b);
      }
      HIPCHK(hipEventRecord(E.events[E.events_used].first, E.ctx->stream));
    }
    PairArgs A = slot_args(E, nl, types, frozen);
    int const lanes = tuned_lanes(E, S, A, d_grad, precision);
    gradient_with(E, S, A, d_grad, precision, lanes, E.posf.get());
    if (E.timing) HIPCHK(hipEventRecord(E.events[E.events_used++].second, E.ctx->stream));
    HIPCHK(hipGetLastError());
  }

  // ---------------------------------------------------------------------------------------------------
  // Allocation-free gradient.  The pass-1 output goes to a file-scope device array instead of the
  // handle's buffer; each GPU holds its own copy of the array (at its own address, looked up once per
  // device when a context is created there) until the process ends, and every call on that GPU shares it,
  // so calls that overlap on one GPU race.  The density pass writes every entry the force pass reads, so
  // nothing needs clearing between calls.
  // ---------------------------------------------------------------------------------------------------

  namespace {
    static __device__ __align__(256) double4 scratch[NOALLOC_MAX_ATOMS];

    constexpr int MAX_DEVICES = 64;
    std::mutex g_scratch_mu;
    double4* g_scratch_ptr[MAX_DEVICES] = {};
  }  // namespace

  void noalloc_scratch_init(int device) {
    require(device >= 0 && device < MAX_DEVICES, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    // The array's name is a host-side stand-in; kernels get this device address (the device is current).
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<double4*>(p);
  }

  void eam_gradient_noalloc(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* d_types, uint8_t const* d_frozen,
                            double* d_grad, int precision) {
    require(nl.built, "List used before rebuild");
    require(E.tab.rc <= nl.rc, "r_cut is bigger than the internal r_cut of the list");
    require(nl.n <= NOALLOC_MAX_ATOMS, "the allocation-free gradient takes at most 131072 atoms");
    double4* pf = nullptr;
    {
      std::lock_guard<std::mutex> lock(g_scratch_mu);
      int const dev = E.ctx->device;
      if (dev >= 0 && dev < MAX_DEVICES) pf = g_scratch_ptr[dev];
    }
    require(pf != nullptr, "no scratch address for this device (its context was not created)");
    Shifts S = shifts_of(nl.box);
    PairArgs A = pair_args(E, nl, d_types, d_frozen);
    // The lanes per atom row change no result bit: a fixed choice, or an earlier tuning, without timing.
    int lanes = E.lanes > 0 ? pack_lanes(E.lanes, E.lanes) : pack_lanes(8, 8);
    auto it = E.tuned.find(std::make_pair(A.n, precision));
    if (E.lanes <= 0 && it != E.tuned.end()) lanes = it->second;
    gradient_with(E, S, A, d_grad, precision, lanes, pf);
    HIPCHK(hipGetLastError());
  }

  void eam_hvp(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen, double const* d_v,
               double* d_hv) {
    require(nl.built, "List used before rebuild");
    require(E.tab.rc <= nl.rc, "r_cut is bigger than the internal r_cut of the list");
    E.dfp.reserve(nl.n);
    Shifts S = shifts_of(nl.box);
    PairArgs A = pair_args(E, nl, types, frozen);
    hipStream_t st = E.ctx->stream;
    unsigned* err = E.ctx->dflags.get();
    unsigned g = grid_for(nl.n, LANES);
    if (A.t0 < 0) {
      k_hvp_density<LANES, true><<<g, NT, 0, st>>>(E.tab, S, A, d_v, E.dfp.get(), err);
      k_hvp_force<LANES, true><<<g, NT, 0, st>>>(E.tab, S, A, d_v, E.dfp.get(), d_hv, err);
    } else {
      k_hvp_density<LANES, false><<<g, NT, 0, st>>>(E.tab, S, A, d_v, E.dfp.get(), err);
      k_hvp_force<LANES, false><<<g, NT, 0, st>>>(E.tab, S, A, d_v, E.dfp.get(), d_hv, err);
    }
    HIPCHK(hipGetLastError());
  }

  void eam_density(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* d_types, double* d_rho) {
    require(nl.built, "List used before rebuild");
    Shifts S = shifts_of(nl.box);
    PairArgs A = pair_args(E, nl, d_types, nullptr);
    unsigned g = grid_for(nl.n, LANES);
    if (A.t0 < 0) {
      k_rho<LANES, true><<<g, NT, 0, E.ctx->stream>>>(E.tab, S, A, d_rho, E.ctx->dflags.get());
    } else {
      k_rho<LANES, false><<<g, NT, 0, E.ctx->stream>>>(E.tab, S, A, d_rho, E.ctx->dflags.get());
    }
    HIPCHK(hipGetLastError());
  }

  void eam_table_eval(flyg_eam_s& E, int kind, int a, int b, double const* x, long m, int deriv, double* out) {
    int n = E.tab.ntypes;
    require(a >= 0 && a < n && b >= 0 && b < n, "type index out of range");
    require(deriv >= 0 && deriv <= 2, "deriv must be 0, 1 or 2");
    SplineDesc s = kind == 0 ? E.tab.F[a] : kind == 1 ? E.tab.phi[a + n * b] : E.tab.V[a > b ? a + n * b : b + n * a];
    DBuf<double> dx, dout;
    dx.reserve(m);
    dout.reserve(m);
    hipStream_t st = E.ctx->stream;
    HIPCHK(hipMemcpyAsync(dx.get(), x, m * sizeof(double), hipMemcpyHostToDevice, st));
    k_table_eval<<<static_cast<unsigned>((m + 255) / 256), 256, 0, st>>>(E.tab, s, dx.get(), m, deriv, dout.get(),
                                                                         E.ctx->dflags.get());
    HIPCHK(hipMemcpyAsync(out, dout.get(), m * sizeof(double), hipMemcpyDeviceToHost, st));
    unsigned fl = E.ctx->take_flags();
    if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
    if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
  }

  namespace {
    // Names one upload of staged host types, unique in the process (lists cache per-slot species by it).
    long next_types_stamp() {
      static std::atomic<long> counter{0};
      return ++counter;
    }
  }  // namespace

  void stage_atoms(flyg_eam_s& E, uint32_t const* types, uint8_t const* frozen, long n, int where,
                   uint32_t const** d_types, uint8_t const** d_frozen) {
    if (!types && !frozen) {   // the handle's staged copy (stage_atoms_device or an earlier host call)
      require(E.staged_n == n, "no staged species and frozen flags for this many atoms");
      *d_types = E.types.get();
      *d_frozen = E.frozen.get();
      return;
    }
    require(types && frozen, "types and frozen must both be given or both be null");
    if (where == FLYG_DEVICE) {
      *d_types = types;
      *d_frozen = frozen;
      return;
    }
    bool same_t = static_cast<long>(E.last_types.size()) == n &&
                  std::memcmp(E.last_types.data(), types, n * sizeof(uint32_t)) == 0;
    bool same_f = static_cast<long>(E.last_frozen.size()) == n &&
                  std::memcmp(E.last_frozen.data(), frozen, n * sizeof(uint8_t)) == 0;
    int single = -1;
    if (!same_t) {   // validate before anything staged changes: a rejected call leaves no stale state
      single = n > 0 ? static_cast<int>(types[0]) : -1;
      for (long i = 0; i < n; ++i) {
        require(types[i] < static_cast<uint32_t>(E.tab.ntypes), "TypeID out of range");
        if (types[i] != types[0]) single = -1;
      }
    }
    // The device buffers may be reallocated below (contents lost): forget what they held first.
    if (!same_t || E.types.cap < static_cast<std::size_t>(n)) {
      E.last_types.clear();
      same_t = false;
    }
    if (!same_f || E.frozen.cap < static_cast<std::size_t>(n)) {
      E.last_frozen.clear();
      same_f = false;
    }
    E.types.reserve(n);
    E.frozen.reserve(n);
    if (!same_t) {
      HIPCHK(hipMemcpyAsync(E.types.get(), types, n * sizeof(uint32_t), hipMemcpyHostToDevice, E.ctx->stream));
      E.types_stamp = next_types_stamp();
      E.staged_single = single;
      E.last_types.assign(types, types + n);
    }
    if (!same_f) {
      HIPCHK(hipMemcpyAsync(E.frozen.get(), frozen, n * sizeof(uint8_t), hipMemcpyHostToDevice, E.ctx->stream));
      E.last_frozen.assign(frozen, frozen + n);
    }
    if (!same_t || !same_f) HIPCHK(hipStreamSynchronize(E.ctx->stream));
    E.staged_n = n;
    *d_types = E.types.get();
    *d_frozen = E.frozen.get();
  }

  void stage_atoms_device(flyg_eam_s& E, uint32_t const* types, uint8_t const* frozen, long n) {
    require(types && frozen, "null argument");
    E.staged_n = -1;
    E.last_types.clear();   // contents no longer match any host array
    E.last_frozen.clear();
    E.types.reserve(n);
    E.frozen.reserve(n);
    HIPCHK(hipMemcpyAsync(E.types.get(), types, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, E.ctx->stream));
    HIPCHK(hipMemcpyAsync(E.frozen.get(), frozen, n * sizeof(uint8_t), hipMemcpyDeviceToDevice, E.ctx->stream));
    E.types_stamp = next_types_stamp();
    E.staged_single = -1;   // not inspected on the host: the kernels treat the species as several
    E.staged_n = n;
  }

}  // namespace flyg
