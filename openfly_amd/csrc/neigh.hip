// Neighbour-list build and update on gfx950 (replaces neigh::List, list.cpp:37-165).
//
// Rebuild pipeline (all on the device, one HIP stream):
//   1. canonicalise real positions                       (List::rebuild copy-in, list.cpp:68-72)
//   2. ghost images, axis by axis, chained               (List::make_ghosts, list.cpp:100-132)
//   3. counting-sort binning of reals + ghosts into the HyperGrid cells (atomic histogram, exclusive
//      scan, scatter, then each cell sorted by packed id so the result does not depend on atomics)
//   4. one thread per real atom scans its 27 cells and appends every (real, image) with d^2 < r_cut^2
//      (strict, list.cpp:146,159) to its row of an ELL table: n rows x K packed 32-bit entries.
// The distance tests, canonical images and ghost chains round exactly like the reference's FP64
// expressions (no FMA contraction), so the pair sets equal the reference's at rebuild time.
// Ghosts are not kept after the build: an entry stores (shift code, real index) and the pair kernels
// add the shift vector to the live real position, so update() touches only the N real atoms.
#include <algorithm>
#include <atomic>
#include <numeric>

#include "eam_host.hpp"

namespace flyg {

  namespace {

    constexpr int TPB = 256;

    __device__ inline void canon_image(BoxDesc const& b, double* x) {
#pragma clang fp contract(off)
      if (b.ortho) {
        for (int k = 0; k < 3; ++k) x[k] = x[k] - b.ext[k] * floor(x[k] * b.inv_ext[k]);
        return;
      }
      double f[3];
      for (int r = 0; r < 3; ++r) f[r] = (b.Binv[r][0] * x[0] + b.Binv[r][1] * x[1]) + b.Binv[r][2] * x[2];
      for (int r = 0; r < 3; ++r) f[r] = f[r] - floor(f[r]);
      for (int r = 0; r < 3; ++r) x[r] = (b.B[r][0] * f[0] + b.B[r][1] * f[1]) + b.B[r][2] * f[2];
    }

    // OrthoGrid::gen_image (orthorhombic.hpp:54-74) / TriGrid::gen_image (triclinic.hpp:47-71).
    __device__ inline bool gen_image(BoxDesc const& b, double rc, int sign, double* x, int ax) {
#pragma clang fp contract(off)
      if (b.ortho) {
        if (sign > 0) {
          if (x[ax] < rc) {
            x[ax] += b.ext[ax];
            return true;
          }
        } else if (x[ax] > b.ext[ax] - rc) {
          x[ax] -= b.ext[ax];
          return true;
        }
        return false;
      }
      if (sign > 0) {
        double d = (x[0] * b.hyper[0][ax] + x[1] * b.hyper[1][ax]) + x[2] * b.hyper[2][ax];
        if (d < rc) {
          for (int r = 0; r < 3; ++r) x[r] = x[r] + b.B[r][ax];
          return true;
        }
      } else {
        double d = ((b.B[0][ax] - x[0]) * b.hyper[0][ax] + (b.B[1][ax] - x[1]) * b.hyper[1][ax])
                   + (b.B[2][ax] - x[2]) * b.hyper[2][ax];
        if (d < rc) {
          for (int r = 0; r < 3; ++r) x[r] = x[r] - b.B[r][ax];
          return true;
        }
      }
      return false;
    }

    // Position of the image with shift code s of a real atom at x: basis columns added axis by axis
    // (the chain make_ghosts produces).
    __device__ inline void image_of(BoxDesc const& b, unsigned s, double* x) {
#pragma clang fp contract(off)
      int sa[3] = {static_cast<int>(s % 3) - 1, static_cast<int>((s / 3) % 3) - 1, static_cast<int>(s / 9) - 1};
      for (int ax = 0; ax < 3; ++ax) {
        if (sa[ax] > 0) {
          for (int r = 0; r < 3; ++r) x[r] = x[r] + b.B[r][ax];
        } else if (sa[ax] < 0) {
          for (int r = 0; r < 3; ++r) x[r] = x[r] - b.B[r][ax];
        }
      }
    }

    // HyperGrid::cell_idx (hypergrid.hpp:74-99): clamp(int((x + cell) / cell)) per axis, x fastest.
    __device__ inline int cell_idx(GridDesc const& g, double const* x, int* c3) {
#pragma clang fp contract(off)
      int idx = 0;
      for (int k = 0; k < 3; ++k) {
        int i = static_cast<int>((x[k] + g.cell[k]) * g.inv_cell[k]);
        i = i < 0 ? 0 : (i > g.shape[k] - 1 ? g.shape[k] - 1 : i);
        c3[k] = i;
        idx += i * g.prod[k];
      }
      return idx;
    }

    __device__ inline double dist2(double const* a, double const* b) {
#pragma clang fp contract(off)
      double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
      return (dx * dx + dy * dy) + dz * dz;
    }

    __global__ void k_update(double4* __restrict__ pos, double const* __restrict__ d, long n, int const* __restrict__ rank,
                             double4* __restrict__ gpos) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i >= n) return;
      double4 p = pos[i];
      p.x -= d[3 * i];
      p.y -= d[3 * i + 1];
      p.z -= d[3 * i + 2];
      pos[i] = p;
      double4* g = gpos + rank[i];   // its w slot may hold the atom's species: kept
      g->x = p.x;
      g->y = p.y;
      g->z = p.z;
    }

    __global__ void k_positions(double4 const* __restrict__ pos, long n, double* __restrict__ out) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i >= n) return;
      double4 p = pos[i];
      out[3 * i] = p.x;
      out[3 * i + 1] = p.y;
      out[3 * i + 2] = p.z;
    }

    // for_neighbours(i, r_cut) with the image chain: count, then write (j, shift, dr) rows.
    template <bool WRITE>
    __global__ void k_export(BoxDesc b, double rc, double4 const* __restrict__ pos, unsigned const* __restrict__ nbr,
                             int const* __restrict__ ncnt, int K, long n, long long const* __restrict__ off,
                             int* __restrict__ counts, int* __restrict__ oj, int* __restrict__ os, double* __restrict__ odr) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i >= n) return;
      double4 p = pos[i];
      double xi[3] = {p.x, p.y, p.z};
      int c = 0;
      long long o = WRITE ? off[i] : 0;
      int const cnt = min(ncnt[i], K);
      for (int k = 0; k < cnt; ++k) {
        unsigned e = nbr[i * static_cast<long>(K) + k];
        unsigned j = e & NB_IDX_MASK, s = e >> NB_SHIFT_BITS;
        double4 q = pos[j];
        double xn[3] = {q.x, q.y, q.z};
        image_of(b, s, xn);
        double dr[3], r2;
        {
#pragma clang fp contract(off)
          for (int r = 0; r < 3; ++r) dr[r] = xn[r] - xi[r];
          r2 = (dr[0] * dr[0] + dr[1] * dr[1]) + dr[2] * dr[2];
        }
        if (r2 < rc * rc) {
          if (WRITE) {
            oj[o + c] = static_cast<int>(j);
            os[o + c] = static_cast<int>(s);
            for (int r = 0; r < 3; ++r) odr[3 * (o + c) + r] = dr[r];
          }
          ++c;
        }
      }
      if (!WRITE) counts[i] = c;
    }

    // ---- single-sync rebuild path ----------------------------------------------------------------
    // lstate[0] ghosts generated, [1] longest row, [3] overflow bits.
    enum : unsigned { OV_GHOST = 1u };

    // Canonical images of the reals; the same launch clears the build's counters (cell counts za[0..na),
    // cell cursors zb[0..nb), the four state words), which only later kernels of the build use.
    __global__ void k_canon_reals(BoxDesc b, double const* __restrict__ xyz, long n, double4* __restrict__ pos,
                                  double4* __restrict__ ent, unsigned* __restrict__ eid, int* __restrict__ za, long na,
                                  int* __restrict__ zb, long nb, unsigned* __restrict__ st) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i < na) za[i] = 0;
      if (i < nb) zb[i] = 0;
      if (i < 4) st[i] = 0u;
      if (i >= n) return;
      double x[3] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
      canon_image(b, x);
      double4 p = make_double4(x[0], x[1], x[2], 0.0);
      pos[i] = p;
      ent[i] = p;
      eid[i] = (13u << NB_SHIFT_BITS) | static_cast<unsigned>(i);
    }

    // Slots for g images per lane, one atomic per wavefront (a prefix sum of the lanes' counts): one per
    // atom on the single counter serialises the kernel.  Slot order is free: the rank placement sorts
    // every cell by packed id afterwards.  Every lane of the wavefront must call it.
    __device__ inline unsigned ghost_slots(unsigned g, unsigned* st) {
      int const wl = static_cast<int>(threadIdx.x & 63);
      unsigned incl = g;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        unsigned t = __shfl_up(incl, o, 64);
        if (wl >= o) incl += t;
      }
      unsigned const total = __shfl(incl, 63, 64);
      unsigned wbase = 0;
      if (wl == 63 && total > 0) wbase = atomicAdd(&st[0], total);
      wbase = __shfl(wbase, 63, 64);
      return wbase + (incl - g);
    }

    // The reference's chain of gen_image calls (any box): images of images, axis by axis.
    __device__ inline void ghosts_chain(BoxDesc const& b, double rc, double4 const& p, bool inr, long i, long n,
                                        double4* __restrict__ ent, unsigned* __restrict__ eid, unsigned* __restrict__ st,
                                        long gcap) {
      double cx[27][3];
      int cs[27];
      int m = 1;
      cx[0][0] = p.x;
      cx[0][1] = p.y;
      cx[0][2] = p.z;
      cs[0] = 13;
      int const pow3[3] = {1, 3, 9};
      for (int ax = 0; ax < 3 && inr; ++ax) {
        if (!b.per[ax]) continue;
        int const end = m;
        for (int sgn = 1; sgn >= -1; sgn -= 2) {
          for (int c = 0; c < end; ++c) {
            double im[3] = {cx[c][0], cx[c][1], cx[c][2]};
            if (gen_image(b, rc, sgn, im, ax)) {
              cx[m][0] = im[0];
              cx[m][1] = im[1];
              cx[m][2] = im[2];
              cs[m] = cs[c] + sgn * pow3[ax];
              ++m;
            }
          }
        }
      }
      unsigned const g = static_cast<unsigned>(m - 1);
      unsigned const base = ghost_slots(g, st);
      if (g == 0) return;
      for (int k = 1; k < m; ++k) {
        long slot = static_cast<long>(base) + (k - 1);
        if (slot >= gcap) {
          atomicOr(&st[3], static_cast<unsigned>(OV_GHOST));
          return;
        }
        ent[n + slot] = make_double4(cx[k][0], cx[k][1], cx[k][2], 0.0);
        eid[n + slot] = (static_cast<unsigned>(cs[k]) << NB_SHIFT_BITS) | static_cast<unsigned>(i);
      }
    }

    // make_ghosts with a fixed capacity; the counter keeps counting past it so the host can resize.
    // Launched with whole wavefronts.  Orthorhombic boxes: whether an axis contributes a +1 / -1 image
    // depends only on the atom's own coordinate on that axis, so the images are the product of the
    // per-axis choices (the same set, and the same coordinates, as the reference's chain) and live in
    // registers; other boxes follow the chain of gen_image calls.  One instantiation per box kind: the
    // chain's per-lane arrays live in scratch memory, which the orthorhombic form does not need.
    template <bool ORTHO>
    __global__ void k_ghosts_cap(BoxDesc b, double rc, double4 const* __restrict__ pos, long n, double4* __restrict__ ent,
                                 unsigned* __restrict__ eid, unsigned* __restrict__ st, long gcap) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      bool const inr = i < n;
      double4 p = inr ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
      if constexpr (ORTHO) {
        double const x[3] = {p.x, p.y, p.z};
        bool up[3], dn[3];
        unsigned g = 1;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          up[ax] = inr && b.per[ax] && x[ax] < rc;
          dn[ax] = inr && b.per[ax] && x[ax] > b.ext[ax] - rc;
          g *= 1u + (up[ax] ? 1u : 0u) + (dn[ax] ? 1u : 0u);
        }
        g -= 1u;
        unsigned const base = ghost_slots(g, st);
        if (g == 0) return;
        unsigned k = 0;
#pragma unroll
        for (int s2 = -1; s2 <= 1; ++s2)
#pragma unroll
          for (int s1 = -1; s1 <= 1; ++s1)
#pragma unroll
            for (int s0 = -1; s0 <= 1; ++s0) {
              if (s0 == 0 && s1 == 0 && s2 == 0) continue;
              bool const use = (s0 == 0 || (s0 > 0 ? up[0] : dn[0])) && (s1 == 0 || (s1 > 0 ? up[1] : dn[1])) &&
                               (s2 == 0 || (s2 > 0 ? up[2] : dn[2]));
              if (!use) continue;
              long const slot = static_cast<long>(base) + k;
              ++k;
              if (slot >= gcap) {
                atomicOr(&st[3], static_cast<unsigned>(OV_GHOST));
                return;
              }
              double im[3] = {x[0], x[1], x[2]};
              {
#pragma clang fp contract(off)
                if (s0 > 0) im[0] += b.ext[0];
                if (s0 < 0) im[0] -= b.ext[0];
                if (s1 > 0) im[1] += b.ext[1];
                if (s1 < 0) im[1] -= b.ext[1];
                if (s2 > 0) im[2] += b.ext[2];
                if (s2 < 0) im[2] -= b.ext[2];
              }
              ent[n + slot] = make_double4(im[0], im[1], im[2], 0.0);
              eid[n + slot] = (static_cast<unsigned>(shift_code(s0, s1, s2)) << NB_SHIFT_BITS) | static_cast<unsigned>(i);
            }
      } else {
        ghosts_chain(b, rc, p, inr, i, n, ent, eid, st, gcap);
      }
    }

    __device__ inline long live_entries(unsigned const* st, long n, long gcap) {
      long g = static_cast<long>(st[0]);
      return n + (g < gcap ? g : gcap);
    }

    // Entries per cell in cnt[0 .. nc1), real atoms per cell in cnt[nc1 .. 2 nc1) (one scan of both gives
    // the cell starts and, less the first half's total, the first slot of each cell's reals).
    __global__ void k_cell_count_cap(GridDesc g, double4 const* __restrict__ ent, long n, long gcap,
                                     unsigned const* __restrict__ st, int* __restrict__ ecell, int* __restrict__ cnt,
                                     long nc1) {
      long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (e >= live_entries(st, n, gcap)) return;
      double4 p = ent[e];
      double x[3] = {p.x, p.y, p.z};
      int c3[3];
      int c = cell_idx(g, x, c3);
      ecell[e] = c;
      atomicAdd(&cnt[c], 1);
      if (e < n) atomicAdd(&cnt[nc1 + c], 1);
    }

    // Counting-sort scatter of the packed ids only (order inside a cell follows the atomics).
    __global__ void k_scatter_ids(unsigned const* __restrict__ eid, int const* __restrict__ ecell, long n, long gcap,
                                  unsigned const* __restrict__ st, int const* __restrict__ cstart, int* __restrict__ cursor,
                                  unsigned* __restrict__ usid) {
      long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (e >= live_entries(st, n, gcap)) return;
      int c = ecell[e];
      usid[cstart[c] + atomicAdd(&cursor[c], 1)] = eid[e];
    }

    // Places every entry at cell start + its rank among the cell's packed ids: cells come out sorted by
    // id whatever order the atomics took, so rows (and every sum over them) are reproducible.  A real
    // atom also gets its slot of the gradient layout: the reals in this sorted order (the cell's first
    // real slot + the reals of the cell with smaller ids), with its position by slot.
    __global__ void k_rank_place(double4 const* __restrict__ ent, unsigned const* __restrict__ eid,
                                 int const* __restrict__ ecell, long n, long gcap, unsigned const* __restrict__ st,
                                 int const* __restrict__ cstart, long nc1, unsigned const* __restrict__ usid,
                                 double4* __restrict__ spos, unsigned* __restrict__ sid, int* __restrict__ rslot,
                                 int* __restrict__ order, int* __restrict__ rank_of, double4* __restrict__ gpos) {
      long e = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (e >= live_entries(st, n, gcap)) return;
      int c = ecell[e];
      unsigned id = eid[e];
      int b = cstart[c], end = cstart[c + 1];
      int rank = 0, reals_before = 0;
      for (int q = b; q < end; ++q) {
        unsigned const u = usid[q];
        bool const lt = u < id;
        rank += lt ? 1 : 0;
        reals_before += (lt && (u >> NB_SHIFT_BITS) == 13u) ? 1 : 0;
      }
      spos[b + rank] = ent[e];
      sid[b + rank] = id;
      if (e < n) {   // the first n entries are the reals, entry e = atom e
        rslot[e] = b + rank;
        int const slot = cstart[nc1 + c] - cstart[nc1] + reals_before;
        order[slot] = static_cast<int>(e);
        rank_of[e] = slot;
        gpos[slot] = ent[e];
      }
    }

    // List::build_neigh_list (list.cpp:134-165): BL lanes per real atom walk the 3x3x3 block of cells
    // around it (z outermost, x innermost; each cell in packed-id order) with the strict r < r_c test
    // of the reference; the lanes split each cell and write their hits in candidate order (ballot +
    // prefix count), so a row is the same sequence a single thread would produce.
    constexpr int BL = 8;
    constexpr int BU = 4;   // candidate chunks loaded per step of the row build

    // Atoms are taken in slot (cell) order, so a workgroup's candidates overlap; each row is written
    // twice: by atom with atom indices (nbr) and by slot with slot indices (grow, the gradient layout).
    __global__ void k_build_rows(GridDesc g, double4 const* __restrict__ spos, unsigned const* __restrict__ sid,
                                 int const* __restrict__ cstart, int const* __restrict__ rslot, long n,
                                 unsigned* __restrict__ nbr, int* __restrict__ ncnt, int K, unsigned* __restrict__ st,
                                 int const* __restrict__ order, int const* __restrict__ rank_of,
                                 unsigned* __restrict__ grow, int* __restrict__ gcnt) {
      long gid = xcd_block(blockIdx.x, gridDim.x) * static_cast<long>(blockDim.x) + threadIdx.x;
      long const slot = gid / BL;
      int const lane = static_cast<int>(gid % BL);
      if (slot >= n) return;   // whole groups leave together (blockDim is a multiple of BL)
      long const i = order[slot];
      int const gbase = static_cast<int>(threadIdx.x & 63) - lane;
      int const self = rslot[i];
      double4 p = spos[self];
      double xi[3] = {p.x, p.y, p.z};
      int c3[3];
      cell_idx(g, xi, c3);
      double const rc2 = g.rc * g.rc;
      int cnt = 0;
      unsigned* row = nbr + i * static_cast<long>(K);
      unsigned* srow = grow + slot * static_cast<long>(K);
      // The up-to-three x-neighbouring cells of a (y, z) line are consecutive cells (x fastest), so their
      // sorted atoms form one contiguous range: 9 ranges instead of 27 cells, in the same order.  Each
      // step loads BU chunks of BL candidates (positions and packed ids) before testing them, then the
      // slots of all its hits, then writes the hits in order (more loads in flight, the same rows).
      int const x0 = max(c3[0] - 1, 0), x1 = min(c3[0] + 2, g.shape[0]);
      for (int z = max(c3[2] - 1, 0); z < min(c3[2] + 2, g.shape[2]); ++z)
        for (int y = max(c3[1] - 1, 0); y < min(c3[1] + 2, g.shape[1]); ++y) {
          int const line = y * g.prod[1] + z * g.prod[2];
          int const b = cstart[x0 * g.prod[0] + line], end = cstart[(x1 - 1) * g.prod[0] + line + 1];
          for (int s0 = b; s0 < end; s0 += BU * BL) {
            bool hit[BU];
            unsigned en[BU];
#pragma unroll
            for (int u = 0; u < BU; ++u) {
              int const s = s0 + u * BL + lane;
              hit[u] = false;
              en[u] = 0;
              if (s < end && s != self) {
                double4 q = spos[s];
                en[u] = sid[s];
                double xn[3] = {q.x, q.y, q.z};
                hit[u] = dist2(xi, xn) < rc2;
              }
            }
            int rk[BU];
#pragma unroll
            for (int u = 0; u < BU; ++u) rk[u] = hit[u] ? rank_of[en[u] & NB_IDX_MASK] : 0;
#pragma unroll
            for (int u = 0; u < BU; ++u) {
              unsigned long long m = __ballot(hit[u]);
              unsigned gm = static_cast<unsigned>((m >> gbase) & ((1ull << BL) - 1ull));
              if (hit[u]) {
                int k = cnt + __popc(gm & ((1u << lane) - 1u));
                if (k < K) {
                  row[k] = en[u];
                  srow[k] = (en[u] & ~NB_IDX_MASK) | static_cast<unsigned>(rk[u]);
                }
              }
              cnt += __popc(gm);
            }
          }
        }
      if (lane == 0) {
        ncnt[i] = cnt;
        gcnt[slot] = cnt;
        // Only overflowing rows report their length (the host widens the rows to the longest one): an
        // atomic per atom on the one counter serialised the kernel (half of its time).
        if (cnt > K) atomicMax(&st[1], static_cast<unsigned>(cnt));
      }
    }

    constexpr int SCAN_B = 1024;

    __global__ void k_scan_block(int* __restrict__ data, int n, int* __restrict__ sums) {
      __shared__ int wsum[SCAN_B / 64];
      int gi = blockIdx.x * SCAN_B + threadIdx.x;
      int v = gi < n ? data[gi] : 0;
      int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      int inc = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      if (lane == 63) wsum[w] = inc;
      __syncthreads();
      if (w == 0) {
        int s = lane < SCAN_B / 64 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          int t = __shfl_up(s, o, 64);
          if (lane >= o) s += t;
        }
        if (lane < SCAN_B / 64) wsum[lane] = s;
      }
      __syncthreads();
      int before = (w > 0 ? wsum[w - 1] : 0) + inc - v;
      if (gi < n) data[gi] = before;
      if (threadIdx.x == SCAN_B - 1) sums[blockIdx.x] = wsum[SCAN_B / 64 - 1];
    }

    // Exclusive scan of up to SCAN_B * SCAN_RUN values in one workgroup: thread t owns a run of
    // consecutive values, the runs' totals are scanned across the block, then each run is written out.
    constexpr int SCAN_RUN = 16;

    __global__ void k_scan_one(int* __restrict__ data, int n) {
      __shared__ int wsum[SCAN_B / 64];
      int const per = (n + SCAN_B - 1) / SCAN_B;
      int const b = static_cast<int>(threadIdx.x) * per;
      int const e = min(n, b + per);
      int v = 0;
      for (int q = b; q < e; ++q) v += data[q];
      int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      int inc = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(inc, o, 64);
        if (lane >= o) inc += t;
      }
      if (lane == 63) wsum[w] = inc;
      __syncthreads();
      if (w == 0) {
        int s = lane < SCAN_B / 64 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          int t = __shfl_up(s, o, 64);
          if (lane >= o) s += t;
        }
        if (lane < SCAN_B / 64) wsum[lane] = s;
      }
      __syncthreads();
      int run = (w > 0 ? wsum[w - 1] : 0) + inc - v;
      for (int q = b; q < e; ++q) {
        int const x = data[q];
        data[q] = run;
        run += x;
      }
    }

    __global__ void k_scan_add(int* __restrict__ data, int n, int const* __restrict__ sums) {
      int gi = blockIdx.x * SCAN_B + threadIdx.x;
      if (gi < n) data[gi] += sums[blockIdx.x];
    }

    inline unsigned nblk(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

  }  // namespace

  void exclusive_scan(flyg_ctx_s& ctx, int* data, int n, DBuf<int>& tmp) {
    if (n <= 0) return;
    if (n <= SCAN_B * SCAN_RUN) {   // one launch: each launch in a stream costs several microseconds
      k_scan_one<<<1, SCAN_B, 0, ctx.stream>>>(data, n);
      HIPCHK(hipGetLastError());
      return;
    }
    int nb = static_cast<int>((n + SCAN_B - 1) / SCAN_B);
    // Workspace layout: [sums of this level | workspace of the next level]
    std::vector<int> offs;
    std::vector<int> sizes;
    int total = 0, cur = nb;
    while (true) {
      offs.push_back(total);
      sizes.push_back(cur);
      total += cur;
      if (cur <= 1) break;
      cur = (cur + SCAN_B - 1) / SCAN_B;
    }
    tmp.reserve(static_cast<std::size_t>(total) + 1);
    // Down-sweep of block sums.
    int* level_data = data;
    int level_n = n;
    for (std::size_t l = 0; l < offs.size(); ++l) {
      int blocks = (level_n + SCAN_B - 1) / SCAN_B;
      k_scan_block<<<blocks, SCAN_B, 0, ctx.stream>>>(level_data, level_n, tmp.get() + offs[l]);
      if (blocks == 1) break;
      level_data = tmp.get() + offs[l];
      level_n = blocks;
    }
    // Up-sweep: add scanned block sums back, deepest level first.
    for (std::size_t l = offs.size(); l-- > 0;) {
      int ln = l == 0 ? n : sizes[l - 1];
      int* ld = l == 0 ? data : tmp.get() + offs[l - 1];
      int blocks = (ln + SCAN_B - 1) / SCAN_B;
      if (blocks == 1) continue;
      k_scan_add<<<blocks, SCAN_B, 0, ctx.stream>>>(ld, ln, tmp.get() + offs[l]);
    }
    HIPCHK(hipGetLastError());
  }

  void shift_table(BoxDesc const& b, double3* out) {
    for (int s = 0; s < 27; ++s) {
      int sa[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
      double v[3] = {0, 0, 0};
      for (int ax = 0; ax < 3; ++ax)
        for (int r = 0; r < 3; ++r) v[r] += sa[ax] * b.B[r][ax];
      out[s] = make_double3(v[0], v[1], v[2]);
    }
  }

  // One rebuild costs one host synchronisation: capacities (ghost slots, row length) come from the
  // previous build; the device records what it needed and the build is redone only on overflow.
  void list_rebuild(flyg_list_s& nl, double const* d_xyz, long n) {
    require(n > 0, "List::rebuild with no atoms");
    require(n < MAX_ATOMS, "too many atoms for the packed neighbour format");
    flyg_ctx_s& ctx = *nl.ctx;
    hipStream_t st = ctx.stream;
    if (!nl.hstate) HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&nl.hstate), 4 * sizeof(unsigned), hipHostMallocDefault));
    nl.lstate.reserve(4);
    nl.built = false;
    if (n != nl.n) {
      nl.gcap = 0;
      nl.K = 0;
    }
    nl.n = n;
    if (nl.gcap == 0) nl.gcap = std::max(4096l, n);
    if (nl.K == 0) nl.K = 64;
    int const nc = nl.grid.ncells;
    nl.pos.reserve(n);
    nl.cstart.reserve(2 * (nc + 1));   // cell starts, then the first real slot of each cell (offset)
    nl.ccursor.reserve(nc);
    nl.ncnt.reserve(n);
    nl.rslot.reserve(n);
    nl.order.reserve(n);
    nl.rank.reserve(n);
    nl.gpos.reserve(n);
    nl.gcnt.reserve(n);
    unsigned* dst = nl.lstate.get();

    for (int attempt = 0; attempt < 3; ++attempt) {
      long const cap = n + nl.gcap;
      nl.ent.reserve(cap);
      nl.eid.reserve(cap);
      nl.ecell.reserve(cap);
      nl.spos.reserve(cap);
      nl.sid.reserve(cap);
      nl.usid.reserve(cap);
      nl.nbr.reserve(static_cast<std::size_t>(n) * nl.K);
      nl.grow.reserve(static_cast<std::size_t>(n) * nl.K);
      long const nc1 = nc + 1;
      long const nz = std::max<long>(std::max<long>(n, 2 * nc1), 4);
      k_canon_reals<<<nblk(nz, TPB), TPB, 0, st>>>(nl.box, d_xyz, n, nl.pos.get(), nl.ent.get(), nl.eid.get(),
                                                   nl.cstart.get(), 2 * nc1, nl.ccursor.get(), nc, dst);
      if (nl.box.ortho) {
        k_ghosts_cap<true><<<nblk(n, TPB), TPB, 0, st>>>(nl.box, nl.rc, nl.pos.get(), n, nl.ent.get(), nl.eid.get(), dst,
                                                         nl.gcap);
      } else {
        k_ghosts_cap<false><<<nblk(n, TPB), TPB, 0, st>>>(nl.box, nl.rc, nl.pos.get(), n, nl.ent.get(), nl.eid.get(), dst,
                                                          nl.gcap);
      }
      k_cell_count_cap<<<nblk(cap, TPB), TPB, 0, st>>>(nl.grid, nl.ent.get(), n, nl.gcap, dst, nl.ecell.get(),
                                                       nl.cstart.get(), nc1);
      exclusive_scan(ctx, nl.cstart.get(), static_cast<int>(2 * nc1), nl.scan_tmp);
      k_scatter_ids<<<nblk(cap, TPB), TPB, 0, st>>>(nl.eid.get(), nl.ecell.get(), n, nl.gcap, dst, nl.cstart.get(),
                                                    nl.ccursor.get(), nl.usid.get());
      k_rank_place<<<nblk(cap, TPB), TPB, 0, st>>>(nl.ent.get(), nl.eid.get(), nl.ecell.get(), n, nl.gcap, dst,
                                                   nl.cstart.get(), nc1, nl.usid.get(), nl.spos.get(), nl.sid.get(),
                                                   nl.rslot.get(), nl.order.get(), nl.rank.get(), nl.gpos.get());
      k_build_rows<<<nblk(n * BL, TPB), TPB, 0, st>>>(nl.grid, nl.spos.get(), nl.sid.get(), nl.cstart.get(), nl.rslot.get(),
                                                 n, nl.nbr.get(), nl.ncnt.get(), nl.K, dst, nl.order.get(),
                                                 nl.rank.get(), nl.grow.get(), nl.gcnt.get());
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(nl.hstate, dst, 4 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      unsigned const G = nl.hstate[0], maxrow = nl.hstate[1], flags = nl.hstate[3];
      if (flags & OV_GHOST) {   // more ghost images than slots: grow and start again
        nl.gcap = static_cast<long>(G) + static_cast<long>(G) / 4 + 4096;
        continue;
      }
      nl.nghost = G;
      if (static_cast<int>(maxrow) > nl.K) {   // a row overflowed: widen the rows and build again
        nl.K = (static_cast<int>(maxrow) + 15) / 16 * 16 + 16;
        continue;
      }
      static std::atomic<long> builds{0};
      nl.build_id = ++builds;
      nl.built = true;
      return;
    }
    throw Error(FLYG_DEVICE_ERROR, "neighbour list build did not converge");
  }

  void list_update(flyg_list_s& nl, double const* d_delta, long n) {
    require(nl.built, "List::update before rebuild");
    require(n == nl.n, "Input to update has the wrong number of atoms");
    k_update<<<nblk(n, TPB), TPB, 0, nl.ctx->stream>>>(nl.pos.get(), d_delta, n, nl.rank.get(), nl.gpos.get());
    HIPCHK(hipGetLastError());
  }

  // Host-facing helpers used by the C-ABI.
  void list_positions(flyg_list_s& nl, double* d_out) {
    k_positions<<<nblk(nl.n, TPB), TPB, 0, nl.ctx->stream>>>(nl.pos.get(), nl.n, d_out);
    HIPCHK(hipGetLastError());
  }

  long list_count(flyg_list_s& nl, double rc, std::vector<int>& counts) {
    require(nl.built, "List used before rebuild");
    require(rc <= nl.rc, "r_cut is bigger than the internal r_cut");
    flyg_ctx_s& ctx = *nl.ctx;
    DBuf<int> dc;
    dc.reserve(nl.n);
    k_export<false><<<nblk(nl.n, TPB), TPB, 0, ctx.stream>>>(nl.box, rc, nl.pos.get(), nl.nbr.get(), nl.ncnt.get(), nl.K,
                                                            nl.n, nullptr, dc.get(), nullptr, nullptr, nullptr);
    counts.resize(nl.n);
    HIPCHK(hipMemcpyAsync(counts.data(), dc.get(), nl.n * sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    long t = 0;
    for (int c : counts) t += c;
    return t;
  }

  void list_export(flyg_list_s& nl, double rc, int64_t* offsets, int32_t* j, int32_t* shift, double* dr) {
    std::vector<int> counts;
    long total = list_count(nl, rc, counts);
    flyg_ctx_s& ctx = *nl.ctx;
    std::vector<long long> off(nl.n + 1);
    off[0] = 0;
    for (long i = 0; i < nl.n; ++i) off[i + 1] = off[i] + counts[i];
    for (long i = 0; i <= nl.n; ++i) offsets[i] = off[i];
    DBuf<long long> doff;
    DBuf<int> dj, ds;
    DBuf<double> ddr;
    doff.reserve(nl.n + 1);
    dj.reserve(total + 1);
    ds.reserve(total + 1);
    ddr.reserve(3 * total + 3);
    HIPCHK(hipMemcpyAsync(doff.get(), off.data(), (nl.n + 1) * sizeof(long long), hipMemcpyHostToDevice, ctx.stream));
    k_export<true><<<nblk(nl.n, TPB), TPB, 0, ctx.stream>>>(nl.box, rc, nl.pos.get(), nl.nbr.get(), nl.ncnt.get(), nl.K,
                                                           nl.n, doff.get(), nullptr, dj.get(), ds.get(), ddr.get());
    HIPCHK(hipMemcpyAsync(j, dj.get(), total * sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipMemcpyAsync(shift, ds.get(), total * sizeof(int), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipMemcpyAsync(dr, ddr.get(), 3 * total * sizeof(double), hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
  }

}  // namespace flyg
