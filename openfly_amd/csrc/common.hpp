// Shared host/device definitions of libflygpu (MI355X / gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/flygpu.h"

namespace flyg {

  // ----------------------------------------------------------------------------------------------
  // Errors.  Host code throws Error(status, message); the C-ABI layer turns it into a status code
  // plus a per-context message.  No exception crosses the C-ABI.
  // ----------------------------------------------------------------------------------------------

  struct Error : std::runtime_error {
    int status;
    Error(int s, std::string const& m) : std::runtime_error(m), status(s) {}
  };

  inline void hip_check(hipError_t e, char const* what) {
    if (e != hipSuccess) throw Error(FLYG_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString(e));
  }
#define HIPCHK(x) ::flyg::hip_check((x), #x)

  inline void require(bool ok, std::string const& msg) {
    if (!ok) throw Error(FLYG_INVALID_ARG, msg);
  }

  // Device-side error flags (bit set, read back by the host after a call).
  enum : unsigned { ERR_SPLINE_OOB = 1u, ERR_LIST_OVERFLOW = 2u, ERR_REGION = 4u, ERR_NAN = 8u, ERR_TYPE = 16u };

  // A device-resident TypeID the host could not check: out of range sets ERR_TYPE (reported as
  // FLYG_INVALID_ARG at the next synchronisation) and species 0 is used so no table index leaves its array.
  __device__ inline int checked_type(uint32_t t, int ntypes, unsigned& err) {
    if (t < static_cast<uint32_t>(ntypes)) return static_cast<int>(t);
    err |= ERR_TYPE;
    return 0;
  }

  // ----------------------------------------------------------------------------------------------
  // Deferred frees.  hipFree and hipHostFree wait until every kernel on the device has finished, so
  // code that runs beside long kernels on other streams (the search server) must not call them.
  // While a FreeLater scope is open on a thread, the buffers that thread's DBufs (and pinned staging)
  // give up are collected in its Graveyard instead, and released by drain() when the device is idle.
  // ----------------------------------------------------------------------------------------------

  struct Graveyard {
    struct Item {
      void* p;
      std::size_t bytes;
      bool host;
    };
    std::vector<Item> items;
    std::size_t bytes = 0;
    Graveyard() = default;
    Graveyard(Graveyard const&) = delete;
    Graveyard& operator=(Graveyard const&) = delete;
    void drain() {
      for (auto const& it : items) (void)(it.host ? hipHostFree(it.p) : hipFree(it.p));
      items.clear();
      bytes = 0;
    }
    ~Graveyard() { drain(); }
  };

  inline thread_local Graveyard* tl_grave = nullptr;

  struct FreeLater {
    Graveyard* prev;
    explicit FreeLater(Graveyard& g) : prev(tl_grave) { tl_grave = &g; }
    ~FreeLater() { tl_grave = prev; }
    FreeLater(FreeLater const&) = delete;
    FreeLater& operator=(FreeLater const&) = delete;
  };

  inline hipError_t free_or_defer(void* p, std::size_t bytes, bool host) {
    if (tl_grave) {
      tl_grave->items.push_back({p, bytes, host});
      tl_grave->bytes += bytes;
      return hipSuccess;
    }
    return host ? hipHostFree(p) : hipFree(p);
  }

  // ----------------------------------------------------------------------------------------------
  // Growable device buffer (hipMalloc'd, never shrinks).
  // ----------------------------------------------------------------------------------------------

  template <typename T>
  struct DBuf {
    T* p = nullptr;
    std::size_t cap = 0;
    DBuf() = default;
    DBuf(DBuf const&) = delete;
    DBuf& operator=(DBuf const&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap) {
      o.p = nullptr;
      o.cap = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
      std::swap(p, o.p);
      std::swap(cap, o.cap);
      return *this;
    }
    ~DBuf() {
      if (p) (void)free_or_defer(p, cap * sizeof(T), false);
    }
    void reserve(std::size_t n) {
      if (n <= cap) return;
      if (p) HIPCHK(free_or_defer(p, cap * sizeof(T), false));
      p = nullptr;
      std::size_t c = n < 64 ? 64 : n;
      HIPCHK(hipMalloc(&p, c * sizeof(T)));
      cap = c;
    }
    T* get() const { return p; }
  };

  // ----------------------------------------------------------------------------------------------
  // Periodic box and cell grid (host + device POD).  The grid follows the reference's HyperGrid:
  // shape = 2 + floor(E / r_cut) per axis (one ghost layer each side), cell = E / floor(E / r_cut),
  // x-index fastest; TriGrid uses the column sums of the basis as E.
  // ----------------------------------------------------------------------------------------------

  struct BoxDesc {
    double B[3][3];      // B[row][col]: column c is basis vector c
    double ext[3], inv_ext[3];
    double Binv[3][3];
    double hyper[3][3];  // column c = unit normal of the plane spanned by the other two basis vectors
    int per[3];
    int ortho;
  };

  struct GridDesc {
    double rc;
    double cell[3], inv_cell[3];
    double gext[3];
    int shape[3], prod[3];
    int ncells;
  };

  BoxDesc make_box(double const* basis_rowmajor, int const* periodic);
  GridDesc make_grid(BoxDesc const& box, double r_cut);
  double box_min_width(BoxDesc const& b);

  // ----------------------------------------------------------------------------------------------
  // Spline tables.  One natural cubic spline = `count` segments of (a,b,c,d) on a uniform grid.
  // FP64 coefficients for the plain path; an FP32 copy for the mixed-precision path.
  // ----------------------------------------------------------------------------------------------

  struct SplineDesc {
    int off;        // first segment in the packed table
    int count;      // number of segments
    double dx, inv_dx;
    float dxf, inv_dxf;
  };

  // Potential tables as the kernels see them.  Index rules follow DataEAM: F[a]; phi(a, b) at a + n*b
  // (density at an atom of type b due to a neighbour of type a); V(a, b) symmetric at max + n*min.
  constexpr int MAX_TYPES = 4;

  struct EAMTables {
    int ntypes;
    double rc, rc2;
    float rcf;
    SplineDesc F[MAX_TYPES];
    SplineDesc phi[MAX_TYPES * MAX_TYPES];
    SplineDesc V[MAX_TYPES * MAX_TYPES];
    double4 const* c64;   // packed FP64 segments
    float4 const* c32;    // packed FP32 segments
  };

  // Packed neighbour entry: bits 0..25 real atom index, bits 26..30 image shift code
  // (s0+1) + 3 (s1+1) + 9 (s2+1), 13 = no shift.
  constexpr int NB_SHIFT_BITS = 26;
  constexpr unsigned NB_IDX_MASK = (1u << NB_SHIFT_BITS) - 1u;
  constexpr long MAX_ATOMS = 1l << NB_SHIFT_BITS;

  __host__ __device__ inline int shift_code(int s0, int s1, int s2) { return (s0 + 1) + 3 * (s1 + 1) + 9 * (s2 + 1); }

  // ----------------------------------------------------------------------------------------------
  // Device spline evaluation (reference Spline::fetch/f/fp/fpp, spline.hpp:81-154): segment
  // i = trunc(x / dx) via the stored reciprocal, t = x - i dx; out of range sets ERR_SPLINE_OOB.
  // ----------------------------------------------------------------------------------------------

  __device__ inline bool seg64(SplineDesc const& s, double x, int& i, double& t) {
    double q = x * s.inv_dx;
    if (!(q < static_cast<double>(s.count)) || q <= -1.0) return false;
    i = static_cast<int>(q);
    t = x - static_cast<double>(i) * s.dx;
    return true;
  }

  template <int D>   // 0 = f, 1 = f', 2 = f''
  __device__ inline double spl64(EAMTables const& T, SplineDesc const& s, double x, unsigned& err) {
    int i;
    double t;
    if (!seg64(s, x, i, t)) {
      err |= ERR_SPLINE_OOB;
      return 0.0;
    }
    double4 c = T.c64[s.off + i];
    if (D == 0) return c.x + t * (c.y + t * (c.z + t * c.w));
    if (D == 1) return c.y + t * (2.0 * c.z + t * 3.0 * c.w);
    return 2.0 * c.z + 6.0 * t * c.w;
  }

  // f and f' from one segment fetch.
  __device__ inline void spl64_ffp(EAMTables const& T, SplineDesc const& s, double x, double& f, double& fp, unsigned& err) {
    int i;
    double t;
    if (!seg64(s, x, i, t)) {
      err |= ERR_SPLINE_OOB;
      f = fp = 0.0;
      return;
    }
    double4 c = T.c64[s.off + i];
    f = c.x + t * (c.y + t * (c.z + t * c.w));
    fp = c.y + t * (2.0 * c.z + t * 3.0 * c.w);
  }

  // FP32 pair arithmetic on FP32 coefficients (mixed-precision form).
  template <int D>
  __device__ inline float spl32f(EAMTables const& T, SplineDesc const& s, float x, unsigned& err) {
    float q = x * s.inv_dxf;
    if (!(q < static_cast<float>(s.count)) || q <= -1.0f) {
      err |= ERR_SPLINE_OOB;
      return 0.0f;
    }
    int i = static_cast<int>(q);
    float t = x - static_cast<float>(i) * s.dxf;
    float4 c = T.c32[s.off + i];
    if (D == 0) return c.x + t * (c.y + t * (c.z + t * c.w));
    if (D == 1) return c.y + t * (2.0f * c.z + t * 3.0f * c.w);
    return 2.0f * c.z + 6.0f * t * c.w;
  }

  // Branch-free spline segment of x (reference Spline::fetch, spline.hpp:81-154): `use` is false for
  // entries that do not count (and for x out of range, which sets ERR_SPLINE_OOB), so a caller can
  // compute every segment of a group of entries, issue all their coefficient loads, and only then do
  // the arithmetic; the value is the same as spl64 / spl32f bit for bit.
  template <bool MIX>
  struct Seg {
    using R = typename std::conditional<MIX, float, double>::type;
    int i;
    R t;
    bool use;
  };

  template <bool MIX>
  __device__ inline Seg<MIX> seg_of(SplineDesc const& s, typename Seg<MIX>::R x, bool in, unsigned& e) {
    using R = typename Seg<MIX>::R;
    R q = x * (MIX ? static_cast<R>(s.inv_dxf) : static_cast<R>(s.inv_dx));
    bool ok = (q < static_cast<R>(s.count)) && !(q <= static_cast<R>(-1));
    if (in && !ok) e |= ERR_SPLINE_OOB;
    Seg<MIX> g;
    g.use = in && ok;
    g.i = g.use ? static_cast<int>(q) : 0;
    g.t = g.use ? x - static_cast<R>(g.i) * (MIX ? static_cast<R>(s.dxf) : static_cast<R>(s.dx)) : static_cast<R>(0);
    return g;
  }

  template <bool MIX>
  __device__ inline auto coef_of(EAMTables const& T, SplineDesc const& s, Seg<MIX> const& g) {
    if constexpr (MIX) {
      return T.c32[s.off + g.i];
    } else {
      return T.c64[s.off + g.i];
    }
  }

  template <typename C, typename R>
  __device__ inline R poly0(C const& c, R t) { return c.x + t * (c.y + t * (c.z + t * c.w)); }
  template <typename C, typename R>
  __device__ inline R poly1(C const& c, R t) { return c.y + t * (static_cast<R>(2) * c.z + t * static_cast<R>(3) * c.w); }

  // ----------------------------------------------------------------------------------------------
  // XCD-aware block order.  The dispatcher hands workgroup b to XCD b mod 8, and each XCD has its own
  // L2; renumbering so that XCD x processes one contiguous range of blocks keeps the atoms (and so
  // the neighbour gathers) of one XCD spatially together instead of spreading every L2 over the
  // whole system.  A bijection on [0, nb).
  // ----------------------------------------------------------------------------------------------
  constexpr unsigned NUM_XCD = 8;

  __device__ inline unsigned xcd_block(unsigned b, unsigned nb) {
    unsigned const per = nb / NUM_XCD, rem = nb % NUM_XCD;
    unsigned const x = b % NUM_XCD, k = b / NUM_XCD;
    return x * per + (x < rem ? x : rem) + k;
  }

  // ----------------------------------------------------------------------------------------------
  // Wavefront helpers (64 lanes on CDNA).
  // ----------------------------------------------------------------------------------------------

  template <int L>
  __device__ inline double group_sum(double v) {   // sum over aligned groups of L lanes
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
  }

  // Row sums that do not depend on how many lanes share a row: entry k of a row always goes to slot
  // k mod 16, each slot adds its entries in increasing k, and the 16 slots meet in the fixed xor tree
  // (8, 4, 2, 1).  A lane of an L-lane group owns slots lane + q L, q < 16 / L; tree levels >= L are
  // done in registers, the rest with cross-lane shuffles.  Every lane gets the row total.
  template <int L>
  struct Acc16 {
    static_assert(L >= 1 && L <= 16 && (16 % L) == 0, "L must divide 16");
    static constexpr int NS = 16 / L;
    double v[NS];
    __device__ Acc16() {
#pragma unroll
      for (int q = 0; q < NS; ++q) v[q] = 0.0;
    }
    // m = index of the entry within this lane's stride (k = lane + m L).
    // (__dadd_rn keeps the accumulation unfused, so every L rounds identically.)
    __device__ void add(int m, double x) {
      if (NS == 1) {
        v[0] = __dadd_rn(v[0], x);
      } else {
#pragma unroll
        for (int q = 0; q < NS; ++q)
          if (q == m % NS) v[q] = __dadd_rn(v[q], x);
      }
    }
    // Add to slot q (q < NS; a compile-time constant at the call sites, so the selects fold away).
    __device__ void add_slot(int q, double x) {
#pragma unroll
      for (int s = 0; s < NS; ++s)
        if (s == q) v[s] = __dadd_rn(v[s], x);
    }
    __device__ double total() {
#pragma unroll
      for (int bit = 8; bit >= 1; bit >>= 1) {
        if (bit >= L) {
#pragma unroll
          for (int q = 0; q < NS; ++q)
            if (((q * L) & bit) == 0 && q + bit / L < NS) v[q] = __dadd_rn(v[q], v[q + bit / L]);
        } else {
          v[0] = __dadd_rn(v[0], __shfl_xor(v[0], bit, 64));
        }
      }
      return v[0];
    }
  };

  // Deterministic block reduction of one double (fixed tree order); result valid in thread 0.
  template <int NT>
  __device__ inline double block_sum(double v, double* sh) {
    v = group_sum<64>(v);
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0) {
      for (int k = 0; k < NT / 64; ++k) s += sh[k];
    }
    __syncthreads();
    return s;
  }

}  // namespace flyg
