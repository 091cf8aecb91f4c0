// Fused vector kernel and deterministic reductions (see vecops.hpp).
#include <cmath>

#include "vecops.hpp"

namespace flyg {

  namespace {

    constexpr int VT = 256;        // threads per block
    constexpr int VMAXB = 240;     // blocks (fixed per length -> reproducible sums)

    __device__ inline double coef_val(Coef const& c) {
      if (!c.s) return c.k;
      double s = *c.s;
      return c.k * (c.inv_sqrt ? 1.0 / sqrt(s) : s);
    }

    // Partials live in part[block * 4 + k]; `counter` elects the finalising block.
    __global__ __launch_bounds__(VT) void k_vec(VecArgs a, double* __restrict__ part, unsigned* __restrict__ counter) {
      __shared__ double sh[VT / 64];
      __shared__ bool last;
      double c0 = a.v[0] ? coef_val(a.c[0]) : 0.0;
      double c1 = a.v[1] ? coef_val(a.c[1]) : 0.0;
      double c2 = a.v[2] ? coef_val(a.c[2]) : 0.0;
      double m3[3] = {0, 0, 0};
      if (a.mean3)
        for (int k = 0; k < 3; ++k) m3[k] = a.mean_k * a.mean3[k];
      bool write = a.out && (a.v[0] || a.mean3);
      double acc[4] = {0, 0, 0, 0};
      long stride = static_cast<long>(gridDim.x) * VT;
      for (long i = blockIdx.x * static_cast<long>(VT) + threadIdx.x; i < a.n; i += stride) {
        double o = 0;
        if (write) {
          o = 0;
          if (a.v[0]) o = c0 * a.v[0][i];
          if (a.v[1]) o += c1 * a.v[1][i];
          if (a.v[2]) o += c2 * a.v[2][i];
          if (!a.v[0]) o = a.out[i];
          if (a.mean3) o += m3[i % 3];
          a.out[i] = o;
        } else if (a.out) {
          o = a.out[i];
        }
        if (a.comp_sums) {
          acc[i % 3] += o;
        } else {
          for (int k = 0; k < a.ndot; ++k) {
            double x = a.X[k] ? a.X[k][i] : o;
            double y = a.Y[k] ? a.Y[k][i] : o;
            acc[k] += x * y;
          }
        }
      }
      int nacc = a.comp_sums ? 3 : a.ndot;
      if (nacc == 0) return;
      for (int k = 0; k < nacc; ++k) {
        double s = block_sum<VT>(acc[k], sh);
        if (threadIdx.x == 0) __hip_atomic_store(&part[blockIdx.x * 4 + k], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __threadfence();
      if (threadIdx.x == 0) {
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
      }
      __syncthreads();
      if (!last || threadIdx.x != 0) return;
      __threadfence();
      double tot[4] = {0, 0, 0, 0};
      for (unsigned b = 0; b < gridDim.x; ++b)
        for (int k = 0; k < nacc; ++k) tot[k] += __hip_atomic_load(&part[b * 4 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *counter = 0u;
      for (int k = 0; k < nacc; ++k) a.slots[k] = tot[k];
      if (a.post == POST_HIST) {
        a.lb->rho[a.pj] = 1.0 / fabs(tot[0]);
        a.lb->h0 = (1.0 / a.lb->rho[a.pj]) / tot[1];
      } else if (a.post == POST_ALPHA) {
        a.lb->alpha[a.pj] = a.lb->rho[a.pj] * tot[0];
      } else if (a.post == POST_BETA) {
        a.lb->coef[a.pj] = a.lb->alpha[a.pj] - a.lb->rho[a.pj] * tot[0];
      }
    }

    __global__ __launch_bounds__(VT) void k_sum(double const* __restrict__ x, long n, double* __restrict__ part,
                                               unsigned* __restrict__ counter, double* __restrict__ out) {
      __shared__ double sh[VT / 64];
      __shared__ bool last;
      double acc = 0;
      long stride = static_cast<long>(gridDim.x) * VT;
      for (long i = blockIdx.x * static_cast<long>(VT) + threadIdx.x; i < n; i += stride) acc += x[i];
      double s = block_sum<VT>(acc, sh);
      if (threadIdx.x == 0) __hip_atomic_store(&part[blockIdx.x * 4], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __threadfence();
      if (threadIdx.x == 0) {
        unsigned t = atomicAdd(counter, 1u);
        last = (t == gridDim.x - 1);
      }
      __syncthreads();
      if (!last || threadIdx.x != 0) return;
      __threadfence();
      double tot = 0;
      for (unsigned b = 0; b < gridDim.x; ++b) tot += __hip_atomic_load(&part[b * 4], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *counter = 0u;
      *out = tot;
    }

    unsigned blocks_for(long n) {
      long b = (n + VT * 4 - 1) / (VT * 4);
      if (b < 1) b = 1;
      if (b > VMAXB) b = VMAXB;
      return static_cast<unsigned>(b);
    }

  }  // namespace

  void vec(flyg_ctx_s& ctx, VecArgs const& a) {
    if (a.n <= 0) return;
    ctx.dpart.reserve(VMAXB * 4);
    k_vec<<<blocks_for(a.n), VT, 0, ctx.stream>>>(a, ctx.dpart.get(), ctx.dflags.get() + 3);
    HIPCHK(hipGetLastError());
  }

  void dev_sum(flyg_ctx_s& ctx, double const* x, long n, double* d_out) {
    ctx.dpart.reserve(VMAXB * 4);
    k_sum<<<blocks_for(n), VT, 0, ctx.stream>>>(x, n, ctx.dpart.get(), ctx.dflags.get() + 3, d_out);
    HIPCHK(hipGetLastError());
  }

}  // namespace flyg
