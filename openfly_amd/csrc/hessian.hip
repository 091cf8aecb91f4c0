// Mass-weighted Hessian spectrum of the non-frozen atoms (the rate prefactor's input): EAM::hessian
// (eam.cpp:122-233), Hessian::mass_weight (hessian.hpp:136-143) and Hessian::eigenvalues
// (hessian.hpp:86-93; the reference calls Eigen's SelfAdjointEigenSolver), as used by Master for
// Mechanism::kinetic_pre (find.cpp:838-871) and the minimum's log-product (find.cpp:1094-1144).
//
// Assembly: column c of the 3M x 3M block (M = non-frozen atoms) is H e_c from the matrix-free
// Hessian-vector product kernels, gathered over the non-frozen rows and divided by sqrt(m_i m_j); the
// lower triangle is kept (the part the reference's solver reads) and mirrored.
// Diagonalisation: cyclic Jacobi rotations (Golub & Van Loan, Matrix Computations, Alg. 8.4.3) in the
// parallel round-robin ordering: each of the m - 1 rounds of a sweep rotates m / 2 disjoint index pairs
// at once (one workgroup per pair updates the pair's two rows, then every row's two columns), so a sweep
// touches every off-diagonal element once; sweeps repeat until the off-diagonal Frobenius norm is below
// 1e-14 of the matrix's.  The eigenvalues (the diagonal) come back sorted.
#include <algorithm>
#include <cmath>
#include <vector>

#include "eam_host.hpp"

namespace flyg {

  void eam_hvp(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* types, uint8_t const* frozen, double const* d_v,
               double* d_hv);

  namespace {

    constexpr int NTB = 256;

    __global__ void k_set(double* v, long i, double x) { v[i] = x; }

    // H[r][c] = hv[dof[r]] / sqrt(m_r m_c) for the non-frozen degrees of freedom r (column c).
    __global__ void k_gather_column(double const* __restrict__ hv, long const* __restrict__ dof,
                                    double const* __restrict__ rmass, long n3m, long c, double* __restrict__ H) {
      long r = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (r >= n3m) return;
      H[r * n3m + c] = hv[dof[r]] * (rmass[r] * rmass[c]);
    }

    // Keep the lower triangle: H[c][r] = H[r][c] for r > c.
    __global__ void k_mirror_lower(double* __restrict__ H, long n) {
      long idx = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (idx >= n * n) return;
      long r = idx / n, c = idx % n;
      if (r > c) H[c * n + r] = H[r * n + c];
    }

    // Pair k of round `round` in the round-robin ordering of m (even) indices: (round, m - 1) for k = 0,
    // ((round + k) mod (m - 1), (round - k) mod (m - 1)) otherwise; returned as p < q.
    __device__ inline void pair_of(int round, int k, int m, int& p, int& q) {
      int a, b;
      if (k == 0) {
        a = round;
        b = m - 1;
      } else {
        a = (round + k) % (m - 1);
        b = (round - k + (m - 1)) % (m - 1);
      }
      p = min(a, b);
      q = max(a, b);
    }

    // Symmetric Schur decomposition of the 2x2 block (Golub & Van Loan Alg. 8.4.1): c, s with
    // [c s; -s c]^T [app apq; apq aqq] [c s; -s c] diagonal.
    __device__ inline void schur2(double app, double aqq, double apq, double& c, double& s) {
      if (apq == 0.0) {
        c = 1.0;
        s = 0.0;
        return;
      }
      double tau = (aqq - app) / (2.0 * apq);
      double t = fabs(tau) > 1e150 ? 0.5 / tau : (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      c = 1.0 / sqrt(1.0 + t * t);
      s = t * c;
    }

    // One workgroup per pair: rotation from the current block, then the pair's two rows.
    __global__ void k_jacobi_rows(double* __restrict__ A, int n, int m, int round, double2* __restrict__ cs) {
      int p, q;
      pair_of(round, blockIdx.x, m, p, q);
      bool const live = q < n;   // an index >= n pads an odd n
      double c = 1.0, s = 0.0;
      if (live) schur2(A[static_cast<long>(p) * n + p], A[static_cast<long>(q) * n + q], A[static_cast<long>(p) * n + q], c, s);
      __syncthreads();   // every thread has read the 2x2 block before the rows change
      if (threadIdx.x == 0) cs[blockIdx.x] = make_double2(c, s);
      if (!live) return;
      double* rp = A + static_cast<long>(p) * n;
      double* rq = A + static_cast<long>(q) * n;
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        double a = rp[j], b = rq[j];
        rp[j] = c * a - s * b;
        rq[j] = s * a + c * b;
      }
    }

    // Every row's two columns of every pair of the round.
    __global__ void k_jacobi_cols(double* __restrict__ A, int n, int m, int round, double2 const* __restrict__ cs) {
      long idx = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      long const npairs = m / 2;
      if (idx >= npairs * n) return;
      int k = static_cast<int>(idx / n);
      long i = idx % n;
      int p, q;
      pair_of(round, k, m, p, q);
      if (q >= n) return;
      double2 r = cs[k];
      double* row = A + i * n;
      double a = row[p], b = row[q];
      row[p] = r.x * a - r.y * b;
      row[q] = r.y * a + r.x * b;
    }

    // Off-diagonal and total squared Frobenius norms, one block, fixed order.
    __global__ void k_norms(double const* __restrict__ A, long n, double* __restrict__ out) {
      __shared__ double so[NTB], st[NTB];
      double o = 0, t = 0;
      for (long idx = threadIdx.x; idx < n * n; idx += NTB) {
        double v = A[idx];
        t += v * v;
        if (idx / n != idx % n) o += v * v;
      }
      so[threadIdx.x] = o;
      st[threadIdx.x] = t;
      __syncthreads();
      for (int w = NTB / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
          so[threadIdx.x] += so[threadIdx.x + w];
          st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        out[0] = so[0];
        out[1] = st[0];
      }
    }

    __global__ void k_diag(double const* __restrict__ A, long n, double* __restrict__ d) {
      long i = blockIdx.x * static_cast<long>(blockDim.x) + threadIdx.x;
      if (i < n) d[i] = A[i * n + i];
    }

    unsigned blocks(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

  }  // namespace

  // Returns the sorted eigenvalues of the mass-weighted Hessian block of the non-frozen atoms.
  std::vector<double> hessian_eigenvalues(flyg_eam_s& E, flyg_list_s& nl, uint32_t const* d_types,
                                          uint8_t const* d_frozen, std::vector<uint8_t> const& h_frozen,
                                          std::vector<uint32_t> const& h_types, int* sweeps_out) {
    flyg_ctx_s& ctx = *E.ctx;
    hipStream_t st = ctx.stream;
    long const n = nl.n;
    std::vector<long> dof;
    std::vector<double> rmass;
    for (long a = 0; a < n; ++a) {
      if (h_frozen[a]) continue;
      uint32_t t = h_types[a];
      require(t < static_cast<uint32_t>(E.data.ntypes), "TypeID out of range");
      double m = E.data.mass[t];
      require(m > 0, "the potential gives a species no positive mass");
      for (int k = 0; k < 3; ++k) {
        dof.push_back(3 * a + k);
        rmass.push_back(1.0 / std::sqrt(m));
      }
    }
    long const n3m = static_cast<long>(dof.size());
    if (n3m == 0) return {};
    require(n3m <= 16384, "the Hessian block is limited to 16384 degrees of freedom");
    DBuf<double> v, hv, H, rm, nrm, diag;
    DBuf<long> ddof;
    DBuf<double2> cs;
    v.reserve(3 * n);
    hv.reserve(3 * n);
    H.reserve(n3m * n3m);
    rm.reserve(n3m);
    ddof.reserve(n3m);
    nrm.reserve(2);
    diag.reserve(n3m);
    long const m = n3m + (n3m % 2);
    cs.reserve(m / 2);
    HIPCHK(hipMemsetAsync(v.get(), 0, 3 * n * sizeof(double), st));
    HIPCHK(hipMemcpyAsync(ddof.get(), dof.data(), n3m * sizeof(long), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(rm.get(), rmass.data(), n3m * sizeof(double), hipMemcpyHostToDevice, st));
    for (long c = 0; c < n3m; ++c) {
      k_set<<<1, 1, 0, st>>>(v.get(), dof[c], 1.0);
      eam_hvp(E, nl, d_types, d_frozen, v.get(), hv.get());
      k_set<<<1, 1, 0, st>>>(v.get(), dof[c], 0.0);
      k_gather_column<<<blocks(n3m, NTB), NTB, 0, st>>>(hv.get(), ddof.get(), rm.get(), n3m, c, H.get());
    }
    k_mirror_lower<<<blocks(n3m * n3m, NTB), NTB, 0, st>>>(H.get(), n3m);
    HIPCHK(hipGetLastError());
    int sweeps = 0;
    double hn[2];
    for (; sweeps < 60; ++sweeps) {
      k_norms<<<1, NTB, 0, st>>>(H.get(), n3m, nrm.get());
      HIPCHK(hipMemcpyAsync(hn, nrm.get(), sizeof(hn), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      if (!(hn[0] > 1e-28 * hn[1])) break;   // off-diagonal norm below 1e-14 of the matrix's
      for (int r = 0; r < m - 1; ++r) {
        k_jacobi_rows<<<static_cast<unsigned>(m / 2), NTB, 0, st>>>(H.get(), static_cast<int>(n3m), static_cast<int>(m), r,
                                                                    cs.get());
        k_jacobi_cols<<<blocks((m / 2) * n3m, NTB), NTB, 0, st>>>(H.get(), static_cast<int>(n3m), static_cast<int>(m), r,
                                                                  cs.get());
      }
      HIPCHK(hipGetLastError());
    }
    require(sweeps < 60, "Jacobi sweeps did not converge");
    std::vector<double> w(n3m);
    k_diag<<<blocks(n3m, NTB), NTB, 0, st>>>(H.get(), n3m, diag.get());
    HIPCHK(hipMemcpyAsync(w.data(), diag.get(), n3m * sizeof(double), hipMemcpyDeviceToHost, st));
    unsigned fl = ctx.take_flags();   // synchronises
    if (fl & ERR_SPLINE_OOB) throw Error(FLYG_SPLINE_OOB, "x is outside tabulated region of a spline");
    if (fl & ERR_TYPE) throw Error(FLYG_INVALID_ARG, "TypeID out of range");
    std::sort(w.begin(), w.end());
    if (sweeps_out) *sweeps_out = sweeps;
    return w;
  }

}  // namespace flyg
