// Vector kernels of the device-resident Krylov solvers: fused Jacobi-PCG
// updates with deterministic FP64 reductions, diagonal inversion, and the
// precision conversions of the defect-correction loop.
//
// Reductions: each kernel runs a grid fixed by the vector length (the CG
// kernels kVecBlocks x kVecThreads, the GMRES basis sweeps at most that, fewer
// blocks for short vectors) with a grid-stride loop, so every partial sum
// covers a fixed index set; the last block to finish (atomic ticket) adds the
// partials in a fixed order. Results are therefore bitwise reproducible run to
// run.
//
// Reference semantics restated: deal.II SolverCG under ReductionControl
// (include/meltpooldg/linear_algebra/linear_solver.hpp:36-38,106-111),
// JacobiPreconditioner::vmult dst = src .* inverse_diag
// (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:51-56),
// and the inversion guard of compute_inverse_diagonal_from_matrixfree
// (source/heat/heat_diffuse_operator.cpp:282-283).
#include "kernels.hpp"
#include "reduce.hpp"

namespace mpdg
{
  namespace
  {
    constexpr int T = kVecThreads;

    __device__ __forceinline__ double wave_sum(double v)
    {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
      return v;
    }

    // Final pass of a multi-column reduction, in the last block: column c
    // (partials[c * kVecBlocks + b], b < gridDim.x) is summed by wave c mod 4,
    // its lanes striding over the blocks, then a fixed butterfly over the lanes,
    // so the order is fixed and the latency chain is gridDim.x / 64 loads.
    __device__ __forceinline__ void ordered_column_sums(const double *partials, int ncol,
                                                        double *out)
    {
      const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
      for (int c = wv; c < ncol; c += T / 64)
        {
          double s = strided_sum(partials + c * kVecBlocks, lane, (int)gridDim.x, 64);
          s = wave_sum(s);
          if (lane == 0)
            out[c] = s;
        }
    }

    // Sums NV values per thread over the block; result valid in thread 0.
    template <int NV>
    __device__ __forceinline__ void block_sums(double (&v)[NV], double (*sh)[T / 64])
    {
#pragma unroll
      for (int k = 0; k < NV; ++k)
        v[k] = wave_sum(v[k]);
      const int w = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          sh[k][w] = v[k];
      __syncthreads();
      if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          {
            double s = 0.0;
            for (int i = 0; i < T / 64; ++i)
              s += sh[k][i];
            v[k] = s;
          }
      __syncthreads();
    }

    // Writes this block's NV partials, and in the last block returns the
    // ordered totals (in thread 0). Returns true in the last block.
    template <int NV>
    __device__ __forceinline__ bool finish_reduction(double (&v)[NV], double *partials,
                                                     unsigned *counter, double (&tot)[NV])
    {
      __shared__ double sh[NV][T / 64];
      __shared__ bool last;
      block_sums<NV>(v, sh);
      if (threadIdx.x == 0)
        {
#pragma unroll
          for (int k = 0; k < NV; ++k)
            store_partial(partials + k * kVecBlocks + blockIdx.x, v[k]);
          wait_vmem();
          const unsigned prev = atomicAdd(counter, 1u);
          last = prev == gridDim.x - 1;
        }
      __syncthreads();
      if (!last)
        return false;
      double s[NV];
#pragma unroll
      for (int k = 0; k < NV; ++k)
        s[k] = strided_sum(partials + k * kVecBlocks, threadIdx.x, (int)gridDim.x, T);
      block_sums<NV>(s, sh);
      if (threadIdx.x == 0)
        {
#pragma unroll
          for (int k = 0; k < NV; ++k)
            tot[k] = s[k];
          *counter = 0u;
        }
      return true;
    }

    __global__ void __launch_bounds__(T) invert_kernel(double *d, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const double v = d[i];
          d[i] = fabs(v) > 1.0e-10 ? 1.0 / v : 1.0;
        }
    }

    __global__ void __launch_bounds__(T)
      jacobi_kernel(double *dst, const double *src, const double *inv, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        dst[i] = src[i] * inv[i];
    }

    // Entry i of an inverse diagonal: stored in the solve's precision, or (the
    // FP32 inner CG of the mixed solve) as bfloat16, the upper half of the
    // FP32 bit pattern rounded to nearest even (to_bf16_kernel).
    __device__ __forceinline__ double inv_value(const double *inv, int64_t i) { return inv[i]; }
    __device__ __forceinline__ float inv_value(const float *inv, int64_t i) { return inv[i]; }
    __device__ __forceinline__ float inv_value(const uint16_t *inv, int64_t i)
    {
      return __uint_as_float((uint32_t)inv[i] << 16);
    }

    // r = b - Ax0, p = D r, rr = r.r, rz = r.p; sets tol, res0, done.
    template <typename Number, typename InvT>
    __global__ void __launch_bounds__(T)
      cg_init_kernel(const double *b, const Number *ax, Number *r, Number *p, const InvT *inv,
                     int64_t n, double rel_tol, double abs_tol, int max_it, double *partials,
                     SolverScalars *scal)
    {
      double v[2] = {0.0, 0.0};
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const Number ri = (Number)(b[i] - (ax ? (double)ax[i] : 0.0));
          const Number zi = ri * (Number)inv_value(inv, i);
          r[i] = ri;
          p[i] = zi;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)zi;
        }
      double tot[2];
      if (finish_reduction<2>(v, partials, &scal->counter, tot) && threadIdx.x == 0)
        {
          const double res = sqrt(tot[0]);
          scal->rr = tot[0];
          scal->rz = tot[1];
          scal->res = res;
          scal->res0 = res;
          scal->tol = fmax(abs_tol, rel_tol * res);
          scal->iteration = 0;
          scal->max_it = max_it;
          scal->status = 1;
          scal->done = 0;
          if (res <= scal->tol)
            {
              scal->done = 1;
              scal->status = 0;
            }
        }
    }

    // x += alpha p, r -= alpha q, rr = r.r, rz = r.(D r); convergence check.
    //
    // XM selects the x update of a pair of iterations (kXEvery, kXDefer,
    // kXPair): every iteration (x += alpha p), none with alpha kept in
    // scal->alpha_prev (the first of the pair), or both directions of the pair,
    // x = (x + alpha_prev p_prev) + alpha p (the second). The pair form
    // performs the same two roundings in the same order as two single
    // updates, so x is bitwise the same, and x is read and written every
    // other iteration only.
    template <typename Number, typename InvT, int XM>
    __global__ void __launch_bounds__(T)
      cg_update_kernel(Number *x, Number *r, const Number *p, const Number *p_prev,
                       const Number *q, const InvT *inv, int64_t n, double *partials,
                       SolverScalars *scal, double *history, int history_len)
    {
      if (scal->done)
        return;
      const double alpha = scal->alpha;
      const Number a = (Number)alpha;
      const Number a0 = XM == kXPair ? (Number)scal->alpha_prev : Number(0);
      double v[2] = {0.0, 0.0};
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          if constexpr (XM == kXEvery)
            x[i] = x[i] + a * p[i];
          else if constexpr (XM == kXPair)
            {
              Number xi = x[i] + a0 * p_prev[i];
              x[i] = xi + a * p[i];
            }
          const Number ri = r[i] - a * q[i];
          r[i] = ri;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)(ri * (Number)inv_value(inv, i));
        }
      double tot[2];
      if (finish_reduction<2>(v, partials, &scal->counter, tot) && threadIdx.x == 0)
        {
          const int it = scal->iteration + 1;
          const double res = sqrt(fabs(tot[0]));
          scal->iteration = it;
          scal->rr = tot[0];
          scal->res = res;
          if (history && it < history_len)
            history[it] = res;
          scal->beta = tot[1] / scal->rz;
          scal->rz_old = scal->rz;
          scal->rz = tot[1];
          if (XM == kXDefer)
            scal->alpha_prev = alpha;
          if (res <= scal->tol)
            {
              scal->done = 1;
              scal->status = 0;
            }
          else if (it >= scal->max_it || isnan(res))
            {
              scal->done = 1;
              scal->status = 1;
            }
        }
    }

    // p = D r + beta p_in (p_in may be p)
    template <typename Number, typename InvT>
    __global__ void __launch_bounds__(T)
      cg_direction_kernel(Number *p, const Number *p_in, const Number *r, const InvT *inv,
                          int64_t n, const SolverScalars *scal)
    {
      if (scal->done)
        return;
      const Number beta = (Number)scal->beta;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        p[i] = r[i] * (Number)inv_value(inv, i) + beta * p_in[i];
    }

    // The x update left over when a solve stops on the first iteration of a
    // pair: x += alpha p (the expression of the single update)
    template <typename Number>
    __global__ void __launch_bounds__(T)
      cg_flush_kernel(Number *x, const Number *p, int64_t n, const SolverScalars *scal)
    {
      const Number a = (Number)scal->alpha;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        x[i] = x[i] + a * p[i];
    }

    // 16-byte vector of Number (2 doubles or 4 floats) for the wide forms of
    // the update kernels; pointers are 16-byte aligned (checked at launch).
    template <typename Number>
    struct Wide;
    template <>
    struct Wide<double>
    {
      typedef double type __attribute__((ext_vector_type(2)));
      static constexpr int W = 2;
    };
    template <>
    struct Wide<float>
    {
      typedef float type __attribute__((ext_vector_type(4)));
      static constexpr int W = 4;
    };

    // W consecutive inverse-diagonal entries from W-entry group j
    template <typename Number>
    __device__ __forceinline__ typename Wide<Number>::type inv_wide(const Number *inv, int64_t j)
    {
      return reinterpret_cast<const typename Wide<Number>::type *>(inv)[j];
    }
    __device__ __forceinline__ Wide<float>::type inv_wide(const uint16_t *inv, int64_t j)
    {
      typedef uint16_t H4 __attribute__((ext_vector_type(4)));
      const H4 h = reinterpret_cast<const H4 *>(inv)[j];
      Wide<float>::type v;
      v[0] = __uint_as_float((uint32_t)h[0] << 16);
      v[1] = __uint_as_float((uint32_t)h[1] << 16);
      v[2] = __uint_as_float((uint32_t)h[2] << 16);
      v[3] = __uint_as_float((uint32_t)h[3] << 16);
      return v;
    }

    // cg_update_kernel over 16-byte vectors: thread t of the fixed grid owns
    // vectors t, t + grid, ... and (for the tail n % W) entry n - n % W + t.
    template <typename Number, typename InvT, int XM>
    __global__ void __launch_bounds__(T)
      cg_update_wide_kernel(Number *x, Number *r, const Number *p, const Number *p_prev,
                            const Number *q, const InvT *inv, int64_t n, double *partials,
                            SolverScalars *scal, double *history, int history_len)
    {
      if (scal->done)
        return;
      using V = typename Wide<Number>::type;
      constexpr int W = Wide<Number>::W;
      const double alpha = scal->alpha;
      const Number a = (Number)alpha;
      const Number a0 = XM == kXPair ? (Number)scal->alpha_prev : Number(0);
      const int64_t nv = n / W;
      V *X = reinterpret_cast<V *>(x);
      V *Rv = reinterpret_cast<V *>(r);
      const V *Pv = reinterpret_cast<const V *>(p);
      const V *P0 = reinterpret_cast<const V *>(p_prev);
      const V *Qv = reinterpret_cast<const V *>(q);
      double v[2] = {0.0, 0.0};
      const int64_t t0 = blockIdx.x * (int64_t)T + threadIdx.x;
      const int64_t stride = (int64_t)gridDim.x * T;
      for (int64_t j = t0; j < nv; j += stride)
        {
          const V qv = Qv[j], iv = inv_wide(inv, j);
          V rv = Rv[j];
          if constexpr (XM == kXEvery)
            {
              const V pv = Pv[j];
              V xv = X[j];
              xv = xv + a * pv;
              X[j] = xv;
            }
          else if constexpr (XM == kXPair)
            {
              const V pv = Pv[j], p0 = P0[j];
              V xv = X[j];
              xv = xv + a0 * p0;
              xv = xv + a * pv;
              X[j] = xv;
            }
          rv = rv - a * qv;
          Rv[j] = rv;
#pragma unroll
          for (int k = 0; k < W; ++k)
            {
              v[0] += (double)rv[k] * (double)rv[k];
              v[1] += (double)rv[k] * (double)(rv[k] * iv[k]);
            }
        }
      const int64_t i = nv * W + t0;
      if (i < n)
        {
          if constexpr (XM == kXEvery)
            x[i] = x[i] + a * p[i];
          else if constexpr (XM == kXPair)
            {
              Number xi = x[i] + a0 * p_prev[i];
              x[i] = xi + a * p[i];
            }
          const Number ri = r[i] - a * q[i];
          r[i] = ri;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)(ri * (Number)inv_value(inv, i));
        }
      double tot[2];
      if (finish_reduction<2>(v, partials, &scal->counter, tot) && threadIdx.x == 0)
        {
          const int it = scal->iteration + 1;
          const double res = sqrt(fabs(tot[0]));
          scal->iteration = it;
          scal->rr = tot[0];
          scal->res = res;
          if (history && it < history_len)
            history[it] = res;
          scal->beta = tot[1] / scal->rz;
          scal->rz_old = scal->rz;
          scal->rz = tot[1];
          if (XM == kXDefer)
            scal->alpha_prev = alpha;
          if (res <= scal->tol)
            {
              scal->done = 1;
              scal->status = 0;
            }
          else if (it >= scal->max_it || isnan(res))
            {
              scal->done = 1;
              scal->status = 1;
            }
        }
    }

    template <typename Number, typename InvT>
    __global__ void __launch_bounds__(T)
      cg_direction_wide_kernel(Number *p, const Number *p_in, const Number *r, const InvT *inv,
                               int64_t n, const SolverScalars *scal)
    {
      if (scal->done)
        return;
      using V = typename Wide<Number>::type;
      constexpr int W = Wide<Number>::W;
      const Number beta = (Number)scal->beta;
      const int64_t nv = n / W;
      V *Pv = reinterpret_cast<V *>(p);
      const V *Pi = reinterpret_cast<const V *>(p_in);
      const V *Rv = reinterpret_cast<const V *>(r);
      const int64_t t0 = blockIdx.x * (int64_t)T + threadIdx.x;
      for (int64_t j = t0; j < nv; j += (int64_t)gridDim.x * T)
        Pv[j] = Rv[j] * inv_wide(inv, j) + beta * Pi[j];
      const int64_t i = nv * W + t0;
      if (i < n)
        p[i] = r[i] * (Number)inv_value(inv, i) + beta * p_in[i];
    }

    bool aligned16(const void *a, const void *b = nullptr, const void *c = nullptr,
                   const void *d = nullptr, const void *e = nullptr)
    {
      auto ok = [](const void *v) { return v == nullptr || (reinterpret_cast<uintptr_t>(v) & 15u) == 0; };
      return ok(a) && ok(b) && ok(c) && ok(d) && ok(e);
    }

    // r = b - ax (FP64), rr -> scal->rr / res.
    __global__ void __launch_bounds__(T)
      residual_kernel(double *r, const double *b, const double *ax, int64_t n, double *partials,
                      SolverScalars *scal)
    {
      double v[1] = {0.0};
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const double ri = b[i] - ax[i];
          r[i] = ri;
          v[0] += ri * ri;
        }
      double tot[1];
      if (finish_reduction<1>(v, partials, &scal->counter, tot) && threadIdx.x == 0)
        {
          scal->rr = tot[0];
          scal->res = sqrt(tot[0]);
        }
    }

    __global__ void __launch_bounds__(T)
      dot_kernel(const double *a, const double *b, int64_t n, double *partials, double *out,
                 unsigned *counter)
    {
      double v[1] = {0.0};
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        v[0] += a[i] * b[i];
      double tot[1];
      if (finish_reduction<1>(v, partials, counter, tot) && threadIdx.x == 0)
        *out = tot[0];
    }

    __global__ void __launch_bounds__(T)
      axpby_kernel(double *dst, double a, const double *x, double b, const double *y, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        dst[i] = a * x[i] + b * y[i];
    }

    // v = 0 on the lattice nodes of the Dirichlet faces (bit 2d: x_d = lo, 2d+1: hi)
    __global__ void __launch_bounds__(T)
      zero_constrained_kernel(double *v, int n1x, int n1y, int n1z, unsigned faces)
    {
      const int64_t n = (int64_t)n1x * n1y * n1z;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const int ix = (int)(i % n1x);
          const int iy = (int)((i / n1x) % n1y);
          const int iz = (int)(i / ((int64_t)n1x * n1y));
          const bool con = ((faces & 1u) && ix == 0) || ((faces & 2u) && ix == n1x - 1) ||
                           ((faces & 4u) && iy == 0) || ((faces & 8u) && iy == n1y - 1) ||
                           ((faces & 16u) && iz == 0) || ((faces & 32u) && iz == n1z - 1);
          if (con)
            v[i] = 0.0;
        }
    }

    // FP64 -> bfloat16 by way of FP32, rounded to nearest even (finite values)
    __global__ void __launch_bounds__(T) to_bf16_kernel(uint16_t *dst, const double *src, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          uint32_t u = __float_as_uint((float)src[i]);
          u += 0x7FFFu + ((u >> 16) & 1u);
          dst[i] = (uint16_t)(u >> 16);
        }
    }

    __global__ void __launch_bounds__(T) to_f32_kernel(float *dst, const double *src, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        dst[i] = (float)src[i];
    }

    __global__ void __launch_bounds__(T) add_f32_kernel(double *x, const float *d, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        x[i] += (double)d[i];
    }

    __global__ void __launch_bounds__(T) fill_kernel(double *x, double v, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        x[i] = v;
    }

    __global__ void __launch_bounds__(T) fill_f32_kernel(float *x, float v, int64_t n)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        x[i] = v;
    }

    // h[c] = V_c . w for c < k (V_c = V + c*ld); deterministic, result in out[c].
    template <typename Number>
    __global__ void __launch_bounds__(T)
      multi_dot_kernel(const Number *V, int64_t ld, int k, const Number *w, int64_t n,
                       double *partials, double *out, unsigned *counter, const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      __shared__ double sh[kMaxBasis][T / 64];
      __shared__ bool last;
      double acc[kMaxBasis];
#pragma unroll
      for (int c = 0; c < kMaxBasis; ++c)
        acc[c] = 0.0;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const double wi = (double)w[i];
#pragma unroll
          for (int c = 0; c < kMaxBasis; ++c)
            if (c < k)
              acc[c] += (double)V[c * ld + i] * wi;
        }
      const int wv = threadIdx.x >> 6;
      for (int c = 0; c < k; ++c)
        {
          const double v = wave_sum(acc[c]);
          if ((threadIdx.x & 63) == 0)
            sh[c][wv] = v;
        }
      __syncthreads();
      if (threadIdx.x < k)
        {
          double s = 0.0;
          for (int i = 0; i < T / 64; ++i)
            s += sh[threadIdx.x][i];
          store_partial(partials + threadIdx.x * kVecBlocks + blockIdx.x, s);
          wait_vmem();
        }
      __syncthreads();
      if (threadIdx.x == 0)
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
      __syncthreads();
      if (!last)
        return;
      ordered_column_sums(partials, k, out);
      if (threadIdx.x == 0)
        *counter = 0u;
    }

    // w -= sum_c h[c] V_c; with nrm2 != nullptr also ||w||^2 -> *nrm2.
    template <typename Number>
    __global__ void __launch_bounds__(T)
      multi_axpy_kernel(Number *w, const Number *V, int64_t ld, int k, const double *h, int64_t n,
                        double *partials, double *nrm2, unsigned *counter, const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      __shared__ double hs[kMaxBasis];
      if (threadIdx.x < k)
        hs[threadIdx.x] = h[threadIdx.x];
      __syncthreads();
      double v[1] = {0.0};
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          double wi = (double)w[i];
          for (int c = 0; c < k; ++c)
            wi -= hs[c] * (double)V[c * ld + i];
          const Number r = (Number)wi;
          w[i] = r;
          v[0] += (double)r * (double)r;
        }
      if (nrm2 == nullptr)
        return;
      double tot[1];
      if (finish_reduction<1>(v, partials, counter, tot) && threadIdx.x == 0)
        *nrm2 = tot[0];
    }

    // w -= sum_c h1[c] V_c, then h2[c] = V_c . w (the classical Gram-Schmidt pass
    // and the dot products of the re-orthogonalisation pass in one sweep; the dots
    // use w rounded to Number, as a separate multi_dot after the axpy would).
    template <typename Number>
    __global__ void __launch_bounds__(T)
      multi_axpy_dot_kernel(Number *w, const Number *V, int64_t ld, int k, const double *h,
                            int64_t n, double *partials, double *out, unsigned *counter,
                            const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      __shared__ double hs[kMaxBasis];
      __shared__ double sh[kMaxBasis][T / 64];
      __shared__ bool last;
      if (threadIdx.x < k)
        hs[threadIdx.x] = h[threadIdx.x];
      __syncthreads();
      double acc[kMaxBasis];
#pragma unroll
      for (int c = 0; c < kMaxBasis; ++c)
        acc[c] = 0.0;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          double wi = (double)w[i];
          for (int c = 0; c < k; ++c)
            wi -= hs[c] * (double)V[c * ld + i];
          const Number r = (Number)wi;
          w[i] = r;
#pragma unroll
          for (int c = 0; c < kMaxBasis; ++c)
            if (c < k)
              acc[c] += (double)V[c * ld + i] * (double)r;
        }
      const int wv = threadIdx.x >> 6;
      for (int c = 0; c < k; ++c)
        {
          const double v = wave_sum(acc[c]);
          if ((threadIdx.x & 63) == 0)
            sh[c][wv] = v;
        }
      __syncthreads();
      if (threadIdx.x < k)
        {
          double s = 0.0;
          for (int i = 0; i < T / 64; ++i)
            s += sh[threadIdx.x][i];
          store_partial(partials + threadIdx.x * kVecBlocks + blockIdx.x, s);
          wait_vmem();
        }
      __syncthreads();
      if (threadIdx.x == 0)
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
      __syncthreads();
      if (!last)
        return;
      ordered_column_sums(partials, k, out);
      if (threadIdx.x == 0)
        *counter = 0u;
    }

    // ---- 16-byte sweeps over the basis -------------------------------------
    // The basis vectors start 16-byte aligned (ld a multiple of the vector
    // width), so a lane moves W entries of every vector per step. MODE:
    //   0  h[c] = V_c . w
    //   1  w -= sum_c h[c] V_c, then out[c] = V_c . w   (w rounded to Number)
    //   2  w -= sum_c h[c] V_c, then out[0] = ||w||^2
    // Lane t of the fixed grid owns vector chunks t, t + grid, ... and (for the
    // tail n % W) entry n - n % W + t; partials are summed in block order.
    enum SweepMode
    {
      kSweepDot = 0,
      kSweepAxpyDot = 1,
      kSweepAxpyNorm = 2
    };

    template <typename Number, int MODE>
    __global__ void __launch_bounds__(T)
      basis_sweep_kernel(Number *w, const Number *V, int64_t ld, int k, const double *h,
                         int64_t n, double *partials, double *out, unsigned *counter,
                         const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      using Vec = typename Wide<Number>::type;
      constexpr int W = Wide<Number>::W;
      constexpr int NACC = MODE == kSweepAxpyNorm ? 1 : kMaxBasis;
      constexpr int G = 8; // basis vectors loaded together
      __shared__ double hs[kMaxBasis];
      __shared__ double sh[NACC][T / 64];
      __shared__ bool last;
      // coefficients past k are 0, so the clamped loads below add nothing
      if (MODE != kSweepDot && threadIdx.x < kMaxBasis)
        hs[threadIdx.x] = (int)threadIdx.x < k ? h[threadIdx.x] : 0.0;
      __syncthreads();
      double acc[NACC];
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        acc[c] = 0.0;
      const int64_t nv = n / W;
      const int64_t ldv = ld / W;
      const Vec *Vv = reinterpret_cast<const Vec *>(V);
      Vec *Wv = reinterpret_cast<Vec *>(w);
      const int64_t t0 = blockIdx.x * (int64_t)T + threadIdx.x;
      const int64_t stride = (int64_t)gridDim.x * T;
      // V_c for c = g*G .. g*G+G-1, indices past k clamped to k-1 (loads stay in
      // bounds and are issued together; their contributions are discarded)
      auto load_group = [&](int g, int64_t j, Vec (&vc)[G]) {
#pragma unroll
        for (int u = 0; u < G; ++u)
          {
            const int c = min(g * G + u, k - 1);
            vc[u] = Vv[c * ldv + j];
          }
      };
      for (int64_t j = t0; j < nv; j += stride)
        {
          const Vec wv = Wv[j];
          double wd[W];
#pragma unroll
          for (int u = 0; u < W; ++u)
            wd[u] = (double)wv[u];
          if constexpr (MODE == kSweepDot)
            {
#pragma unroll
              for (int g = 0; g < kMaxBasis / G; ++g)
                if (g * G < k)
                  {
                    Vec vc[G];
                    load_group(g, j, vc);
#pragma unroll
                    for (int u = 0; u < G; ++u)
                      {
                        double d = 0.0;
#pragma unroll
                        for (int e = 0; e < W; ++e)
                          d += (double)vc[u][e] * wd[e];
                        acc[g * G + u] = g * G + u < k ? acc[g * G + u] + d : acc[g * G + u];
                      }
                  }
            }
          else
            {
              // after the loop vk holds the last group (g = (k-1)/G): the dot
              // pass below reuses it instead of loading it again
              Vec vk[G];
#pragma unroll
              for (int g = 0; g < kMaxBasis / G; ++g)
                if (g * G < k)
                  {
                    load_group(g, j, vk);
#pragma unroll
                    for (int u = 0; u < G; ++u)
#pragma unroll
                      for (int e = 0; e < W; ++e)
                        wd[e] -= hs[g * G + u] * (double)vk[u][e];
                  }
              Vec r;
#pragma unroll
              for (int e = 0; e < W; ++e)
                r[e] = (Number)wd[e];
              Wv[j] = r;
              if constexpr (MODE == kSweepAxpyNorm)
                {
#pragma unroll
                  for (int e = 0; e < W; ++e)
                    acc[0] += (double)r[e] * (double)r[e];
                }
              else
                {
                  const int gl = (k - 1) / G;
#pragma unroll
                  for (int g = 0; g < kMaxBasis / G; ++g)
                    if (g * G < k)
                      {
                        Vec vc[G];
                        if (g == gl)
                          {
#pragma unroll
                            for (int u = 0; u < G; ++u)
                              vc[u] = vk[u];
                          }
                        else
                          load_group(g, j, vc);
#pragma unroll
                        for (int u = 0; u < G; ++u)
                          {
                            double d = 0.0;
#pragma unroll
                            for (int e = 0; e < W; ++e)
                              d += (double)vc[u][e] * (double)r[e];
                            acc[g * G + u] = g * G + u < k ? acc[g * G + u] + d : acc[g * G + u];
                          }
                      }
                }
            }
        }
      const int64_t i = nv * W + t0;
      if (i < n)
        {
          double wi = (double)w[i];
          if constexpr (MODE == kSweepDot)
            {
#pragma unroll
              for (int c = 0; c < kMaxBasis; ++c)
                if (c < k)
                  acc[c] += (double)V[c * ld + i] * wi;
            }
          else
            {
              for (int c = 0; c < k; ++c)
                wi -= hs[c] * (double)V[c * ld + i];
              const Number r = (Number)wi;
              w[i] = r;
              if constexpr (MODE == kSweepAxpyNorm)
                acc[0] += (double)r * (double)r;
              else
                {
#pragma unroll
                  for (int c = 0; c < kMaxBasis; ++c)
                    if (c < k)
                      acc[c] += (double)V[c * ld + i] * (double)r;
                }
            }
        }
      const int nacc = MODE == kSweepAxpyNorm ? 1 : k;
      const int wv = threadIdx.x >> 6;
      // compile-time indices keep acc in registers
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        if (c < nacc)
          {
            const double v = wave_sum(acc[c]);
            if ((threadIdx.x & 63) == 0)
              sh[c][wv] = v;
          }
      __syncthreads();
      if (threadIdx.x < nacc)
        {
          double s = 0.0;
          for (int q = 0; q < T / 64; ++q)
            s += sh[threadIdx.x][q];
          store_partial(partials + threadIdx.x * kVecBlocks + blockIdx.x, s);
          wait_vmem();
        }
      __syncthreads();
      if (threadIdx.x == 0)
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
      __syncthreads();
      if (!last)
        return;
      ordered_column_sums(partials, nacc, out);
      if (threadIdx.x == 0)
        *counter = 0u;
    }

    // z = D .* (w / sqrt(nrm2)) and v = w / sqrt(nrm2) in one pass (the new
    // basis vector and its right-preconditioned image), 16-byte accesses.
    template <typename Number>
    __global__ void __launch_bounds__(T)
      scale_precondition_kernel(Number *v, Number *z, const Number *w, const Number *d,
                                const double *nrm2, int64_t n, const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      using Vec = typename Wide<Number>::type;
      constexpr int W = Wide<Number>::W;
      const double inv = 1.0 / sqrt(*nrm2);
      const int64_t nv = n / W;
      const Vec *Wv = reinterpret_cast<const Vec *>(w);
      const Vec *Dv = reinterpret_cast<const Vec *>(d);
      Vec *Vv = reinterpret_cast<Vec *>(v);
      Vec *Zv = reinterpret_cast<Vec *>(z);
      const int64_t t0 = blockIdx.x * (int64_t)T + threadIdx.x;
      for (int64_t j = t0; j < nv; j += (int64_t)gridDim.x * T)
        {
          const Vec wv = Wv[j], dv = Dv[j];
          Vec a, b;
#pragma unroll
          for (int u = 0; u < W; ++u)
            {
              a[u] = (Number)((double)wv[u] * inv);
              b[u] = a[u] * dv[u];
            }
          Vv[j] = a;
          Zv[j] = b;
        }
      const int64_t i = nv * W + t0;
      if (i < n)
        {
          const Number a = (Number)((double)w[i] * inv);
          v[i] = a;
          z[i] = a * d[i];
        }
    }

    // v = w / sqrt(nrm2)
    template <typename Number>
    __global__ void __launch_bounds__(T)
      scale_kernel(Number *v, const Number *w, const double *nrm2, int64_t n, const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      const double inv = 1.0 / sqrt(*nrm2);
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        v[i] = (Number)((double)w[i] * inv);
    }

    // z = D .* v (right Jacobi preconditioner)
    template <typename Number>
    __global__ void __launch_bounds__(T)
      diag_scale_kernel(Number *z, const Number *v, const Number *d, int64_t n, const int *stop)
    {
      if (stop != nullptr && *stop)
        return;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        z[i] = v[i] * d[i];
    }

    // x += D .* (sum_c y[c] V_c)
    template <typename Number, typename XNumber>
    __global__ void __launch_bounds__(T)
      gmres_update_kernel(XNumber *x, const Number *V, int64_t ld, int k, const double *y,
                          const Number *d, int64_t n, const int *kdev)
    {
      if (kdev != nullptr)
        k = min(max(*kdev, 0), kMaxBasis);
      __shared__ double ys[kMaxBasis];
      if (threadIdx.x < k)
        ys[threadIdx.x] = y[threadIdx.x];
      __syncthreads();
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          double s = 0.0;
          for (int c = 0; c < k; ++c)
            s += ys[c] * (double)V[c * ld + i];
          x[i] = (XNumber)((double)x[i] + (double)d[i] * s);
        }
    }

    // 16-byte sweeps need every basis vector and w 16-byte aligned
    template <typename Number>
    bool wide_basis(const Number *V, int64_t ld, const Number *w)
    {
      return aligned16(V, w) && ld % Wide<Number>::W == 0;
    }

    inline int grid_for(int64_t n)
    {
      const int64_t g = (n + T - 1) / T;
      return (int)(g < kVecBlocks ? (g < 1 ? 1 : g) : kVecBlocks);
    }

    // grid of a pure streaming kernel (no block partials): one element per
    // thread, so every load of the vector is in flight at once
    inline int stream_grid(int64_t n)
    {
      const int64_t g = (n + T - 1) / T;
      return (int)(g < 1 ? 1 : (g > (1 << 24) ? (1 << 24) : g));
    }

    // grid of the 16-byte basis sweeps (W entries per lane and step)
    template <typename Number>
    int sweep_grid(int64_t n)
    {
      return grid_for((n + Wide<Number>::W - 1) / Wide<Number>::W);
    }
  } // namespace

  hipError_t launch_invert_diagonal(double *d, int64_t n, hipStream_t s)
  {
    invert_kernel<<<stream_grid(n), T, 0, s>>>(d, n);
    return hipGetLastError();
  }

  hipError_t launch_jacobi(double *dst, const double *src, const double *inv, int64_t n,
                           hipStream_t s)
  {
    jacobi_kernel<<<stream_grid(n), T, 0, s>>>(dst, src, inv, n);
    return hipGetLastError();
  }

  // Reductions always use the full fixed grid so each partial covers a fixed index set.
  template <typename Number, typename InvT>
  hipError_t launch_cg_init(const double *b, const Number *ax, Number *r, Number *p,
                            const InvT *inv_diag, int64_t n, double rel_tol, double abs_tol,
                            int max_it, double *partials, SolverScalars *scal, hipStream_t s)
  {
    cg_init_kernel<Number, InvT><<<kVecBlocks, T, 0, s>>>(b, ax, r, p, inv_diag, n, rel_tol,
                                                          abs_tol, max_it, partials, scal);
    return hipGetLastError();
  }

  template <typename Number, typename InvT>
  hipError_t launch_cg_update(Number *x, Number *r, const Number *p, const Number *q,
                              const InvT *inv_diag, int64_t n, double *partials,
                              SolverScalars *scal, double *history, int history_len, hipStream_t s,
                              int x_mode, const Number *p_prev)
  {
    if (x_mode == kXPair && p_prev == nullptr)
      return hipErrorInvalidValue;
    const bool wide = aligned16(x, r, p, q, inv_diag) && aligned16(p_prev);
    auto go = [&](auto xm) {
      constexpr int XM = decltype(xm)::value;
      if (wide)
        cg_update_wide_kernel<Number, InvT, XM><<<kVecBlocks, T, 0, s>>>(
          x, r, p, p_prev, q, inv_diag, n, partials, scal, history, history_len);
      else
        cg_update_kernel<Number, InvT, XM><<<kVecBlocks, T, 0, s>>>(
          x, r, p, p_prev, q, inv_diag, n, partials, scal, history, history_len);
    };
    if (x_mode == kXDefer)
      go(std::integral_constant<int, kXDefer>());
    else if (x_mode == kXPair)
      go(std::integral_constant<int, kXPair>());
    else
      go(std::integral_constant<int, kXEvery>());
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_cg_flush(Number *x, const Number *p, int64_t n, const SolverScalars *scal,
                             hipStream_t s)
  {
    cg_flush_kernel<Number><<<stream_grid(n), T, 0, s>>>(x, p, n, scal);
    return hipGetLastError();
  }

  template <typename Number, typename InvT>
  hipError_t launch_cg_direction(Number *p, const Number *r, const InvT *inv_diag, int64_t n,
                                 const SolverScalars *scal, hipStream_t s, const Number *p_in)
  {
    if (p_in == nullptr)
      p_in = p;
    if (aligned16(p, r, inv_diag, p_in))
      {
        const int64_t nv = (n + Wide<Number>::W - 1) / Wide<Number>::W;
        cg_direction_wide_kernel<Number, InvT><<<stream_grid(nv), T, 0, s>>>(p, p_in, r, inv_diag,
                                                                             n, scal);
      }
    else
      cg_direction_kernel<Number, InvT><<<stream_grid(n), T, 0, s>>>(p, p_in, r, inv_diag, n,
                                                                     scal);
    return hipGetLastError();
  }

  hipError_t launch_residual(double *r, const double *b, const double *ax, int64_t n,
                             double *partials, SolverScalars *scal, hipStream_t s)
  {
    residual_kernel<<<kVecBlocks, T, 0, s>>>(r, b, ax, n, partials, scal);
    return hipGetLastError();
  }

  hipError_t launch_dot(const double *a, const double *b, int64_t n, double *partials, double *out,
                        unsigned *counter, hipStream_t s)
  {
    dot_kernel<<<kVecBlocks, T, 0, s>>>(a, b, n, partials, out, counter);
    return hipGetLastError();
  }

  hipError_t launch_zero_constrained(double *v, int n1x, int n1y, int n1z, unsigned faces,
                                     hipStream_t s)
  {
    zero_constrained_kernel<<<stream_grid((int64_t)n1x * n1y * n1z), T, 0, s>>>(v, n1x, n1y, n1z,
                                                                            faces);
    return hipGetLastError();
  }

  hipError_t launch_axpby(double *dst, double a, const double *x, double b, const double *y,
                          int64_t n, hipStream_t s)
  {
    axpby_kernel<<<stream_grid(n), T, 0, s>>>(dst, a, x, b, y, n);
    return hipGetLastError();
  }

  hipError_t launch_to_bf16(uint16_t *dst, const double *src, int64_t n, hipStream_t s)
  {
    to_bf16_kernel<<<stream_grid(n), T, 0, s>>>(dst, src, n);
    return hipGetLastError();
  }

  hipError_t launch_to_f32(float *dst, const double *src, int64_t n, hipStream_t s)
  {
    to_f32_kernel<<<stream_grid(n), T, 0, s>>>(dst, src, n);
    return hipGetLastError();
  }

  hipError_t launch_add_f32(double *x, const float *d, int64_t n, hipStream_t s)
  {
    add_f32_kernel<<<stream_grid(n), T, 0, s>>>(x, d, n);
    return hipGetLastError();
  }

  hipError_t launch_fill(double *x, double v, int64_t n, hipStream_t s)
  {
    fill_kernel<<<stream_grid(n), T, 0, s>>>(x, v, n);
    return hipGetLastError();
  }

  namespace
  {
    __global__ void __launch_bounds__(T)
      permute_kernel(double *dst, const double *src, const uint32_t *index, int64_t n,
                     bool to_lattice)
    {
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const uint32_t j = index[i];
          if ((int64_t)j >= n) // not a lattice index: left out rather than read or written
            continue;
          if (to_lattice)
            dst[j] = src[i];
          else
            dst[i] = src[j];
        }
    }
  } // namespace

  hipError_t launch_permute(double *dst, const double *src, const uint32_t *index, int64_t n,
                            bool to_lattice, hipStream_t s)
  {
    permute_kernel<<<stream_grid(n), T, 0, s>>>(dst, src, index, n, to_lattice);
    return hipGetLastError();
  }

  namespace
  {
    // One direction of the FE_Q_iso_Q1 -> FE_Q interpolation: lattice index i
    // of that direction, n cells of degree p: the index of the left equidistant
    // node of the sub-interval holding i's GLL point, and the point's fraction
    // of the way to the next node.
    __device__ __forceinline__ int iso_q1_left(const IsoQ1Map &m, int i, int n, int p, double &t)
    {
      const int c = min(i / p, n - 1);
      const int a = i - c * p;
      t = m.t[a];
      return c * p + m.s[a];
    }

    __global__ void __launch_bounds__(T)
      iso_q1_kernel(double *dst, const double *src, int n1x, int n1y, int n1z, int p,
                    IsoQ1Map m)
    {
      const int64_t n = (int64_t)n1x * n1y * n1z;
      const int nx = (n1x - 1) / p, ny = (n1y - 1) / p, nz = (n1z - 1) / p;
      for (int64_t i = blockIdx.x * (int64_t)T + threadIdx.x; i < n; i += (int64_t)gridDim.x * T)
        {
          const int ix = (int)(i % n1x);
          const int iy = (int)((i / n1x) % n1y);
          const int iz = (int)(i / ((int64_t)n1x * n1y));
          double tx, ty, tz;
          const int jx = iso_q1_left(m, ix, nx, p, tx);
          const int jy = iso_q1_left(m, iy, ny, p, ty);
          const int jz = iso_q1_left(m, iz, nz, p, tz);
          // trilinear within the sub-cell [jx, jx+1] x [jy, jy+1] x [jz, jz+1]
          // (jx + 1 <= nx * p = n1x - 1: always a lattice node)
          double v = 0.0;
#pragma unroll
          for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
              for (int a = 0; a < 2; ++a)
                {
                  const double w = (a ? tx : 1.0 - tx) * (b ? ty : 1.0 - ty) * (c ? tz : 1.0 - tz);
                  v += w * src[(jx + a) + (int64_t)n1x * ((jy + b) + (int64_t)n1y * (jz + c))];
                }
          dst[i] = v;
        }
    }
  } // namespace

  hipError_t launch_iso_q1_to_gll(double *dst, const double *src, int n1x, int n1y, int n1z,
                                  int p, const IsoQ1Map &m, hipStream_t s)
  {
    const int64_t n = (int64_t)n1x * n1y * n1z;
    iso_q1_kernel<<<stream_grid(n), T, 0, s>>>(dst, src, n1x, n1y, n1z, p, m);
    return hipGetLastError();
  }

  hipError_t launch_fill_f32(float *x, float v, int64_t n, hipStream_t s)
  {
    fill_f32_kernel<<<stream_grid(n), T, 0, s>>>(x, v, n);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_multi_dot(const Number *V, int64_t ld, int k, const Number *w, int64_t n,
                              double *partials, double *out, unsigned *counter, hipStream_t s,
                              const int *stop)
  {
    if (k < 1 || k > kMaxBasis)
      return hipErrorInvalidValue;
    if (wide_basis(V, ld, w))
      basis_sweep_kernel<Number, kSweepDot><<<sweep_grid<Number>(n), T, 0, s>>>(
        const_cast<Number *>(w), V, ld, k, nullptr, n, partials, out, counter, stop);
    else
      multi_dot_kernel<Number><<<grid_for(n), T, 0, s>>>(V, ld, k, w, n, partials, out, counter,
                                                        stop);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_multi_axpy(Number *w, const Number *V, int64_t ld, int k, const double *h,
                               int64_t n, double *partials, double *nrm2, unsigned *counter,
                               hipStream_t s, const int *stop)
  {
    if (k < 1 || k > kMaxBasis)
      return hipErrorInvalidValue;
    if (nrm2 != nullptr && wide_basis(V, ld, w))
      basis_sweep_kernel<Number, kSweepAxpyNorm><<<sweep_grid<Number>(n), T, 0, s>>>(
        w, V, ld, k, h, n, partials, nrm2, counter, stop);
    else
      multi_axpy_kernel<Number><<<grid_for(n), T, 0, s>>>(w, V, ld, k, h, n, partials, nrm2,
                                                         counter, stop);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_multi_axpy_dot(Number *w, const Number *V, int64_t ld, int k, const double *h,
                                   int64_t n, double *partials, double *out, unsigned *counter,
                                   hipStream_t s, const int *stop)
  {
    if (k < 1 || k > kMaxBasis)
      return hipErrorInvalidValue;
    if (wide_basis(V, ld, w))
      basis_sweep_kernel<Number, kSweepAxpyDot><<<sweep_grid<Number>(n), T, 0, s>>>(
        w, V, ld, k, h, n, partials, out, counter, stop);
    else
      multi_axpy_dot_kernel<Number><<<grid_for(n), T, 0, s>>>(w, V, ld, k, h, n, partials, out,
                                                             counter, stop);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_scale_precondition(Number *v, Number *z, const Number *w, const Number *d,
                                       const double *nrm2, int64_t n, hipStream_t s,
                                       const int *stop)
  {
    if (aligned16(v, z, w, d))
      {
        const int64_t nv = (n + Wide<Number>::W - 1) / Wide<Number>::W;
        scale_precondition_kernel<Number><<<stream_grid(nv), T, 0, s>>>(v, z, w, d, nrm2, n, stop);
      }
    else
      {
        scale_kernel<Number><<<stream_grid(n), T, 0, s>>>(v, w, nrm2, n, stop);
        diag_scale_kernel<Number><<<stream_grid(n), T, 0, s>>>(z, v, d, n, stop);
      }
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_scale(Number *v, const Number *w, const double *nrm2, int64_t n, hipStream_t s)
  {
    scale_kernel<Number><<<stream_grid(n), T, 0, s>>>(v, w, nrm2, n, nullptr);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_diag_scale(Number *z, const Number *v, const Number *d, int64_t n, hipStream_t s)
  {
    diag_scale_kernel<Number><<<stream_grid(n), T, 0, s>>>(z, v, d, n, nullptr);
    return hipGetLastError();
  }

  template <typename Number, typename XNumber>
  hipError_t launch_gmres_update(XNumber *x, const Number *V, int64_t ld, int k, const double *y,
                                 const Number *d, int64_t n, hipStream_t s, const int *kdev)
  {
    if (k < 1 || k > kMaxBasis)
      return hipErrorInvalidValue;
    gmres_update_kernel<Number, XNumber><<<stream_grid(n), T, 0, s>>>(x, V, ld, k, y, d, n, kdev);
    return hipGetLastError();
  }

  // ---- GMRES restart-cycle control on the device (one thread each) ----------
  namespace
  {
    __global__ void gmres_init_kernel(GmresState *st, double tol, int max_it)
    {
      st->tol = tol;
      st->max_it = max_it;
      st->it = 0;
      st->res = 0.0;
      st->sc.done = 0;
    }

    __global__ void gmres_start_kernel(GmresState *st, const SolverScalars *rs, int restart)
    {
      st->sc.done = 0;
      st->inner_conv = 0;
      st->k = restart;
      st->g[0] = rs->res;
      for (int l = 1; l <= kMaxBasis; ++l)
        st->g[l] = 0.0;
    }

    // The arithmetic of deal.II's Givens update restated: rotate the new
    // column by the earlier rotations, annihilate its subdiagonal entry, rotate g.
    __global__ void gmres_givens_kernel(GmresState *st, const double *h1, const double *h2,
                                        const double *nrm2, int j, int restart)
    {
      if (st->sc.done)
        return;
      st->it += 1;
      double col[kMaxBasis + 1];
      for (int l = 0; l <= j; ++l)
        col[l] = h1[l] + h2[l];
      const double hn = sqrt(*nrm2);
      col[j + 1] = hn;
      for (int l = 0; l < j; ++l)
        {
          const double a = st->cs[l] * col[l] + st->sn[l] * col[l + 1];
          const double b = -st->sn[l] * col[l] + st->cs[l] * col[l + 1];
          col[l] = a;
          col[l + 1] = b;
        }
      const double den = hypot(col[j], col[j + 1]);
      st->cs[j] = col[j] / den;
      st->sn[j] = col[j + 1] / den;
      col[j] = den;
      st->g[j + 1] = -st->sn[j] * st->g[j];
      st->g[j] = st->cs[j] * st->g[j];
      for (int l = 0; l <= j; ++l)
        st->hm[l * kMaxBasis + j] = col[l];
      const double res = fabs(st->g[j + 1]);
      st->res = res;
      if (res <= st->tol || st->it >= st->max_it || isnan(res) || hn == 0.0)
        {
          st->sc.done = 1;
          st->k = j + 1;
          st->inner_conv = res <= st->tol ? 1 : 0;
        }
      else if (j + 1 == restart)
        st->k = restart;
    }

    __global__ void gmres_solve_kernel(GmresState *st)
    {
      const int k = st->k;
      for (int l = k - 1; l >= 0; --l)
        {
          double sum = st->g[l];
          for (int c = l + 1; c < k; ++c)
            sum -= st->hm[l * kMaxBasis + c] * st->y[c];
          st->y[l] = sum / st->hm[l * kMaxBasis + l];
        }
    }
  } // namespace

  hipError_t launch_gmres_init(GmresState *st, double tol, int max_it, hipStream_t s)
  {
    gmres_init_kernel<<<1, 1, 0, s>>>(st, tol, max_it);
    return hipGetLastError();
  }

  hipError_t launch_gmres_start(GmresState *st, const SolverScalars *rs, int restart, hipStream_t s)
  {
    if (restart < 1 || restart > kMaxBasis - 1)
      return hipErrorInvalidValue;
    gmres_start_kernel<<<1, 1, 0, s>>>(st, rs, restart);
    return hipGetLastError();
  }

  hipError_t launch_gmres_givens(GmresState *st, const double *h1, const double *h2,
                                 const double *nrm2, int j, int restart, hipStream_t s)
  {
    if (j < 0 || j >= restart || restart > kMaxBasis - 1)
      return hipErrorInvalidValue;
    gmres_givens_kernel<<<1, 1, 0, s>>>(st, h1, h2, nrm2, j, restart);
    return hipGetLastError();
  }

  hipError_t launch_gmres_solve(GmresState *st, hipStream_t s)
  {
    gmres_solve_kernel<<<1, 1, 0, s>>>(st);
    return hipGetLastError();
  }

#define MPDG_GMRES_INST(NUM)                                                                     \
  template hipError_t launch_multi_dot<NUM>(const NUM *, int64_t, int, const NUM *, int64_t,     \
                                            double *, double *, unsigned *, hipStream_t,         \
                                            const int *);                                        \
  template hipError_t launch_multi_axpy<NUM>(NUM *, const NUM *, int64_t, int, const double *,   \
                                             int64_t, double *, double *, unsigned *,            \
                                             hipStream_t, const int *);                          \
  template hipError_t launch_multi_axpy_dot<NUM>(NUM *, const NUM *, int64_t, int, const double *,\
                                                 int64_t, double *, double *, unsigned *,        \
                                                 hipStream_t, const int *);                      \
  template hipError_t launch_scale_precondition<NUM>(NUM *, NUM *, const NUM *, const NUM *,      \
                                                     const double *, int64_t, hipStream_t,       \
                                                     const int *);                               \
  template hipError_t launch_scale<NUM>(NUM *, const NUM *, const double *, int64_t, hipStream_t); \
  template hipError_t launch_diag_scale<NUM>(NUM *, const NUM *, const NUM *, int64_t, hipStream_t);
  MPDG_GMRES_INST(double)
  MPDG_GMRES_INST(float)
#undef MPDG_GMRES_INST
  template hipError_t launch_gmres_update<double, double>(double *, const double *, int64_t, int,
                                                          const double *, const double *, int64_t,
                                                          hipStream_t, const int *);
  template hipError_t launch_gmres_update<float, float>(float *, const float *, int64_t, int,
                                                        const double *, const float *, int64_t,
                                                        hipStream_t, const int *);

  template hipError_t launch_cg_init<double>(const double *, const double *, double *, double *,
                                             const double *, int64_t, double, double, int,
                                             double *, SolverScalars *, hipStream_t);
  template hipError_t launch_cg_init<float>(const double *, const float *, float *, float *,
                                            const float *, int64_t, double, double, int, double *,
                                            SolverScalars *, hipStream_t);
  template hipError_t launch_cg_update<double>(double *, double *, const double *, const double *,
                                               const double *, int64_t, double *, SolverScalars *,
                                               double *, int, hipStream_t, int, const double *);
  template hipError_t launch_cg_update<float>(float *, float *, const float *, const float *,
                                              const float *, int64_t, double *, SolverScalars *,
                                              double *, int, hipStream_t, int, const float *);
  template hipError_t launch_cg_direction<double>(double *, const double *, const double *, int64_t,
                                                  const SolverScalars *, hipStream_t,
                                                  const double *);
  template hipError_t launch_cg_direction<float>(float *, const float *, const float *, int64_t,
                                                 const SolverScalars *, hipStream_t, const float *);
  template hipError_t launch_cg_flush<double>(double *, const double *, int64_t,
                                              const SolverScalars *, hipStream_t);
  template hipError_t launch_cg_flush<float>(float *, const float *, int64_t,
                                             const SolverScalars *, hipStream_t);
  template hipError_t launch_cg_init<float, uint16_t>(const double *, const float *, float *,
                                                      float *, const uint16_t *, int64_t, double,
                                                      double, int, double *, SolverScalars *,
                                                      hipStream_t);
  template hipError_t launch_cg_update<float, uint16_t>(float *, float *, const float *,
                                                        const float *, const uint16_t *, int64_t,
                                                        double *, SolverScalars *, double *, int,
                                                        hipStream_t, int, const float *);
  template hipError_t launch_cg_direction<float, uint16_t>(float *, const float *, const uint16_t *,
                                                           int64_t, const SolverScalars *,
                                                           hipStream_t, const float *);
} // namespace mpdg
