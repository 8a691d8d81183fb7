// Segmented vector kernels of the batched Jacobi-PCG: one launch updates the
// vectors of every solve of a batch (concatenated), each solve with its own
// scalars, its own deterministic reductions and its own convergence flag.
//
// Solve s owns entries [seg.off, seg.off + seg.n) and the workgroups
// [seg.blk0, seg.blk0 + seg.nblk) of every launch. Each workgroup writes its
// partial sums to partials[k * nblocks + workgroup]; the last workgroup of a
// solve (per-solve atomic ticket) adds that solve's partials in workgroup
// order, so results do not depend on scheduling.
//
// Same recurrence as the single-solve PCG (deal.II SolverCG under
// ReductionControl, include/meltpooldg/linear_algebra/linear_solver.hpp:36-38,
// 106-111, Jacobi dst = src .* inverse_diag, preconditioner_jacobi.hpp:51-56).
#include "kernels.hpp"
#include "reduce.hpp"

namespace mpdg
{
  namespace
  {
    constexpr int T = kBatchThreads;

    __device__ __forceinline__ double wave_sum(double v)
    {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
      return v;
    }

    __device__ __forceinline__ int find_solve(const BatchSeg *seg, int nsolve, int block)
    {
      int lo = 0, hi = nsolve;
      while (hi - lo > 1)
        {
          const int mid = (lo + hi) / 2;
          if (seg[mid].blk0 <= block)
            lo = mid;
          else
            hi = mid;
        }
      return lo;
    }

    // Block sums of NV values -> partials; true (thread 0 of) the solve's last
    // workgroup, which receives the ordered per-solve totals in tot.
    template <int NV>
    __device__ __forceinline__ bool seg_reduce(double (&v)[NV], const BatchSeg &sg, int nblocks,
                                               double *partials, unsigned *counter,
                                               double (&tot)[NV])
    {
      __shared__ double sh[NV][T / 64];
      __shared__ bool last;
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
        {
#pragma unroll
          for (int k = 0; k < NV; ++k)
            {
              double s = 0.0;
              for (int i = 0; i < T / 64; ++i)
                s += sh[k][i];
              store_partial(partials + k * nblocks + blockIdx.x, s);
            }
          wait_vmem();
          last = atomicAdd(counter, 1u) == (unsigned)sg.nblk - 1;
        }
      __syncthreads();
      if (!last || threadIdx.x != 0)
        return false;
#pragma unroll
      for (int k = 0; k < NV; ++k)
        {
          tot[k] = strided_sum(partials + k * nblocks, sg.blk0, sg.blk0 + sg.nblk, 1);
        }
      *counter = 0u;
      return true;
    }

    template <typename Number>
    __global__ void __launch_bounds__(T)
      batch_cg_init_kernel(const double *b, const Number *ax, Number *r, Number *p,
                           const Number *inv, const BatchSeg *seg, int nsolve, int nblocks,
                           double *partials, SolverScalars *scal, const double *rel_tol,
                           double abs_tol, int max_it, const int *active)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      const BatchSeg sg = seg[s];
      const bool on = active == nullptr || active[s] != 0;
      double v[2] = {0.0, 0.0};
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        {
          const Number ri = on ? (Number)(b[i] - (ax ? (double)ax[i] : 0.0)) : Number(0);
          const Number zi = ri * inv[i];
          r[i] = ri;
          p[i] = zi;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)zi;
        }
      double tot[2];
      if (seg_reduce<2>(v, sg, nblocks, partials, &scal[s].counter, tot))
        {
          SolverScalars &sc = scal[s];
          const double res = sqrt(tot[0]);
          sc.rr = tot[0];
          sc.rz = tot[1];
          sc.res = res;
          sc.res0 = res;
          sc.tol = fmax(abs_tol, rel_tol[s] * res);
          sc.iteration = 0;
          sc.max_it = max_it;
          sc.status = 1;
          sc.done = 0;
          if (!on || res <= sc.tol)
            {
              sc.done = 1;
              sc.status = on ? 0 : 2;
              sc.t_done = wall_clock64();
            }
        }
    }

    template <typename Number>
    __global__ void __launch_bounds__(T)
      batch_cg_update_kernel(Number *x, Number *r, const Number *p, const Number *q,
                             const Number *inv, const BatchSeg *seg, int nsolve, int nblocks,
                             double *partials, SolverScalars *scal)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      if (scal[s].done)
        return;
      const BatchSeg sg = seg[s];
      const Number a = (Number)scal[s].alpha;
      double v[2] = {0.0, 0.0};
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        {
          x[i] = x[i] + a * p[i];
          const Number ri = r[i] - a * q[i];
          r[i] = ri;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)(ri * inv[i]);
        }
      double tot[2];
      if (seg_reduce<2>(v, sg, nblocks, partials, &scal[s].counter, tot))
        {
          SolverScalars &sc = scal[s];
          const int it = sc.iteration + 1;
          const double res = sqrt(fabs(tot[0]));
          sc.iteration = it;
          sc.rr = tot[0];
          sc.res = res;
          sc.beta = tot[1] / sc.rz;
          sc.rz_old = sc.rz;
          sc.rz = tot[1];
          if (res <= sc.tol)
            {
              sc.done = 1;
              sc.status = 0;
              sc.t_done = wall_clock64();
            }
          else if (it >= sc.max_it || isnan(res))
            {
              sc.done = 1;
              sc.status = 1;
              sc.t_done = wall_clock64();
            }
        }
    }

    template <typename Number>
    __global__ void __launch_bounds__(T)
      batch_cg_direction_kernel(Number *p, const Number *r, const Number *inv,
                                const BatchSeg *seg, int nsolve, const SolverScalars *scal)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      if (scal[s].done)
        return;
      const BatchSeg sg = seg[s];
      const Number beta = (Number)scal[s].beta;
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        p[i] = r[i] * inv[i] + beta * p[i];
    }

    __global__ void __launch_bounds__(T)
      batch_residual_kernel(double *r, const double *b, const double *ax, const BatchSeg *seg,
                            int nsolve, int nblocks, double *partials, unsigned *counters,
                            double *res)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      const BatchSeg sg = seg[s];
      double v[1] = {0.0};
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        {
          const double ri = b[i] - ax[i];
          r[i] = ri;
          v[0] += ri * ri;
        }
      double tot[1];
      if (seg_reduce<1>(v, sg, nblocks, partials, counters + s, tot))
        res[s] = sqrt(tot[0]);
    }
    // One ticket per workgroup of solve sg: true in the last one to arrive
    // (every other workgroup of the solve has read its state by then).
    __device__ __forceinline__ bool seg_ticket(const BatchSeg &sg, unsigned *counter)
    {
      __shared__ bool last;
      __syncthreads();
      if (threadIdx.x == 0)
        {
          wait_vmem();
          last = atomicAdd(counter, 1u) == (unsigned)sg.nblk - 1;
          if (last)
            *counter = 0u;
        }
      __syncthreads();
      return last && threadIdx.x == 0;
    }

    __global__ void batch_begin_kernel(BatchOuter *st, SolverScalars *scal, SolverScalars *ogate,
                                       int nsolve)
    {
      const int s = blockIdx.x * blockDim.x + threadIdx.x;
      if (s >= nsolve)
        return;
      BatchOuter o{};
      o.phase = kPhaseOuter;
      o.status = 1;
      o.t_start = wall_clock64();
      st[s] = o;
      scal[s].done = 1;
      scal[s].counter = 0u;
      ogate[s].done = 0;
    }

    // The phase is read by every workgroup of a solve before its ticket and
    // changed only by the last one, so all workgroups of a solve take the same
    // branch.
    __global__ void __launch_bounds__(T)
      batch_outer_step_kernel(double *r, const double *b, const double *ax, const BatchSeg *seg,
                              int nsolve, int nblocks, double *partials, BatchOuter *st,
                              SolverScalars *ogate, double abs_tol, double rel_tol,
                              double inner_rel, int max_it)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      if (st[s].phase != kPhaseOuter)
        return;
      const BatchSeg sg = seg[s];
      double v[1] = {0.0};
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        {
          const double ri = b[i] - ax[i];
          r[i] = ri;
          v[0] += ri * ri;
        }
      double tot[1];
      if (seg_reduce<1>(v, sg, nblocks, partials, &st[s].counter, tot))
        {
          BatchOuter &o = st[s];
          const double res = sqrt(tot[0]);
          if (o.outer == 0)
            {
              o.res0 = res;
              o.tol = fmax(abs_tol, rel_tol * res);
            }
          o.res = res;
          ogate[s].done = 1;
          if (res <= o.tol || !isfinite(res) || o.inner_total >= max_it)
            {
              o.status = res <= o.tol ? 0 : 1;
              o.phase = kPhaseDone;
              o.t_done = wall_clock64();
            }
          else
            {
              // the inner solve need not go below half the outer tolerance
              o.inner_rel = fmax(inner_rel, 0.5 * o.tol / res);
              o.outer += 1;
              o.phase = kPhaseInnerStart;
            }
        }
    }

    __global__ void __launch_bounds__(T)
      batch_inner_start_kernel(const double *r, float *r32, float *p32, float *x32,
                               const float *inv, const BatchSeg *seg, int nsolve, int nblocks,
                               double *partials, BatchOuter *st, SolverScalars *scal, int max_it)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      if (st[s].phase != kPhaseInnerStart)
        return;
      const BatchSeg sg = seg[s];
      double v[2] = {0.0, 0.0};
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        {
          const float ri = (float)r[i];
          const float zi = ri * inv[i];
          r32[i] = ri;
          p32[i] = zi;
          x32[i] = 0.0f;
          v[0] += (double)ri * (double)ri;
          v[1] += (double)ri * (double)zi;
        }
      double tot[2];
      if (seg_reduce<2>(v, sg, nblocks, partials, &st[s].counter, tot))
        {
          BatchOuter &o = st[s];
          SolverScalars &sc = scal[s];
          const double res = sqrt(tot[0]);
          sc.rr = tot[0];
          sc.rz = tot[1];
          sc.res = res;
          sc.res0 = res;
          sc.tol = o.inner_rel * res;
          sc.iteration = 0;
          sc.max_it = max(1, max_it - o.inner_total);
          sc.status = 1;
          sc.done = res <= sc.tol ? 1 : 0;
          o.phase = kPhaseInner;
        }
    }

    __global__ void __launch_bounds__(T)
      batch_inner_finish_kernel(double *x, const float *x32, const BatchSeg *seg, int nsolve,
                                BatchOuter *st, const SolverScalars *scal, SolverScalars *ogate)
    {
      const int s = find_solve(seg, nsolve, blockIdx.x);
      if (st[s].phase != kPhaseInner || !scal[s].done)
        return;
      const BatchSeg sg = seg[s];
      const int64_t end = sg.off + sg.n;
      for (int64_t i = sg.off + (int64_t)(blockIdx.x - sg.blk0) * T + threadIdx.x; i < end;
           i += (int64_t)sg.nblk * T)
        x[i] += (double)x32[i];
      if (seg_ticket(sg, &st[s].counter))
        {
          st[s].inner_total += scal[s].iteration;
          st[s].phase = kPhaseOuter;
          ogate[s].done = 0;
        }
    }
  } // namespace

  hipError_t launch_batch_begin(BatchOuter *st, SolverScalars *scal, SolverScalars *ogate,
                                int nsolve, hipStream_t s)
  {
    batch_begin_kernel<<<(nsolve + 63) / 64, 64, 0, s>>>(st, scal, ogate, nsolve);
    return hipGetLastError();
  }

  hipError_t launch_batch_outer_step(double *r, const double *b, const double *ax,
                                     const BatchSeg *seg, int nsolve, int nblocks,
                                     double *partials, BatchOuter *st, SolverScalars *ogate,
                                     double abs_tol, double rel_tol, double inner_rel, int max_it,
                                     hipStream_t s)
  {
    batch_outer_step_kernel<<<nblocks, T, 0, s>>>(r, b, ax, seg, nsolve, nblocks, partials, st,
                                                  ogate, abs_tol, rel_tol, inner_rel, max_it);
    return hipGetLastError();
  }

  hipError_t launch_batch_inner_start(const double *r, float *r32, float *p32, float *x32,
                                      const float *inv, const BatchSeg *seg, int nsolve,
                                      int nblocks, double *partials, BatchOuter *st,
                                      SolverScalars *scal, int max_it, hipStream_t s)
  {
    batch_inner_start_kernel<<<nblocks, T, 0, s>>>(r, r32, p32, x32, inv, seg, nsolve, nblocks,
                                                   partials, st, scal, max_it);
    return hipGetLastError();
  }

  hipError_t launch_batch_inner_finish(double *x, const float *x32, const BatchSeg *seg,
                                       int nsolve, int nblocks, BatchOuter *st,
                                       const SolverScalars *scal, SolverScalars *ogate,
                                       hipStream_t s)
  {
    batch_inner_finish_kernel<<<nblocks, T, 0, s>>>(x, x32, seg, nsolve, st, scal, ogate);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_batch_cg_init(const double *b, const Number *ax, Number *r, Number *p,
                                  const Number *inv, const BatchSeg *seg, int nsolve, int nblocks,
                                  double *partials, SolverScalars *scal, const double *rel_tol,
                                  double abs_tol, int max_it, const int *active, hipStream_t s)
  {
    batch_cg_init_kernel<Number><<<nblocks, T, 0, s>>>(b, ax, r, p, inv, seg, nsolve, nblocks,
                                                       partials, scal, rel_tol, abs_tol, max_it,
                                                       active);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_batch_cg_update(Number *x, Number *r, const Number *p, const Number *q,
                                    const Number *inv, const BatchSeg *seg, int nsolve,
                                    int nblocks, double *partials, SolverScalars *scal,
                                    hipStream_t s)
  {
    batch_cg_update_kernel<Number><<<nblocks, T, 0, s>>>(x, r, p, q, inv, seg, nsolve, nblocks,
                                                         partials, scal);
    return hipGetLastError();
  }

  template <typename Number>
  hipError_t launch_batch_cg_direction(Number *p, const Number *r, const Number *inv,
                                       const BatchSeg *seg, int nsolve, int nblocks,
                                       const SolverScalars *scal, hipStream_t s)
  {
    batch_cg_direction_kernel<Number><<<nblocks, T, 0, s>>>(p, r, inv, seg, nsolve, scal);
    return hipGetLastError();
  }

  hipError_t launch_batch_residual(double *r, const double *b, const double *ax,
                                   const BatchSeg *seg, int nsolve, int nblocks, double *partials,
                                   double *res, hipStream_t s)
  {
    // res doubles as the per-solve ticket array: [0, nsolve) results, then counters
    unsigned *counters = reinterpret_cast<unsigned *>(res + nsolve);
    batch_residual_kernel<<<nblocks, T, 0, s>>>(r, b, ax, seg, nsolve, nblocks, partials,
                                                counters, res);
    return hipGetLastError();
  }

#define MPDG_BATCH_INST(NUM)                                                                     \
  template hipError_t launch_batch_cg_init<NUM>(const double *, const NUM *, NUM *, NUM *,       \
                                                const NUM *, const BatchSeg *, int, int,         \
                                                double *, SolverScalars *, const double *,       \
                                                double, int, const int *, hipStream_t);          \
  template hipError_t launch_batch_cg_update<NUM>(NUM *, NUM *, const NUM *, const NUM *,        \
                                                  const NUM *, const BatchSeg *, int, int,       \
                                                  double *, SolverScalars *, hipStream_t);       \
  template hipError_t launch_batch_cg_direction<NUM>(NUM *, const NUM *, const NUM *,            \
                                                     const BatchSeg *, int, int,                 \
                                                     const SolverScalars *, hipStream_t);
  MPDG_BATCH_INST(double)
  MPDG_BATCH_INST(float)
#undef MPDG_BATCH_INST
} // namespace mpdg
