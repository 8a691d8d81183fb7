// reduce.hpp - fixed-order device reductions shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

namespace mpdg
{
  namespace
  {
    // vmcnt(0): this wavefront's vector memory operations have completed.
    __device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

    // Ticket protocol of the last-block reductions: every block stores its
    // partial with an agent-coherent store, waits until the store has
    // completed (wait_vmem), then takes a ticket (atomicAdd on the counter).
    // The block that draws the last ticket reads all partials with
    // agent-coherent loads (strided_sum). No cache-wide write-back or
    // invalidate fence is needed: only the partials cross workgroups.
    __device__ __forceinline__ void store_partial(double *p, double v)
    {
      __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // p[i0] + p[i0 + stride] + ... (indices < count), added in that order, with
    // the agent-coherent loads issued sixteen at a time: a plain loop waits for
    // every load before the next add, a chain of memory latencies. (Adding 0.0
    // for the missing tail entries changes nothing: the sum starts at +0.0, so
    // it is never -0.0.)
    __device__ __forceinline__ double strided_sum(const double *p, int i0, int count, int stride)
    {
      constexpr int kBatch = 16;
      double v = 0.0;
      for (int i = i0; i < count; i += kBatch * stride)
        {
          double t[kBatch];
#pragma unroll
          for (int u = 0; u < kBatch; ++u)
            {
              const int k = i + u * stride;
              t[u] = k < count ? __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : 0.0;
            }
#pragma unroll
          for (int u = 0; u < kBatch; ++u)
            v += t[u];
        }
      return v;
    }
  } // namespace
} // namespace mpdg
