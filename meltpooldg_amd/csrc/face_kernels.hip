// Robin boundary faces of the heat operator: convection alpha (T - T_inf) and
// radiation eps sigma (T^4 - T_inf^4) on faces of the box
// (source/heat/heat_diffuse_operator.cpp: tangent_boundary_loop :243-269,
// tangent_local_boundary_operation :1088-1129, the boundary operation of the
// diagonal :352-356, rhs_boundary_loop :719-781).
//
// Per face cell the reference evaluates the face values at the (p+1)^2 Gauss
// points, applies the quadrature-point physics and integrates against the face
// shape functions. The Robin part touches only the boundary layer of the
// lattice (O(N^(2/3)) nodes), so it runs as one gather launch after the cell
// loop: a thread per node of a Robin face recomputes the face cells around its
// node (one to four per face) and keeps the integral of its own shape function.
// A node on several Robin faces (an edge or corner) belongs to the lowest of
// them, which adds every Robin face through the node in ascending face order;
// each row is therefore written by exactly one thread, in a fixed order, with
// no atomics, and runs are bitwise reproducible. Face physics in FP64 for both
// vector precisions; the temperature is the FP64 solution.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    __device__ __forceinline__ bool on_face(const RobinParams &rp, const int (&I)[3], int f)
    {
      const int d = f >> 1;
      return (f & 1) ? I[d] == rp.n1[d] - 1 : I[d] == 0;
    }

    __device__ __forceinline__ bool is_dirichlet(const RobinParams &rp, const int (&I)[3])
    {
      bool c = false;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        c = c || (((rp.dirichlet >> (2 * d)) & 1u) && I[d] == 0) ||
            (((rp.dirichlet >> (2 * d + 1)) & 1u) && I[d] == rp.n1[d] - 1);
      return c;
    }

    __device__ __forceinline__ int64_t lattice(const RobinParams &rp, const int (&I)[3])
    {
      return I[0] + (int64_t)rp.n1[0] * (I[1] + (int64_t)rp.n1[1] * I[2]);
    }

    // Cells along one direction that hold lattice index i, with i's local index.
    template <int P>
    __device__ __forceinline__ int cells_of(int i, int ncell, int (&c)[2], int (&l)[2])
    {
      const int q = i / P, r = i - q * P;
      if (r != 0)
        {
          c[0] = q;
          l[0] = r;
          return 1;
        }
      int k = 0;
      if (q >= 1)
        {
          c[k] = q - 1;
          l[k] = P;
          ++k;
        }
      if (q < ncell)
        {
          c[k] = q;
          l[k] = 0;
          ++k;
        }
      return k;
    }

    // Row l of a compile-time table column-wise: out[q] = M[q][l].
    template <int N>
    __device__ __forceinline__ void column(const double (&M)[N][N], int l, double (&out)[N])
    {
#pragma unroll
      for (int q = 0; q < N; ++q)
        {
          double v = M[q][0];
#pragma unroll
          for (int j = 1; j < N; ++j)
            v = l == j ? M[q][j] : v;
          out[q] = v;
        }
    }

    // Contribution of face f's face cell (c1, c2) to the row of its local node
    // (l1, l2).
    template <int P, int MODE, typename Number>
    __device__ __forceinline__ double face_cell(const RobinParams &rp, int f, int e1, int e2,
                                                int c1, int c2, int l1, int l2, int layer,
                                                const Number *__restrict__ src,
                                                const double *__restrict__ temp)
    {
      constexpr int N = P + 1;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;
      const int d = f >> 1;
      const bool conv = (rp.conv >> f) & 1u;
      const bool rad = (rp.rad >> f) & 1u;

      // nodal values of the face cell: T (tangent/diagonal: constraint-resolved,
      // rhs: plain) and the increment u (tangent, constraint-resolved). The
      // tangent and the diagonal read T only on radiating faces (the only
      // place it enters them; without radiation no temperature is registered).
      const bool readT = MODE == kRobinRhs || rad;
      double Tn[N][N], un[N][N];
#pragma unroll
      for (int b2 = 0; b2 < N; ++b2)
#pragma unroll
        for (int b1 = 0; b1 < N; ++b1)
          {
            int L[3];
            L[d] = layer;
            L[e1] = c1 * P + b1;
            L[e2] = c2 * P + b2;
            const int64_t g = lattice(rp, L);
            const bool con = MODE != kRobinRhs && is_dirichlet(rp, L);
            Tn[b1][b2] = (con || !readT) ? 0.0 : temp[g];
            if constexpr (MODE == kRobinTangent)
              un[b1][b2] = con ? 0.0 : (double)src[g];
          }

      // A[q1][b2] = sum_b1 S[q1][b1] v[b1][b2]; v(q1, q2) = sum_b2 S[q2][b2] A[q1][b2]
      double AT[N][N], AU[N][N];
#pragma unroll
      for (int q1 = 0; q1 < N; ++q1)
#pragma unroll
        for (int b2 = 0; b2 < N; ++b2)
          {
            double t = 0.0, u = 0.0;
#pragma unroll
            for (int b1 = 0; b1 < N; ++b1)
              {
                t += T1.S[q1][b1] * Tn[b1][b2];
                if constexpr (MODE == kRobinTangent)
                  u += T1.S[q1][b1] * un[b1][b2];
              }
            AT[q1][b2] = t;
            AU[q1][b2] = u;
          }

      double s1[N], s2[N];
      column<N>(T1.S, l1, s1);
      column<N>(T1.S, l2, s2);
      const double area = rp.h[e1] * rp.h[e2];
      double y = 0.0;
#pragma unroll
      for (int q1 = 0; q1 < N; ++q1)
        {
          double row = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < N; ++q2)
            {
              double Tq = 0.0, uq = 0.0;
#pragma unroll
              for (int b2 = 0; b2 < N; ++b2)
                {
                  Tq += T1.S[q2][b2] * AT[q1][b2];
                  if constexpr (MODE == kRobinTangent)
                    uq += T1.S[q2][b2] * AU[q1][b2];
                }
              double val = 0.0;
              if constexpr (MODE == kRobinRhs)
                {
                  const double Ti = rp.t_inf_rad;
                  if (rad)
                    val -= rp.eps_sigma * ((Tq * Tq) * (Tq * Tq) - (Ti * Ti) * (Ti * Ti));
                  if (conv)
                    val -= rp.alpha * (Tq - rp.t_inf_conv);
                }
              else
                {
                  // tangent coefficient alpha + 4 eps sigma T^3; the diagonal
                  // takes the unit vector of the node itself as increment
                  const double u = MODE == kRobinTangent ? uq : s1[q1] * s2[q2];
                  if (conv)
                    val += rp.alpha * u;
                  if (rad)
                    val += 4. * rp.eps_sigma * (Tq * Tq * Tq) * u;
                }
              row += s2[q2] * T1.w[q2] * val;
            }
          y += s1[q1] * T1.w[q1] * row;
        }
      return y * area;
    }

    // Sum over the face cells of face f around node I.
    template <int P, int MODE, typename Number>
    __device__ __forceinline__ double face_node(const RobinParams &rp, int f, const int (&I)[3],
                                                const Number *__restrict__ src,
                                                const double *__restrict__ temp)
    {
      const int d = f >> 1;
      const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
      int ca[2], la[2], cb[2], lb[2];
      const int na = cells_of<P>(I[e1], rp.n[e1], ca, la);
      const int nb = cells_of<P>(I[e2], rp.n[e2], cb, lb);
      double y = 0.0;
      for (int ib = 0; ib < nb; ++ib)
        for (int ia = 0; ia < na; ++ia)
          y += face_cell<P, MODE, Number>(rp, f, e1, e2, ca[ia], cb[ib], la[ia], lb[ib], I[d], src,
                                          temp);
      return y;
    }

    template <int P, int MODE, typename Number>
    __global__ void __launch_bounds__(kRobinThreads)
      robin_kernel(Number *__restrict__ dst, const Number *__restrict__ src,
                   const double *__restrict__ temp, const RobinParams rp,
                   double *__restrict__ partials, SolverScalars *__restrict__ scal)
    {
      constexpr bool DOT = MODE == kRobinTangent;
      if (DOT && scal != nullptr && scal->done)
        return; // uniform over the grid: done is not written during the launch
      const int64_t tid = (int64_t)blockIdx.x * kRobinThreads + threadIdx.x;
      double pdot = 0.0;
      if (tid < rp.node_off[rp.nfaces])
        {
          int k = 0;
          while (k + 1 < rp.nfaces && tid >= rp.node_off[k + 1])
            ++k;
          const int f = rp.face[k];
          const int d = f >> 1;
          const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
          const int64_t loc = tid - rp.node_off[k];
          int I[3];
          I[d] = (f & 1) ? rp.n1[d] - 1 : 0;
          I[e1] = (int)(loc % rp.n1[e1]);
          I[e2] = (int)(loc / rp.n1[e1]);
          bool owned = true;
          for (int g = 0; g < f; ++g)
            if (((rp.robin >> g) & 1u) && on_face(rp, I, g))
              owned = false;
          if (owned && !is_dirichlet(rp, I))
            {
              double y = 0.0;
              for (int g = f; g < 6; ++g)
                if (((rp.robin >> g) & 1u) && on_face(rp, I, g))
                  y += face_node<P, MODE, Number>(rp, g, I, src, temp);
              const int64_t gi = lattice(rp, I);
              if constexpr (DOT)
                pdot = (double)src[gi] * y;
              dst[gi] = (Number)((double)dst[gi] + y);
            }
        }
      if constexpr (DOT)
        {
          if (scal == nullptr)
            return;
          // p.(A_face p): block sums, then the last workgroup adds them in
          // index order to the cell loop's p.Ap and resets alpha
          __shared__ double red[kRobinThreads / 64];
          __shared__ bool is_last;
          const double bs = block_sum<kRobinThreads>(pdot, red);
          if (threadIdx.x == 0)
            {
              partials[blockIdx.x] = bs;
              __threadfence();
              const unsigned prev = atomicAdd(&scal->counter, 1u);
              is_last = prev == gridDim.x - 1;
            }
          __syncthreads();
          if (is_last)
            {
              __threadfence();
              const double s = ordered_sum<kRobinThreads>(partials, gridDim.x, red);
              if (threadIdx.x == 0)
                {
                  const double pap = scal->pap + s;
                  scal->pap = pap;
                  scal->alpha = scal->rz / pap;
                  scal->counter = 0u;
                }
            }
        }
    }

    template <int P, typename Number>
    hipError_t robin_deg(int mode, const RobinParams &rp, Number *dst, const Number *src,
                         const double *temp, double *partials, SolverScalars *scal,
                         hipStream_t s)
    {
      const int blocks = robin_blocks(rp);
      if (blocks == 0)
        return hipSuccess;
      switch (mode)
        {
          case kRobinTangent:
            robin_kernel<P, kRobinTangent, Number><<<blocks, kRobinThreads, 0, s>>>(
              dst, src, temp, rp, partials, scal);
            break;
          case kRobinDiagonal:
            robin_kernel<P, kRobinDiagonal, Number><<<blocks, kRobinThreads, 0, s>>>(
              dst, src, temp, rp, nullptr, nullptr);
            break;
          case kRobinRhs:
            robin_kernel<P, kRobinRhs, Number><<<blocks, kRobinThreads, 0, s>>>(
              dst, src, temp, rp, nullptr, nullptr);
            break;
          default: return hipErrorInvalidValue;
        }
      return hipGetLastError();
    }
  } // namespace

  int robin_blocks(const RobinParams &rp)
  {
    const int64_t nodes = rp.nfaces > 0 ? rp.node_off[rp.nfaces] : 0;
    return (int)((nodes + kRobinThreads - 1) / kRobinThreads);
  }

  template <typename Number>
  hipError_t launch_robin(int degree, int mode, const RobinParams &rp, Number *dst,
                          const Number *src, const double *temperature, double *partials,
                          SolverScalars *scal, hipStream_t stream)
  {
    if (rp.nfaces == 0)
      return hipSuccess;
    switch (degree)
      {
        case 1: return robin_deg<1, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        case 2: return robin_deg<2, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        case 3: return robin_deg<3, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        case 4: return robin_deg<4, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        case 5: return robin_deg<5, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        case 6: return robin_deg<6, Number>(mode, rp, dst, src, temperature, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  template hipError_t launch_robin<double>(int, int, const RobinParams &, double *,
                                           const double *, const double *, double *,
                                           SolverScalars *, hipStream_t);
  template hipError_t launch_robin<float>(int, int, const RobinParams &, float *, const float *,
                                          const double *, double *, SolverScalars *, hipStream_t);
} // namespace mpdg
