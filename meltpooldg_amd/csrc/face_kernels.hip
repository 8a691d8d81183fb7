// Robin boundary faces of the heat operator: convection alpha (T - T_inf) and
// radiation eps sigma (T^4 - T_inf^4) on faces of the box
// (source/heat/heat_diffuse_operator.cpp: tangent_boundary_loop :243-269,
// tangent_local_boundary_operation :1088-1129, the boundary operation of the
// diagonal :352-356, rhs_boundary_loop :719-781).
//
// Per face cell the reference evaluates the face values at the (p+1)^2 Gauss
// points, applies the quadrature-point physics and integrates against the face
// shape functions. The Robin part touches only the boundary layer of the
// lattice (O(N^(2/3)) nodes) and runs after the cell loop as two launches:
//   1. face_cell_kernel: one 64-lane workgroup per face cell. The lanes stage
//      the (p+1)^2 nodal values in LDS, then each lane evaluates one Gauss
//      point, then each lane integrates against one shape function; the row of
//      (p+1)^2 results goes to a scratch array (one row per face cell).
//   2. face_node_kernel: one thread per node of a Robin face adds the rows of
//      the face cells around it (one to four per face) into dst. A node on
//      several Robin faces (an edge or corner) belongs to the lowest of them,
//      which adds every Robin face through the node in ascending face order.
// Each row of dst is written by exactly one thread in a fixed order, with no
// atomics, so runs are bitwise reproducible. Face physics in FP64 for both
// vector precisions; the temperature is the FP64 solution.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    constexpr int kFaceCellThreads = 64; // (p+1)^2 <= 49 lanes active

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

    __device__ __forceinline__ void tangential(int f, int &e1, int &e2)
    {
      const int d = f >> 1;
      e1 = d == 0 ? 1 : 0;
      e2 = d == 2 ? 1 : 2;
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

    // One face cell per workgroup: scratch[cell][a1 + N a2] =
    //   tangent:  int_F phi_a (alpha + 4 eps sigma T^3) u
    //   diagonal: int_F phi_a^2 (alpha + 4 eps sigma T^3)
    //   rhs:      -int_F phi_a (eps sigma (T^4 - T_inf^4) + alpha (T - T_inf) + q . n)
    // with T and u read constraint-resolved (tangent, diagonal) or T plain (rhs).
    template <int P, int MODE, typename Number>
    __global__ void __launch_bounds__(kFaceCellThreads)
      face_cell_kernel(const Number *__restrict__ src, const double *__restrict__ temp,
                       const RobinParams rp, double *__restrict__ scratch,
                       const SolverScalars *__restrict__ scal)
    {
      constexpr int N = P + 1;
      constexpr int NN = N * N;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;
      if (MODE == kRobinTangent && scal != nullptr && scal->done)
        return;
      __shared__ double sS[N][N], sW[N], sT[NN], sU[NN], sV[NN];
      const int l = threadIdx.x;
      if (l == 0)
        {
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              sW[q] = T1.w[q];
#pragma unroll
              for (int b = 0; b < N; ++b)
                sS[q][b] = T1.S[q][b];
            }
        }
      const int64_t cell = blockIdx.x;
      int k = 0;
      while (k + 1 < rp.nfaces && cell >= rp.cell_off[k + 1])
        ++k;
      const int f = rp.face[k];
      const int d = f >> 1;
      int e1, e2;
      tangential(f, e1, e2);
      const int64_t loc = cell - rp.cell_off[k];
      const int c1 = (int)(loc % rp.n[e1]);
      const int c2 = (int)(loc / rp.n[e1]);
      const bool conv = (rp.conv >> f) & 1u;
      const bool rad = (rp.rad >> f) & 1u;
      // the tangent and the diagonal read T only on radiating faces (without
      // radiation no temperature is registered)
      const bool readT = MODE == kRobinRhs || rad;
      const int i1 = l % N, i2 = l / N;
      if (l < NN)
        {
          int L[3];
          L[d] = (f & 1) ? rp.n1[d] - 1 : 0;
          L[e1] = c1 * P + i1;
          L[e2] = c2 * P + i2;
          const int64_t g = lattice(rp, L);
          const bool con = MODE != kRobinRhs && is_dirichlet(rp, L);
          sT[l] = (con || !readT) ? 0.0 : temp[g];
          if constexpr (MODE == kRobinTangent)
            sU[l] = con ? 0.0 : (double)src[g];
        }
      __syncthreads();
      // lane = Gauss point (q1, q2) = (i1, i2): values, physics, JxW
      if (l < NN)
        {
          double Tq = 0.0, uq = 0.0;
#pragma unroll
          for (int b2 = 0; b2 < N; ++b2)
            {
              double t = 0.0, u = 0.0;
#pragma unroll
              for (int b1 = 0; b1 < N; ++b1)
                {
                  t += sS[i1][b1] * sT[b1 + N * b2];
                  if constexpr (MODE == kRobinTangent)
                    u += sS[i1][b1] * sU[b1 + N * b2];
                }
              Tq += sS[i2][b2] * t;
              uq += sS[i2][b2] * u;
            }
          double val = 0.0;
          if constexpr (MODE == kRobinRhs)
            {
              // Neumann face: -q . n (a constant per face)
              if ((rp.neumann >> f) & 1u)
                val -= rp.qn[f];
              const double Ti = rp.t_inf_rad;
              if (rad)
                val -= rp.eps_sigma * ((Tq * Tq) * (Tq * Tq) - (Ti * Ti) * (Ti * Ti));
              if (conv)
                val -= rp.alpha * (Tq - rp.t_inf_conv);
            }
          else
            {
              // tangent coefficient alpha + 4 eps sigma T^3 (times u for the
              // tangent; the diagonal applies it to each node's own shape function)
              const double u = MODE == kRobinTangent ? uq : 1.0;
              if (conv)
                val += rp.alpha * u;
              if (rad)
                val += 4. * rp.eps_sigma * (Tq * Tq * Tq) * u;
            }
          sV[l] = val * sW[i1] * sW[i2] * (rp.h[e1] * rp.h[e2]);
        }
      __syncthreads();
      // lane = node (a1, a2) = (i1, i2): integrate against its shape function
      if (l < NN)
        {
          double y = 0.0;
#pragma unroll
          for (int q2 = 0; q2 < N; ++q2)
            {
              double r = 0.0;
#pragma unroll
              for (int q1 = 0; q1 < N; ++q1)
                {
                  const double s = sS[q1][i1];
                  r += (MODE == kRobinDiagonal ? s * s : s) * sV[q1 + N * q2];
                }
              const double s = sS[q2][i2];
              y += (MODE == kRobinDiagonal ? s * s : s) * r;
            }
          scratch[cell * NN + l] = y;
        }
    }

    // Sum of face f's scratch rows at node I (face cells around I in a fixed order).
    template <int P>
    __device__ __forceinline__ double face_rows(const RobinParams &rp, int f,
                                                const int (&I)[3],
                                                const double *__restrict__ scratch)
    {
      constexpr int N = P + 1;
      int e1, e2;
      tangential(f, e1, e2);
      const int k = __popc(rp.robin & ((1u << f) - 1u)); // position of f among the Robin faces
      int ca[2], la[2], cb[2], lb[2];
      const int na = cells_of<P>(I[e1], rp.n[e1], ca, la);
      const int nb = cells_of<P>(I[e2], rp.n[e2], cb, lb);
      double y = 0.0;
      for (int ib = 0; ib < nb; ++ib)
        for (int ia = 0; ia < na; ++ia)
          {
            const int64_t cell = rp.cell_off[k] + ca[ia] + (int64_t)rp.n[e1] * cb[ib];
            y += scratch[cell * (N * N) + la[ia] + N * lb[ib]];
          }
      return y;
    }

    template <int P, int MODE, typename Number>
    __global__ void __launch_bounds__(kRobinThreads)
      face_node_kernel(Number *__restrict__ dst, const Number *__restrict__ src,
                       const RobinParams rp, const double *__restrict__ scratch,
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
          int e1, e2;
          tangential(f, e1, e2);
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
                  y += face_rows<P>(rp, g, I, scratch);
              const int64_t gi = lattice(rp, I);
              if constexpr (DOT)
                pdot = (double)src[gi] * y;
              dst[gi] = (Number)((double)dst[gi] + y);
            }
        }
      if constexpr (DOT)
        {
          if (scal == nullptr || partials == nullptr)
            return;
          // p.(A_face p): block sums, then the last workgroup adds them in
          // index order to the cell loop's p.Ap and resets alpha
          __shared__ double red[kRobinThreads / 64];
          __shared__ bool is_last;
          const double bs = block_sum<kRobinThreads>(pdot, red);
          if (threadIdx.x == 0)
            {
              store_partial(partials + blockIdx.x, bs);
              wait_vmem();
              const unsigned prev = atomicAdd(&scal->counter, 1u);
              is_last = prev == gridDim.x - 1;
            }
          __syncthreads();
          if (is_last)
            {
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

    template <int P, int MODE, typename Number>
    hipError_t robin_mode(const RobinParams &rp, Number *dst, const Number *src,
                          const double *temp, double *scratch, double *partials,
                          SolverScalars *scal, hipStream_t s)
    {
      const int64_t cells = rp.cell_off[rp.nfaces];
      face_cell_kernel<P, MODE, Number><<<(unsigned)cells, kFaceCellThreads, 0, s>>>(
        src, temp, rp, scratch, scal);
      face_node_kernel<P, MODE, Number><<<robin_blocks(rp), kRobinThreads, 0, s>>>(
        dst, src, rp, scratch, partials, scal);
      return hipGetLastError();
    }

    template <int P, typename Number>
    hipError_t robin_deg(int mode, const RobinParams &rp, Number *dst, const Number *src,
                         const double *temp, double *scratch, double *partials,
                         SolverScalars *scal, hipStream_t s)
    {
      switch (mode)
        {
          case kRobinTangent:
            return robin_mode<P, kRobinTangent, Number>(rp, dst, src, temp, scratch, partials,
                                                        scal, s);
          case kRobinDiagonal:
            return robin_mode<P, kRobinDiagonal, Number>(rp, dst, src, temp, scratch, nullptr,
                                                         nullptr, s);
          case kRobinRhs:
            return robin_mode<P, kRobinRhs, Number>(rp, dst, src, temp, scratch, nullptr,
                                                    nullptr, s);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  int robin_blocks(const RobinParams &rp)
  {
    const int64_t nodes = rp.nfaces > 0 ? rp.node_off[rp.nfaces] : 0;
    return (int)((nodes + kRobinThreads - 1) / kRobinThreads);
  }

  int64_t robin_scratch(const RobinParams &rp, int degree)
  {
    return rp.nfaces > 0 ? rp.cell_off[rp.nfaces] * (int64_t)(degree + 1) * (degree + 1) : 0;
  }

  template <typename Number>
  hipError_t launch_robin(int degree, int mode, const RobinParams &rp, Number *dst,
                          const Number *src, const double *temperature, double *scratch,
                          double *partials, SolverScalars *scal, hipStream_t stream)
  {
    if (rp.nfaces == 0)
      return hipSuccess;
    switch (degree)
      {
        case 1: return robin_deg<1, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        case 2: return robin_deg<2, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        case 3: return robin_deg<3, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        case 4: return robin_deg<4, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        case 5: return robin_deg<5, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        case 6: return robin_deg<6, Number>(mode, rp, dst, src, temperature, scratch, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  template hipError_t launch_robin<double>(int, int, const RobinParams &, double *,
                                           const double *, const double *, double *, double *,
                                           SolverScalars *, hipStream_t);
  template hipError_t launch_robin<float>(int, int, const RobinParams &, float *, const float *,
                                          const double *, double *, double *, SolverScalars *,
                                          hipStream_t);
} // namespace mpdg
