// Residual right-hand side of the heat operator (create_rhs), FP64.
//
// dst += -( rho_cp/dt M (T_new - T_old) - M_1 f + K_k T_new + rho_cp C_u T_new )
// on the free rows, with T_new, T_old and the heat source f read plain
// (constrained values included), as source/heat/heat_diffuse_operator.cpp
// create_rhs :783-806 / rhs_cell_loop :393-529 do without evaporation, the
// material evaluated at the heaviside and, for liquid_solid / gas_liquid_solid,
// at T_new(q) (get_material_parameters :1133-1165): at q
//   val      = dt^-1 rho_cp (T_new - T_old) - f  (+ rho_cp grad T_new . u),
//   val_grad = k grad T_new,
// submitted with the factor -1 and integrated against values and gradients.
//
// It runs once per Newton step, so it keeps the simple block-synchronous form:
// one workgroup per row segment of CPB cells (thread = cell x x-line), the 1D
// contractions in registers, LDS between line directions with workgroup
// barriers, shared x-faces summed inside the workgroup and every other shared
// node by eight ordered colour passes. All passes add into dst (the
// reference's loop does not zero dst), and constrained rows are never written.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    template <int P, int COEF, int VM, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      rhs_kernel(double *__restrict__ dst, const double *__restrict__ tnew,
                 const double *__restrict__ told, const double *__restrict__ fsrc,
                 const double *__restrict__ heav, const double *__restrict__ vel,
                 const CellParams<double> prm, const int colour)
    {
      using Number = double;
      constexpr bool VEL = VM != kVelNone;
      constexpr bool NV = VM == kVelNodal;
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      // QC: rho_cp and k vary per quadrature point (staged in KB / RB)
      constexpr bool QC = COEF == kCoefHeaviside || COEF == kCoefTemperature;
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      constexpr ct::EO<N> TSt = ct::VTab<P>::St;
      constexpr ct::EO<N> TC = ct::VTab<P>::C;
      constexpr ct::EO<N> TCt = ct::VTab<P>::Ct;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;

      __shared__ Number A[CPB * STR];  // T_new (nodal -> q)
      __shared__ Number Bd[CPB * STR]; // T_new - T_old; later the q-space accumulator
      __shared__ Number Cf[CPB * STR]; // heat source
      __shared__ Number KB[QC ? CPB * STR : 1];
      __shared__ Number RB[QC ? CPB * STR : 1];
      __shared__ Number XE[CPB * NL];
      // nodal velocity at q (kVelNodal)
      __shared__ Number VX[NV ? CPB * STR : 1], VY[NV ? CPB * STR : 1], VZ[NV ? CPB * STR : 1];

      const Segment g = segment_of<Number, CPB>(prm, colour);
      if constexpr (NV)
        nodal_vector_at_q<P, CPB>(prm, g, vel, VX, VY, VZ);
      const int vo = NV ? (int)(threadIdx.x % CPB) * STR : 0;
      const Number *VXl = VX + vo, *VYl = VY + vo, *VZl = VZ + vo;
      // the heaviside is read for gas_liquid and gas_liquid_solid
      const bool use_h = COEF == kCoefHeaviside || (COEF == kCoefTemperature && prm.tmat == 3);
      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int cx = min(g.cx0 + l, prm.nx - 1);
      Number *Al = A + l * STR;
      Number *Bl = Bd + l * STR;
      Number *Cl = Cf + l * STR;
      Number *KBl = KB + (QC ? l * STR : 0);
      Number *RBl = RB + (QC ? l * STR : 0);

      // ---- x-lines (y = s0, z = s1): plain gathers, interpolate in x ----------
      Number dq[N], fq[N];
      if (active)
        {
          const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
          const int64_t base = (int64_t)(g.cx0 + l) * P + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);
          Number tn[N], dl[N], fl[N], hl[N], o[N];
#pragma unroll
          for (int a = 0; a < N; ++a)
            {
              tn[a] = cell_ok ? tnew[base + a] : Number(0);
              dl[a] = cell_ok ? tn[a] - told[base + a] : Number(0);
              fl[a] = (cell_ok && fsrc) ? fsrc[base + a] : Number(0);
              hl[a] = (QC && use_h && cell_ok) ? heav[base + a] : Number(0);
            }
          const int row = (s1 * N + s0) * N;
          eo_apply<N, 1>(TS, tn, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Al[row + q] = o[q];
          eo_apply<N, 1>(TS, dl, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Bl[row + q] = o[q];
          eo_apply<N, 1>(TS, fl, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Cl[row + q] = o[q];
          if constexpr (QC)
            if (use_h)
              {
                eo_apply<N, 1>(TS, hl, o);
#pragma unroll
                for (int q = 0; q < N; ++q)
                  KBl[row + q] = o[q];
              }
        }
      __syncthreads();

      // ---- y-lines (qx = s0, z = s1): interpolate in y ----------------------------
      if (active)
        {
          auto sweep = [&](Number *T) {
            Number v[N], o[N];
#pragma unroll
            for (int b = 0; b < N; ++b)
              v[b] = T[(s1 * N + b) * N + s0];
            eo_apply<N, 1>(TS, v, o);
#pragma unroll
            for (int b = 0; b < N; ++b)
              T[(s1 * N + b) * N + s0] = o[b];
          };
          sweep(Al);
          sweep(Bl);
          sweep(Cl);
          if constexpr (QC)
            if (use_h)
              sweep(KBl);
        }
      __syncthreads();

      // ---- z-lines (qx = s0, qy = s1): values, z-gradient, material at q ---------
      Number U[N], Gz[N];
      if (active)
        {
          Number v[N];
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Al[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, U);
          eo_apply<N, -1>(TC, U, Gz);
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Bl[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, dq);
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Cl[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, fq);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Al[(q * N + s1) * N + s0] = U[q];
          if constexpr (QC)
            {
              Number Hq[N];
              if (use_h)
                {
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    v[c] = KBl[(c * N + s1) * N + s0];
                  eo_apply<N, 1>(TS, v, Hq);
                }
              else
                {
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    Hq[q] = Number(0);
                }
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  Number rc, k;
                  if constexpr (COEF == kCoefTemperature)
                    {
                      Number drc, dk;
                      solid_liquid_material(prm, Hq[q], U[q], rc, k, drc, dk);
                    }
                  else
                    two_phase(prm, Hq[q], rc, k);
                  KBl[(q * N + s1) * N + s0] = k;
                  RBl[(q * N + s1) * N + s0] = rc;
                }
            }
        }
      __syncthreads();

      // ---- x-lines (qy = s0, qz = s1): -x-flux, -x-convection -> Bd ------------
      if (active)
        {
          Number v[N], gx[N], f[N], o[N], cv[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + s0) * N + q];
          eo_apply<N, -1>(TC, v, gx);
          const Number wyz = Number(T1.w[s0] * T1.w[s1]);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const Number w3 = Number(T1.w[q]) * wyz;
              const Number k = conductivity<Number, COEF, N>(prm, KBl, cx, g.cy, g.cz, q, s0, s1);
              f[q] = -k * w3 * prm.kf[0] * gx[q];
              if constexpr (VEL)
                {
                  const Number rc = QC ? RBl[(s1 * N + s0) * N + q] : prm.rho_cp0;
                  const Number vx = NV ? VXl[(s1 * N + s0) * N + q] :
                                         prm.vel_x[0][cx * N + q] * prm.vel_y[0][g.cy * N + s0];
                  cv[q] = -rc * w3 * prm.vf[0] * vx * gx[q];
                }
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Bl[(s1 * N + s0) * N + q] = VEL ? o[q] + cv[q] : o[q];
        }
      __syncthreads();

      // ---- y-lines (qx = s0, qz = s1): -y-flux, -y-convection += Bd --------------
      if (active)
        {
          Number v[N], gy[N], f[N], o[N], cv[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + q) * N + s0];
          eo_apply<N, -1>(TC, v, gy);
          const Number wxz = Number(T1.w[s0] * T1.w[s1]);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const Number w3 = Number(T1.w[q]) * wxz;
              const Number k = conductivity<Number, COEF, N>(prm, KBl, cx, g.cy, g.cz, s0, q, s1);
              f[q] = -k * w3 * prm.kf[1] * gy[q];
              if constexpr (VEL)
                {
                  const Number rc = QC ? RBl[(s1 * N + q) * N + s0] : prm.rho_cp0;
                  const Number vy = NV ? VYl[(s1 * N + q) * N + s0] :
                                         prm.vel_x[1][cx * N + s0] * prm.vel_y[1][g.cy * N + q];
                  cv[q] = -rc * w3 * prm.vf[1] * vy * gy[q];
                }
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (s1 * N + q) * N + s0;
              Bl[idx] = Bl[idx] + (VEL ? o[q] + cv[q] : o[q]);
            }
        }
      __syncthreads();

      // ---- z-lines (qx = s0, qy = s1): -z-flux, -value terms, integrate z -------
      if (active)
        {
          Number f[N], o[N], V[N], y1[N];
          const Number wxy = Number(T1.w[s0] * T1.w[s1]);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const Number w3 = Number(T1.w[q]) * wxy;
              const Number k = conductivity<Number, COEF, N>(prm, KBl, cx, g.cy, g.cz, s0, s1, q);
              f[q] = -k * w3 * prm.kf[2] * Gz[q];
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const Number w3 = Number(T1.w[q]) * wxy;
              const Number rc = QC ? RBl[(q * N + s1) * N + s0] : prm.rho_cp0;
              const Number mass = QC ? rc * prm.dtinv_vol : prm.mass;
              Number val = -(mass * w3 * dq[q] - prm.vol * w3 * fq[q]);
              if constexpr (VEL)
                val -= rc * w3 * prm.vf[2] * (NV ? VZl[(q * N + s1) * N + s0] : prm.w0) * Gz[q];
              V[q] = val + o[q] + Bl[(q * N + s1) * N + s0];
            }
          eo_apply<N, 1>(TSt, V, y1);
#pragma unroll
          for (int c = 0; c < N; ++c)
            Al[(c * N + s1) * N + s0] = y1[c];
        }
      __syncthreads();

      // ---- y-lines (qx = s0, c = s1): integrate y ----------------------------------
      if (active)
        {
          Number v[N], o[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + q) * N + s0];
          eo_apply<N, 1>(TSt, v, o);
#pragma unroll
          for (int b = 0; b < N; ++b)
            Al[(s1 * N + b) * N + s0] = o[b];
        }
      __syncthreads();

      // ---- x-lines (b = s0, c = s1): integrate x, x-face exchange, add -----------
      Number Y[N];
      if (active)
        {
          Number v[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + s0) * N + q];
          eo_apply<N, 1>(TSt, v, Y);
          XE[s * CPB + l] = Y[N - 1];
        }
      __syncthreads();
      if (cell_ok)
        {
          if (l > 0)
            Y[0] += XE[s * CPB + l - 1];
          const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
          const bool rowc = on_dirichlet(prm, Yi, 1) || on_dirichlet(prm, Zi, 2);
          const int X0 = (g.cx0 + l) * P;
          const int64_t base = X0 + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);
          const bool last = l == g.ncell - 1;
          if (!rowc)
#pragma unroll
            for (int a = 0; a < N; ++a)
              {
                if (a == N - 1 && !last)
                  continue;
                if (on_dirichlet(prm, X0 + a, 0))
                  continue;
                dst[base + a] += Y[a];
              }
        }
    }

    // Squared L2 norm of the FE field u: one workgroup per row segment of CPB
    // cells (all segments in one launch), values interpolated to the Gauss
    // points through LDS, JxW u^2 summed per workgroup into partials[block];
    // the workgroup that finishes last (ticket) adds the partials in block
    // order into *out, so the result is reproducible.
    template <int P, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      l2_norm_kernel(const double *__restrict__ u, const CellParams<double> prm,
                     double *__restrict__ partials, unsigned *__restrict__ ticket,
                     double *__restrict__ out)
    {
      using Number = double;
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;
      __shared__ Number A[CPB * STR];
      __shared__ double red[THREADS / 64];
      __shared__ bool is_last;

      const int nsx = (prm.nx + CPB - 1) / CPB;
      int b = blockIdx.x;
      const int sx = b % nsx;
      b /= nsx;
      const int cy = b % prm.ny;
      const int cz = b / prm.ny;
      const int cx0 = sx * CPB;
      const int ncell = min(CPB, prm.nx - cx0);
      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < ncell;
      Number *Al = A + l * STR;

      if (active)
        {
          const int64_t base = (int64_t)(cx0 + l) * P +
                               (int64_t)prm.n1x * (cy * P + s0 + (int64_t)prm.n1y * (cz * P + s1));
          Number v[N], o[N];
#pragma unroll
          for (int a = 0; a < N; ++a)
            v[a] = cell_ok ? u[base + a] : Number(0);
          eo_apply<N, 1>(TS, v, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Al[(s1 * N + s0) * N + q] = o[q];
        }
      __syncthreads();
      if (active)
        {
          Number v[N], o[N];
#pragma unroll
          for (int b2 = 0; b2 < N; ++b2)
            v[b2] = Al[(s1 * N + b2) * N + s0];
          eo_apply<N, 1>(TS, v, o);
#pragma unroll
          for (int b2 = 0; b2 < N; ++b2)
            Al[(s1 * N + b2) * N + s0] = o[b2];
        }
      __syncthreads();
      double acc = 0.0;
      if (active)
        {
          Number v[N], o[N];
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Al[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, o);
          const double wxy = T1.w[s0] * T1.w[s1] * (double)prm.vol;
#pragma unroll
          for (int q = 0; q < N; ++q)
            acc += wxy * T1.w[q] * o[q] * o[q];
        }
      const double bs = block_sum<THREADS>(acc, red);
      if (threadIdx.x == 0)
        {
          store_partial(partials + blockIdx.x, bs);
          wait_vmem();
          is_last = atomicAdd(ticket, 1u) == gridDim.x - 1;
        }
      __syncthreads();
      if (is_last)
        {
          const double total = ordered_sum<THREADS>(partials, gridDim.x, red);
          if (threadIdx.x == 0)
            {
              *out = total;
              *ticket = 0u;
            }
        }
    }

    template <int P>
    hipError_t l2_norm_p(const CellParams<double> &prm, const double *u, double *partials,
                         unsigned *ticket, double *out, hipStream_t s)
    {
      constexpr int CPB = cells_per_block_for(P, false);
      constexpr int THREADS = ((CPB * (P + 1) * (P + 1) + 63) / 64) * 64;
      const int blocks = ((prm.nx + CPB - 1) / CPB) * prm.ny * prm.nz;
      l2_norm_kernel<P, CPB, THREADS><<<blocks, THREADS, 0, s>>>(u, prm, partials, ticket, out);
      return hipGetLastError();
    }

    // cells per workgroup: three more cell tensors in LDS with a nodal
    // velocity (then half as many cells from degree 4 on)
    template <int P, int VM>
    constexpr int rhs_cpb()
    {
      return cells_per_block_for(P, false) / (VM == kVelNodal && P >= 4 ? 2 : 1);
    }

    template <int P, int COEF, int VM>
    hipError_t rhs_passes(const CellParams<double> &prm, double *dst, const double *tn,
                          const double *to, const double *f, const double *heav,
                          const double *vel, hipStream_t s)
    {
      constexpr int CPB = rhs_cpb<P, VM>();
      constexpr int THREADS = ((CPB * (P + 1) * (P + 1) + 63) / 64) * 64;
      const ColourGrid cg = colour_grid(prm.nx, prm.ny, prm.nz, CPB);
      for (int c = 0; c < 8; ++c)
        {
          if (cg.blocks[c] == 0)
            continue;
          rhs_kernel<P, COEF, VM, CPB, THREADS><<<cg.blocks[c], THREADS, 0, s>>>(dst, tn, to, f,
                                                                                heav, vel, prm, c);
        }
      return hipGetLastError();
    }

    template <int P>
    hipError_t rhs_deg(int coef, int vmode, const CellParams<double> &prm, double *dst,
                       const double *tn, const double *to, const double *f, const double *heav,
                       const double *vel, hipStream_t s)
    {
      // the separable-sine conductivity has no nodal-velocity form (creation
      // refuses that combination)
      switch (coef * 3 + vmode)
        {
          case 0: return rhs_passes<P, 0, kVelNone>(prm, dst, tn, to, f, heav, vel, s);
          case 1: return rhs_passes<P, 0, kVelTables>(prm, dst, tn, to, f, heav, vel, s);
          case 2: return rhs_passes<P, 0, kVelNodal>(prm, dst, tn, to, f, heav, vel, s);
          case 3: return rhs_passes<P, 1, kVelNone>(prm, dst, tn, to, f, heav, vel, s);
          case 4: return rhs_passes<P, 1, kVelTables>(prm, dst, tn, to, f, heav, vel, s);
          case 6: return rhs_passes<P, 2, kVelNone>(prm, dst, tn, to, f, heav, vel, s);
          case 7: return rhs_passes<P, 2, kVelTables>(prm, dst, tn, to, f, heav, vel, s);
          case 8: return rhs_passes<P, 2, kVelNodal>(prm, dst, tn, to, f, heav, vel, s);
          case 9: return rhs_passes<P, 3, kVelNone>(prm, dst, tn, to, f, heav, vel, s);
          case 10: return rhs_passes<P, 3, kVelTables>(prm, dst, tn, to, f, heav, vel, s);
          case 11: return rhs_passes<P, 3, kVelNodal>(prm, dst, tn, to, f, heav, vel, s);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  int l2_norm_blocks(int degree, int nx, int ny, int nz)
  {
    const int cpb = cells_per_block_for(degree, false);
    return ((nx + cpb - 1) / cpb) * ny * nz;
  }

  hipError_t launch_l2_norm_sq(int degree, const CellParams<double> &prm, const double *u,
                               double *partials, unsigned *ticket, double *out, hipStream_t s)
  {
    switch (degree)
      {
        case 1: return l2_norm_p<1>(prm, u, partials, ticket, out, s);
        case 2: return l2_norm_p<2>(prm, u, partials, ticket, out, s);
        case 3: return l2_norm_p<3>(prm, u, partials, ticket, out, s);
        case 4: return l2_norm_p<4>(prm, u, partials, ticket, out, s);
        case 5: return l2_norm_p<5>(prm, u, partials, ticket, out, s);
        case 6: return l2_norm_p<6>(prm, u, partials, ticket, out, s);
        default: return hipErrorInvalidValue;
      }
  }

  hipError_t launch_create_rhs(int degree, int coef, int vmode, const CellParams<double> &prm,
                               double *dst, const double *temperature,
                               const double *temperature_old, const double *heat_source,
                               const double *heaviside, const double *velocity,
                               hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return rhs_deg<1>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        case 2: return rhs_deg<2>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        case 3: return rhs_deg<3>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        case 4: return rhs_deg<4>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        case 5: return rhs_deg<5>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        case 6: return rhs_deg<6>(coef, vmode, prm, dst, temperature, temperature_old, heat_source, heaviside, velocity, stream);
        default: return hipErrorInvalidValue;
      }
  }
} // namespace mpdg
