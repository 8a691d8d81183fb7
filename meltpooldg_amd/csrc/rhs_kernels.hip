// Residual right-hand side of the heat operator (create_rhs), FP64.
//
// dst += -( rho_cp/dt M (T_new - T_old) - M_1 f + K_k T_new + rho_cp C_u T_new )
// on the free rows, with T_new, T_old and the heat source f read plain
// (constrained values included), as source/heat/heat_diffuse_operator.cpp
// create_rhs :783-806 / rhs_cell_loop :393-529 do for a temperature-independent
// material without evaporation: at q
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
    template <int P, int COEF, bool VEL, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      rhs_kernel(double *__restrict__ dst, const double *__restrict__ tnew,
                 const double *__restrict__ told, const double *__restrict__ fsrc,
                 const double *__restrict__ heav, const CellParams<double> prm, const int colour)
    {
      using Number = double;
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr bool H2 = COEF == kCoefHeaviside;
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      constexpr ct::EO<N> TSt = ct::VTab<P>::St;
      constexpr ct::EO<N> TC = ct::VTab<P>::C;
      constexpr ct::EO<N> TCt = ct::VTab<P>::Ct;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;

      __shared__ Number A[CPB * STR];  // T_new (nodal -> q)
      __shared__ Number Bd[CPB * STR]; // T_new - T_old; later the q-space accumulator
      __shared__ Number Cf[CPB * STR]; // heat source
      __shared__ Number KB[H2 ? CPB * STR : 1];
      __shared__ Number RB[H2 ? CPB * STR : 1];
      __shared__ Number XE[CPB * NL];

      const Segment g = segment_of<Number, CPB>(prm, colour);
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
      Number *KBl = KB + (H2 ? l * STR : 0);
      Number *RBl = RB + (H2 ? l * STR : 0);

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
              hl[a] = (H2 && cell_ok) ? heav[base + a] : Number(0);
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
          if constexpr (H2)
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
          if constexpr (H2)
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
          if constexpr (H2)
            {
              Number Hq[N];
#pragma unroll
              for (int c = 0; c < N; ++c)
                v[c] = KBl[(c * N + s1) * N + s0];
              eo_apply<N, 1>(TS, v, Hq);
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  Number rc, k;
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
                  const Number rc = H2 ? RBl[(s1 * N + s0) * N + q] : prm.rho_cp0;
                  const Number vx = -prm.sinpi2[0][cx * N + q] * prm.sin2pi[1][g.cy * N + s0];
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
                  const Number rc = H2 ? RBl[(s1 * N + q) * N + s0] : prm.rho_cp0;
                  const Number vy = prm.sin2pi[0][cx * N + s0] * prm.sinpi2[1][g.cy * N + q];
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
              const Number rc = H2 ? RBl[(q * N + s1) * N + s0] : prm.rho_cp0;
              const Number mass = H2 ? rc * prm.dtinv_vol : prm.mass;
              Number val = -(mass * w3 * dq[q] - prm.vol * w3 * fq[q]);
              if constexpr (VEL)
                val -= rc * w3 * prm.vf[2] * prm.w0 * Gz[q];
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

    template <int P, int COEF, bool VEL>
    hipError_t rhs_passes(const CellParams<double> &prm, double *dst, const double *tn,
                          const double *to, const double *f, const double *heav, hipStream_t s)
    {
      constexpr int CPB = cells_per_block_for(P, false);
      constexpr int THREADS = ((CPB * (P + 1) * (P + 1) + 63) / 64) * 64;
      const ColourGrid cg = colour_grid(prm.nx, prm.ny, prm.nz, CPB);
      for (int c = 0; c < 8; ++c)
        {
          if (cg.blocks[c] == 0)
            continue;
          rhs_kernel<P, COEF, VEL, CPB, THREADS><<<cg.blocks[c], THREADS, 0, s>>>(dst, tn, to, f,
                                                                                 heav, prm, c);
        }
      return hipGetLastError();
    }

    template <int P>
    hipError_t rhs_deg(int coef, bool vel, const CellParams<double> &prm, double *dst,
                       const double *tn, const double *to, const double *f, const double *heav,
                       hipStream_t s)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return rhs_passes<P, 0, false>(prm, dst, tn, to, f, heav, s);
          case 1: return rhs_passes<P, 0, true>(prm, dst, tn, to, f, heav, s);
          case 2: return rhs_passes<P, 1, false>(prm, dst, tn, to, f, heav, s);
          case 3: return rhs_passes<P, 1, true>(prm, dst, tn, to, f, heav, s);
          case 4: return rhs_passes<P, 2, false>(prm, dst, tn, to, f, heav, s);
          case 5: return rhs_passes<P, 2, true>(prm, dst, tn, to, f, heav, s);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  hipError_t launch_create_rhs(int degree, int coef, bool vel, const CellParams<double> &prm,
                               double *dst, const double *temperature,
                               const double *temperature_old, const double *heat_source,
                               const double *heaviside, hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return rhs_deg<1>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        case 2: return rhs_deg<2>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        case 3: return rhs_deg<3>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        case 4: return rhs_deg<4>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        case 5: return rhs_deg<5>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        case 6: return rhs_deg<6>(coef, vel, prm, dst, temperature, temperature_old, heat_source, heaviside, stream);
        default: return hipErrorInvalidValue;
      }
  }
} // namespace mpdg
