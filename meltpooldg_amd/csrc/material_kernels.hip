// Tangent and diagonal of the heat operator for temperature-dependent
// (solidification) materials, FP64: liquid_solid and gas_liquid_solid.
//
// tangent_local_cell_operation (source/heat/heat_diffuse_operator.cpp:955-1086)
// with do_solidification reads the linearisation temperature T and the old
// temperature T_old plain (:1003-1014) and evaluates the material and its
// temperature derivatives at T(q) (get_material_parameters_and_derivatives,
// :1167-1210). At q, for the trial function u:
//   val      = dt^-1 rho_cp u + rho_cp grad u . v
//              + dt^-1 d(rho_cp)/dT (T - T_old) u + d(rho_cp)/dT (grad T . v) u
//   val_grad = k grad u + dk/dT grad T u                           (:1029-1056)
//
// The coefficients depend on T, T_old and the heaviside, not on u, so each
// workgroup first evaluates them at the Gauss points of its cells into LDS:
//   CM = w3 (rho_cp |J|/dt + d(rho_cp)/dT (|J|/dt (T - T_old) + |J| v . grad T))
//   CK = w3 k,  CR = w3 rho_cp,  CB_d = w3 dk/dT dT/dx_d (reference-cell derivative)
// (w3 the Gauss weight product; the metric factors |J|/h_d^2 and |J|/h_d are
// applied with the derivative of u). Everything is recomputed per launch from
// the three nodal fields, so nothing per quadrature point is stored in HBM.
//
// The same kernels run the constant-density materials (gas, gas_liquid: the
// material of the heaviside alone, no temperature terms) when the velocity is
// a nodal DoF-vector field (VM == kVelNodal): its three components are read
// plain at the cell nodes and interpolated to the Gauss points
// (vel_eval.read_dof_values_plain + evaluate(values),
// heat_diffuse_operator.cpp:975-980), and w3 rho_cp v_d is kept per point.
//
// Block-synchronous layout, as the residual kernel: one workgroup per row
// segment of CPB cells of one colour (thread = cell x line), 1D contractions
// in registers with the compile-time even-odd tables, LDS between the line
// directions, x-faces shared inside the segment summed through LDS and every
// other shared node by the coloured first-touch scatter, so the result does
// not depend on scheduling. CPB is smaller than the residual kernel's because
// eight cell tensors per cell are live in LDS.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    // cells per workgroup: eight cell tensors live in LDS, eleven with a nodal
    // velocity (then half as many cells from degree 5 on)
    constexpr int tmat_cpb(int P, int VM = 0)
    {
      return (P <= 2 ? 32 : (P == 3 ? 16 : (P <= 5 ? 8 : 4))) / (VM == kVelNodal && P >= 5 ? 2 : 1);
    }

    template <int P, int VM>
    constexpr int tmat_threads()
    {
      return ((tmat_cpb(P, VM) * (P + 1) * (P + 1) + 63) / 64) * 64;
    }

    // Coefficients of the linearised material at the Gauss points of the
    // segment's cells (layout of each cell tensor: (qz * N + qy) * N + qx).
    // Every thread of the workgroup must call it; ends with a barrier.
    template <int P, int VM, int CPB>
    __device__ __forceinline__ void tmat_coefficients(const CellParams<double> &prm,
                                                      const Segment &g, const double *T,
                                                      const double *To, const double *heav,
                                                      const double *vel, double *CM, double *CK,
                                                      double *CR, double *CB0, double *CB1,
                                                      double *CB2, double *V0, double *V1,
                                                      double *V2)
    {
      if constexpr (VM == kVelNodal)
        nodal_vector_at_q<P, CPB>(prm, g, vel, V0, V1, V2);
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      constexpr ct::EO<N> TC = ct::VTab<P>::C;
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;
      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int cx = min(g.cx0 + l, prm.nx - 1);
      const bool use_h = heav != nullptr;        // gas_liquid, gas_liquid_solid
      const bool use_t = prm.tmat >= 2;           // temperature-dependent material
      // staging: T -> CK, T_old -> CM, H -> CR, grad T -> CB0..2
      double *Tl = CK + l * STR, *Ol = CM + l * STR, *Hl = CR + l * STR;
      double *G0 = CB0 + l * STR, *G1 = CB1 + l * STR, *G2 = CB2 + l * STR;

      // x-lines (y = s0, z = s1): plain gathers, interpolate in x
      if (active)
        {
          const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
          const int64_t base = (int64_t)(g.cx0 + l) * P + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);
          double tl[N], ol[N], hl[N], o[N];
#pragma unroll
          for (int a = 0; a < N; ++a)
            {
              tl[a] = (use_t && cell_ok) ? T[base + a] : 0.0;
              ol[a] = (use_t && cell_ok) ? To[base + a] : 0.0;
              hl[a] = (use_h && cell_ok) ? heav[base + a] : 0.0;
            }
          const int row = (s1 * N + s0) * N;
          eo_apply<N, 1>(TS, tl, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Tl[row + q] = o[q];
          eo_apply<N, 1>(TS, ol, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Ol[row + q] = o[q];
          if (use_h)
            {
              eo_apply<N, 1>(TS, hl, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                Hl[row + q] = o[q];
            }
        }
      __syncthreads();

      // y-lines (qx = s0, z = s1): interpolate in y
      if (active)
        {
          auto sweep = [&](double *X) {
            double v[N], o[N];
#pragma unroll
            for (int b = 0; b < N; ++b)
              v[b] = X[(s1 * N + b) * N + s0];
            eo_apply<N, 1>(TS, v, o);
#pragma unroll
            for (int b = 0; b < N; ++b)
              X[(s1 * N + b) * N + s0] = o[b];
          };
          sweep(Tl);
          sweep(Ol);
          if (use_h)
            sweep(Hl);
        }
      __syncthreads();

      // z-lines (qx = s0, qy = s1): values at q, z-derivative of T
      if (active)
        {
          double v[N], o[N], gz[N];
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Tl[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, o);
          eo_apply<N, -1>(TC, o, gz);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              Tl[(q * N + s1) * N + s0] = o[q];
              G2[(q * N + s1) * N + s0] = gz[q];
            }
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Ol[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Ol[(q * N + s1) * N + s0] = o[q];
          if (use_h)
            {
#pragma unroll
              for (int c = 0; c < N; ++c)
                v[c] = Hl[(c * N + s1) * N + s0];
              eo_apply<N, 1>(TS, v, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                Hl[(q * N + s1) * N + s0] = o[q];
            }
        }
      __syncthreads();

      // x-derivative (row qy = s0, qz = s1) and y-derivative (column qx = s0,
      // qz = s1) of T at q
      if (active)
        {
          double v[N], d[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Tl[(s1 * N + s0) * N + q];
          eo_apply<N, -1>(TC, v, d);
#pragma unroll
          for (int q = 0; q < N; ++q)
            G0[(s1 * N + s0) * N + q] = d[q];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Tl[(s1 * N + q) * N + s0];
          eo_apply<N, -1>(TC, v, d);
#pragma unroll
          for (int q = 0; q < N; ++q)
            G1[(s1 * N + q) * N + s0] = d[q];
        }
      __syncthreads();

      // material at q (column qx = s0, qy = s1), in place
      if (active)
        {
          const double wxy = T1.w[s0] * T1.w[s1];
          double ux = 0.0, uy = 0.0;
          if constexpr (VM == kVelTables)
            {
              ux = prm.vel_x[0][cx * N + s0] * prm.vel_y[0][g.cy * N + s1];
              uy = prm.vel_x[1][cx * N + s0] * prm.vel_y[1][g.cy * N + s1];
            }
          double *W0 = V0 + l * STR, *W1 = V1 + l * STR, *W2 = V2 + l * STR;
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (q * N + s1) * N + s0;
              const double Tq = Tl[idx], Oq = Ol[idx], Hq = use_h ? Hl[idx] : 0.0;
              const double g0 = G0[idx], g1 = G1[idx], g2 = G2[idx];
              double rc, k, drc = 0.0, dk = 0.0;
              if (use_t)
                solid_liquid_material(prm, Hq, Tq, rc, k, drc, dk);
              else
                two_phase(prm, Hq, rc, k);
              const double w3 = T1.w[q] * wxy;
              double m = rc * prm.dtinv_vol + drc * prm.dtinv_vol * (Tq - Oq);
              if constexpr (VM == kVelTables)
                m += drc * (prm.vf[0] * ux * g0 + prm.vf[1] * uy * g1 + prm.vf[2] * prm.w0 * g2);
              if constexpr (VM == kVelNodal)
                {
                  const double vx = W0[idx], vy = W1[idx], vz = W2[idx];
                  m += drc * (prm.vf[0] * vx * g0 + prm.vf[1] * vy * g1 + prm.vf[2] * vz * g2);
                  // the advection coefficients w3 rho_cp v_d of the cell operator
                  W0[idx] = w3 * rc * vx;
                  W1[idx] = w3 * rc * vy;
                  W2[idx] = w3 * rc * vz;
                }
              Ol[idx] = w3 * m;
              Tl[idx] = w3 * k;
              Hl[idx] = w3 * rc;
              G0[idx] = w3 * dk * g0;
              G1[idx] = w3 * dk * g1;
              G2[idx] = w3 * dk * g2;
            }
        }
      __syncthreads();
    }

    // dst = J src, one colour pass; partial p.Ap of the pass when partials is
    // given (then scal receives p.Ap); a given scal alone only stops the launch
    // once scal->done is set.
    template <int P, int VM, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      tmat_vmult_kernel(const double *__restrict__ src, double *__restrict__ dst,
                        const double *__restrict__ T, const double *__restrict__ To,
                        const double *__restrict__ heav, const double *__restrict__ vel,
                        const CellParams<double> prm, const int colour,
                        double *__restrict__ partials, const int offset, const int total,
                        const int finalize, SolverScalars *scal)
    {
      constexpr bool VEL = VM != kVelNone;
      if (scal != nullptr && scal->done)
        return; // uniform over the grid: done is not written during the launch
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      constexpr ct::EO<N> TSt = ct::VTab<P>::St;
      constexpr ct::EO<N> TC = ct::VTab<P>::C;
      constexpr ct::EO<N> TCt = ct::VTab<P>::Ct;

      __shared__ double A[CPB * STR];  // u: nodal -> q; after the z-integration the output
      __shared__ double B[CPB * STR];  // q-space accumulator of the x and y passes
      __shared__ double CM[CPB * STR], CK[CPB * STR], CR[CPB * STR];
      __shared__ double CB0[CPB * STR], CB1[CPB * STR], CB2[CPB * STR];
      __shared__ double V0[VM == kVelNodal ? CPB * STR : 1], V1[VM == kVelNodal ? CPB * STR : 1],
        V2[VM == kVelNodal ? CPB * STR : 1];
      __shared__ double XE[CPB * NL];

      const Segment g = segment_of<double, CPB>(prm, colour);
      tmat_coefficients<P, VM, CPB>(prm, g, T, To, heav, vel, CM, CK, CR, CB0, CB1, CB2, V0, V1,
                                    V2);

      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int cx = min(g.cx0 + l, prm.nx - 1);
      const int vo = VM == kVelNodal ? l * STR : 0;
      const double *W0 = V0 + vo, *W1 = V1 + vo, *W2 = V2 + vo; // w3 rho_cp v_d at q
      double *Al = A + l * STR;
      double *Bl = B + l * STR;
      const double *Ml = CM + l * STR, *Kl = CK + l * STR, *Rl = CR + l * STR;
      const double *E0 = CB0 + l * STR, *E1 = CB1 + l * STR, *E2 = CB2 + l * STR;
      const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
      const int X0 = (g.cx0 + l) * P;
      const bool rowc = on_dirichlet(prm, Yi, 1) || on_dirichlet(prm, Zi, 2);
      const int64_t base = X0 + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);

      // x-lines (y = s0, z = s1): read_dof_values (constrained -> 0), interpolate x
      double ul[N];
      if (active)
        {
          double o[N];
#pragma unroll
          for (int a = 0; a < N; ++a)
            ul[a] = (cell_ok && !rowc && !on_dirichlet(prm, X0 + a, 0)) ? src[base + a] : 0.0;
          eo_apply<N, 1>(TS, ul, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Al[(s1 * N + s0) * N + q] = o[q];
        }
      __syncthreads();

      // y-lines (qx = s0, z = s1): interpolate y
      if (active)
        {
          double v[N], o[N];
#pragma unroll
          for (int b = 0; b < N; ++b)
            v[b] = Al[(s1 * N + b) * N + s0];
          eo_apply<N, 1>(TS, v, o);
#pragma unroll
          for (int b = 0; b < N; ++b)
            Al[(s1 * N + b) * N + s0] = o[b];
        }
      __syncthreads();

      // z-lines (qx = s0, qy = s1): values and z-derivative of u at q
      double U[N], Gz[N];
      if (active)
        {
          double v[N];
#pragma unroll
          for (int c = 0; c < N; ++c)
            v[c] = Al[(c * N + s1) * N + s0];
          eo_apply<N, 1>(TS, v, U);
          eo_apply<N, -1>(TC, U, Gz);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Al[(q * N + s1) * N + s0] = U[q];
        }
      __syncthreads();

      // x-rows (qy = s0, qz = s1): x-flux k du/dx + dk/dT dT/dx u, x-convection
      if (active)
        {
          double v[N], gx[N], f[N], o[N], cv[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + s0) * N + q];
          eo_apply<N, -1>(TC, v, gx);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (s1 * N + s0) * N + q;
              f[q] = prm.kf[0] * (Kl[idx] * gx[q] + E0[idx] * v[q]);
              cv[q] = 0.0;
              if constexpr (VM == kVelTables)
                {
                  const double vx = prm.vel_x[0][cx * N + q] * prm.vel_y[0][g.cy * N + s0];
                  cv[q] = prm.vf[0] * Rl[idx] * vx * gx[q];
                }
              if constexpr (VM == kVelNodal)
                cv[q] = prm.vf[0] * W0[idx] * gx[q];
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            Bl[(s1 * N + s0) * N + q] = o[q] + cv[q];
        }
      __syncthreads();

      // y-columns (qx = s0, qz = s1): y-flux, y-convection
      if (active)
        {
          double v[N], gy[N], f[N], o[N], cv[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + q) * N + s0];
          eo_apply<N, -1>(TC, v, gy);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (s1 * N + q) * N + s0;
              f[q] = prm.kf[1] * (Kl[idx] * gy[q] + E1[idx] * v[q]);
              cv[q] = 0.0;
              if constexpr (VM == kVelTables)
                {
                  const double vy = prm.vel_x[1][cx * N + s0] * prm.vel_y[1][g.cy * N + q];
                  cv[q] = prm.vf[1] * Rl[idx] * vy * gy[q];
                }
              if constexpr (VM == kVelNodal)
                cv[q] = prm.vf[1] * W1[idx] * gy[q];
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (s1 * N + q) * N + s0;
              Bl[idx] = Bl[idx] + o[q] + cv[q];
            }
        }
      __syncthreads();

      // z-columns (qx = s0, qy = s1): z-flux, value terms, integrate z
      if (active)
        {
          double f[N], o[N], V[N], y1[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (q * N + s1) * N + s0;
              f[q] = prm.kf[2] * (Kl[idx] * Gz[q] + E2[idx] * U[q]);
            }
          eo_apply<N, -1>(TCt, f, o);
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (q * N + s1) * N + s0;
              double val = Ml[idx] * U[q];
              if constexpr (VM == kVelTables)
                val += prm.vf[2] * Rl[idx] * prm.w0 * Gz[q];
              if constexpr (VM == kVelNodal)
                val += prm.vf[2] * W2[idx] * Gz[q];
              V[q] = val + o[q] + Bl[idx];
            }
          eo_apply<N, 1>(TSt, V, y1);
#pragma unroll
          for (int c = 0; c < N; ++c)
            Al[(c * N + s1) * N + s0] = y1[c];
        }
      __syncthreads();

      // y-lines (qx = s0, c = s1): integrate y
      if (active)
        {
          double v[N], o[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + q) * N + s0];
          eo_apply<N, 1>(TSt, v, o);
#pragma unroll
          for (int b = 0; b < N; ++b)
            Al[(s1 * N + b) * N + s0] = o[b];
        }
      __syncthreads();

      // x-lines (b = s0, c = s1): integrate x, in-segment x-face exchange, scatter
      double Y[N];
      if (active)
        {
          double v[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            v[q] = Al[(s1 * N + s0) * N + q];
          eo_apply<N, 1>(TSt, v, Y);
          XE[s * CPB + l] = Y[N - 1];
        }
      __syncthreads();
      double acc = 0.0;
      if (cell_ok)
        {
          if (l > 0)
            Y[0] += XE[s * CPB + l - 1];
          scatter_xline<P, double, CPB>(prm, g, l, s0, s1, Y, dst);
          // p.Ap: src . (this cell row's added values) over the nodes it writes
          const bool last = l == g.ncell - 1;
#pragma unroll
          for (int a = 0; a < N; ++a)
            if (a < N - 1 || last)
              acc += ul[a] * Y[a];
        }
      if (partials != nullptr)
        pass_partials<THREADS>(acc, partials, offset, total, finalize, scal);
    }

    // Diagonal of the tangent, one colour pass:
    // diag_i = sum_q [ CM phi_i^2 + sum_d kf_d CK (d_d phi_i)^2
    //                  + sum_d (vf_d CR v_d + kf_d CB_d) phi_i d_d phi_i ]
    // contracted direction by direction with S o S, D o D and S o D.
    template <int P, int VM, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      tmat_diag_kernel(double *__restrict__ dst, const double *__restrict__ T,
                       const double *__restrict__ To, const double *__restrict__ heav,
                       const double *__restrict__ vel, const DiagTables<double> tab,
                       const CellParams<double> prm, const int colour)
    {
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      __shared__ double CM[CPB * STR], CK[CPB * STR], CR[CPB * STR];
      __shared__ double CB0[CPB * STR], CB1[CPB * STR], CB2[CPB * STR];
      __shared__ double V0[VM == kVelNodal ? CPB * STR : 1], V1[VM == kVelNodal ? CPB * STR : 1],
        V2[VM == kVelNodal ? CPB * STR : 1];
      __shared__ double XE[CPB * NL];

      const Segment g = segment_of<double, CPB>(prm, colour);
      tmat_coefficients<P, VM, CPB>(prm, g, T, To, heav, vel, CM, CK, CR, CB0, CB1, CB2, V0, V1,
                                    V2);
      const int vo0 = VM == kVelNodal ? (int)(threadIdx.x % CPB) * STR : 0;
      const double *W0 = V0 + vo0, *W1 = V1 + vo0, *W2 = V2 + vo0; // w3 rho_cp v_d at q

      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int cx = min(g.cx0 + l, prm.nx - 1);
      // the contraction fields overwrite the coefficient tensors in place: each
      // thread reads its row's coefficients before it writes the row
      double *Am = CM + l * STR, *Bk = CK + l * STR, *Cr = CR + l * STR;
      double *D0 = CB0 + l * STR, *D1 = CB1 + l * STR, *D2 = CB2 + l * STR;

      // x-rows (qy = s0, qz = s1): fields of the three directions, contract x
      if (active)
        {
          double Fm[N], Fky[N], Fkz[N], Fvy[N], Fvz[N], Fx2[N], Fxs[N];
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const int idx = (s1 * N + s0) * N + q;
              const double k = Bk[idx];
              double vx = 0.0, vy = 0.0, vz = 0.0;
              if constexpr (VM == kVelTables)
                {
                  vx = prm.vel_x[0][cx * N + q] * prm.vel_y[0][g.cy * N + s0];
                  vy = prm.vel_x[1][cx * N + q] * prm.vel_y[1][g.cy * N + s0];
                  vz = prm.w0;
                }
              // w3 rho_cp v_d: the tables' velocity times Cr, or the nodal one
              const double a0 = VM == kVelNodal ? W0[idx] : Cr[idx] * vx;
              const double a1 = VM == kVelNodal ? W1[idx] : Cr[idx] * vy;
              const double a2 = VM == kVelNodal ? W2[idx] : Cr[idx] * vz;
              Fm[q] = Am[idx];
              Fx2[q] = prm.kf[0] * k;
              Fxs[q] = prm.vf[0] * a0 + prm.kf[0] * D0[idx];
              Fky[q] = prm.kf[1] * k;
              Fkz[q] = prm.kf[2] * k;
              Fvy[q] = prm.vf[1] * a1 + prm.kf[1] * D1[idx];
              Fvz[q] = prm.vf[2] * a2 + prm.kf[2] * D2[idx];
            }
          double g1[N], t1[N], g2[N], g3[N], g4[N], g5[N];
          eo_apply<N, 1>(tab.S2t, Fm, g1);
          eo_apply<N, 1>(tab.D2t, Fx2, t1);
#pragma unroll
          for (int a = 0; a < N; ++a)
            g1[a] += t1[a];
          eo_apply<N, -1>(tab.SDt, Fxs, t1);
#pragma unroll
          for (int a = 0; a < N; ++a)
            g1[a] += t1[a];
          eo_apply<N, 1>(tab.S2t, Fky, g2);
          eo_apply<N, 1>(tab.S2t, Fvy, g3);
          eo_apply<N, 1>(tab.S2t, Fkz, g4);
          eo_apply<N, 1>(tab.S2t, Fvz, g5);
#pragma unroll
          for (int a = 0; a < N; ++a)
            {
              const int idx = (s1 * N + s0) * N + a;
              Am[idx] = g1[a];
              Bk[idx] = g2[a];
              Cr[idx] = g3[a];
              D0[idx] = g4[a];
              D1[idx] = g5[a];
            }
        }
      __syncthreads();

      // y-columns (a = s0, qz = s1): contract y
      if (active)
        {
          double v1[N], v2[N], v3[N], v4[N], v5[N], h1[N], t1[N], h2[N], h3[N];
#pragma unroll
          for (int r = 0; r < N; ++r)
            {
              const int idx = (s1 * N + r) * N + s0;
              v1[r] = Am[idx];
              v2[r] = Bk[idx];
              v3[r] = Cr[idx];
              v4[r] = D0[idx];
              v5[r] = D1[idx];
            }
          eo_apply<N, 1>(tab.S2t, v1, h1);
          eo_apply<N, 1>(tab.D2t, v2, t1);
#pragma unroll
          for (int b = 0; b < N; ++b)
            h1[b] += t1[b];
          eo_apply<N, -1>(tab.SDt, v3, t1);
#pragma unroll
          for (int b = 0; b < N; ++b)
            h1[b] += t1[b];
          eo_apply<N, 1>(tab.S2t, v4, h2);
          eo_apply<N, 1>(tab.S2t, v5, h3);
#pragma unroll
          for (int b = 0; b < N; ++b)
            {
              const int idx = (s1 * N + b) * N + s0;
              Am[idx] = h1[b];
              Bk[idx] = h2[b];
              Cr[idx] = h3[b];
            }
        }
      __syncthreads();

      // z-columns (a = s0, b = s1): contract z
      if (active)
        {
          double v1[N], v2[N], v3[N], d[N], t1[N];
#pragma unroll
          for (int r = 0; r < N; ++r)
            {
              const int idx = (r * N + s1) * N + s0;
              v1[r] = Am[idx];
              v2[r] = Bk[idx];
              v3[r] = Cr[idx];
            }
          eo_apply<N, 1>(tab.S2t, v1, d);
          eo_apply<N, 1>(tab.D2t, v2, t1);
#pragma unroll
          for (int c = 0; c < N; ++c)
            d[c] += t1[c];
          eo_apply<N, -1>(tab.SDt, v3, t1);
#pragma unroll
          for (int c = 0; c < N; ++c)
            d[c] += t1[c];
#pragma unroll
          for (int c = 0; c < N; ++c)
            Am[(c * N + s1) * N + s0] = d[c];
        }
      __syncthreads();

      // x-lines (b = s0, c = s1): exchange the shared x-face, scatter
      double Y[N];
      if (active)
        {
#pragma unroll
          for (int a = 0; a < N; ++a)
            Y[a] = Am[(s1 * N + s0) * N + a];
          XE[s * CPB + l] = Y[N - 1];
        }
      __syncthreads();
      if (cell_ok)
        {
          if (l > 0)
            Y[0] += XE[s * CPB + l - 1];
          scatter_xline<P, double, CPB>(prm, g, l, s0, s1, Y, dst);
        }
    }

    template <int P, int VM>
    hipError_t tmat_vmult_passes(const CellParams<double> &prm, const double *src, double *dst,
                                 const double *T, const double *To, const double *heav,
                                 const double *vel, double *partials, SolverScalars *scal,
                                 hipStream_t s)
    {
      constexpr int CPB = tmat_cpb(P, VM);
      constexpr int THREADS = tmat_threads<P, VM>();
      const ColourGrid cg = colour_grid(prm.nx, prm.ny, prm.nz, CPB);
      for (int c = 0; c < 8; ++c)
        {
          if (cg.blocks[c] == 0)
            continue;
          tmat_vmult_kernel<P, VM, CPB, THREADS><<<cg.blocks[c], THREADS, 0, s>>>(
            src, dst, T, To, heav, vel, prm, c, partials, cg.offset[c], cg.total,
            c == cg.last ? 1 : 0, scal);
        }
      return hipGetLastError();
    }

    template <int P, int VM>
    hipError_t tmat_diag_passes(const CellParams<double> &prm, const DiagTables<double> &tab,
                                double *dst, const double *T, const double *To,
                                const double *heav, const double *vel, hipStream_t s)
    {
      constexpr int CPB = tmat_cpb(P, VM);
      constexpr int THREADS = tmat_threads<P, VM>();
      const ColourGrid cg = colour_grid(prm.nx, prm.ny, prm.nz, CPB);
      for (int c = 0; c < 8; ++c)
        {
          if (cg.blocks[c] == 0)
            continue;
          tmat_diag_kernel<P, VM, CPB, THREADS><<<cg.blocks[c], THREADS, 0, s>>>(
            dst, T, To, heav, vel, tab, prm, c);
        }
      return hipGetLastError();
    }

    template <int P>
    hipError_t tmat_vmult_vm(int vmode, const CellParams<double> &prm, const double *src,
                             double *dst, const double *T, const double *To, const double *heav,
                             const double *vel, double *partials, SolverScalars *scal,
                             hipStream_t s)
    {
      switch (vmode)
        {
          case kVelNone: return tmat_vmult_passes<P, kVelNone>(prm, src, dst, T, To, heav, vel, partials, scal, s);
          case kVelTables: return tmat_vmult_passes<P, kVelTables>(prm, src, dst, T, To, heav, vel, partials, scal, s);
          case kVelNodal: return tmat_vmult_passes<P, kVelNodal>(prm, src, dst, T, To, heav, vel, partials, scal, s);
          default: return hipErrorInvalidValue;
        }
    }

    template <int P>
    hipError_t tmat_diag_vm(int vmode, const CellParams<double> &prm,
                            const DiagTables<double> &tab, double *dst, const double *T,
                            const double *To, const double *heav, const double *vel, hipStream_t s)
    {
      switch (vmode)
        {
          case kVelNone: return tmat_diag_passes<P, kVelNone>(prm, tab, dst, T, To, heav, vel, s);
          case kVelTables: return tmat_diag_passes<P, kVelTables>(prm, tab, dst, T, To, heav, vel, s);
          case kVelNodal: return tmat_diag_passes<P, kVelNodal>(prm, tab, dst, T, To, heav, vel, s);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  int tmat_blocks(int degree, int nx, int ny, int nz)
  {
    // the nodal-velocity kernels' smaller segments give the larger count
    return colour_grid(nx, ny, nz, tmat_cpb(degree, kVelNodal)).total;
  }

  hipError_t launch_tmat_vmult(int degree, int vmode, const CellParams<double> &prm,
                               const double *src, double *dst, const double *temperature,
                               const double *temperature_old, const double *heaviside,
                               const double *velocity, double *partials, SolverScalars *scal,
                               hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return tmat_vmult_vm<1>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        case 2: return tmat_vmult_vm<2>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        case 3: return tmat_vmult_vm<3>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        case 4: return tmat_vmult_vm<4>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        case 5: return tmat_vmult_vm<5>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        case 6: return tmat_vmult_vm<6>(vmode, prm, src, dst, temperature, temperature_old, heaviside, velocity, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  hipError_t launch_tmat_diagonal(int degree, int vmode, const CellParams<double> &prm,
                                  const DiagTables<double> &tab, double *dst,
                                  const double *temperature, const double *temperature_old,
                                  const double *heaviside, const double *velocity,
                                  hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return tmat_diag_vm<1>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        case 2: return tmat_diag_vm<2>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        case 3: return tmat_diag_vm<3>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        case 4: return tmat_diag_vm<4>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        case 5: return tmat_diag_vm<5>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        case 6: return tmat_diag_vm<6>(vmode, prm, tab, dst, temperature, temperature_old, heaviside, velocity, stream);
        default: return hipErrorInvalidValue;
      }
  }
} // namespace mpdg
