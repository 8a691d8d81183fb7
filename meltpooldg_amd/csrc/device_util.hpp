// Device helpers shared by the cell kernels: even-odd 1D contractions with
// compile-time tables, raw buffer loads/stores with out-of-range masking, the
// two-phase material blend, and deterministic block reductions.
#pragma once

#include "ctables.hpp"
#include "kernels.hpp"
#include "reduce.hpp"

namespace mpdg
{
  // Non-returning atomic add of a double through a buffer resource (the LLVM
  // intrinsic behind buffer_atomic_add_f64; out-of-range offsets are dropped).
  __device__ double buffer_atomic_fadd_f64(double v, __amdgpu_buffer_rsrc_t r, int off, int soff,
                                           int aux) __asm("llvm.amdgcn.raw.ptr.buffer.atomic.fadd.f64");

  namespace
  {
    // acc + c x, where a compile-time zero c adds nothing (GLL nodes and Gauss
    // points share the cell centre at odd point counts, so the interpolation
    // matrix's middle row and its transpose's middle column are zero off the
    // diagonal). `set` says whether acc holds a term yet; after unrolling, c
    // and set are constants and the tests fold away.
    template <typename Number>
    __device__ __forceinline__ void eo_term(Number &acc, bool &set, double c, Number x)
    {
      if (c == 0.0)
        return;
      if (c == 1.0)
        acc = set ? acc + x : x;
      else
        acc = set ? acc + Number(c) * x : Number(c) * x;
      set = true;
    }

    // y = M x with a compile-time even-odd table (entries folded into the code).
    template <int N, int SIGN, typename Number>
    __device__ __forceinline__ void eo_apply(const ct::EO<N> &M, const Number (&x)[N], Number (&y)[N])
    {
      constexpr int H = N / 2;
      Number E[H], O[H];
#pragma unroll
      for (int j = 0; j < H; ++j)
        {
          E[j] = x[j] + x[N - 1 - j];
          O[j] = x[j] - x[N - 1 - j];
        }
#pragma unroll
      for (int i = 0; i < H; ++i)
        {
          Number a = Number(0), b = Number(0);
          bool sa = false, sb = false;
#pragma unroll
          for (int j = 0; j < H; ++j)
            {
              eo_term(a, sa, M.e[i][j], E[j]);
              eo_term(b, sb, M.o[i][j], O[j]);
            }
          if constexpr (N % 2 == 1)
            eo_term(a, sa, M.mc[i], x[H]);
          y[i] = a + b;
          y[N - 1 - i] = SIGN > 0 ? a - b : b - a;
        }
      if constexpr (N % 2 == 1)
        {
          Number m = Number(0);
          bool sm = false;
#pragma unroll
          for (int j = 0; j < H; ++j)
            eo_term(m, sm, M.mr[j], SIGN > 0 ? E[j] : O[j]);
          eo_term(m, sm, M.mm, x[H]);
          y[H] = m;
        }
    }

    // y = M x with M in even-odd form held at run time (kernel argument), SIGN
    // the symmetry sign of M.
    template <int N, int SIGN, typename Number>
    __device__ __forceinline__ void eo_apply(const EvenOdd<Number> &M, const Number (&x)[N],
                                             Number (&y)[N])
    {
      constexpr int H = N / 2;
      Number E[H], O[H];
#pragma unroll
      for (int j = 0; j < H; ++j)
        {
          E[j] = x[j] + x[N - 1 - j];
          O[j] = x[j] - x[N - 1 - j];
        }
#pragma unroll
      for (int i = 0; i < H; ++i)
        {
          Number a = M.e[i][0] * E[0];
          Number b = M.o[i][0] * O[0];
#pragma unroll
          for (int j = 1; j < H; ++j)
            {
              a += M.e[i][j] * E[j];
              b += M.o[i][j] * O[j];
            }
          if constexpr (N % 2 == 1)
            a += M.mc[i] * x[H];
          y[i] = a + b;
          y[N - 1 - i] = SIGN > 0 ? a - b : b - a;
        }
      if constexpr (N % 2 == 1)
        {
          Number m = M.mm * x[H];
#pragma unroll
          for (int j = 0; j < H; ++j)
            m += M.mr[j] * (SIGN > 0 ? E[j] : O[j]);
          y[H] = m;
        }
    }

    // Raw buffer access: an offset at or past the buffer's byte size reads 0
    // and drops the store, so masked lanes need no branch.
    constexpr uint32_t kOutOfRange = 0xFFFFFF00u;

    __device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes)
    {
      return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                               0x00020000);
    }

    typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

    __device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, double)
    {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    }
    __device__ __forceinline__ float buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off, float)
    {
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
    __device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, double v)
    {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 0);
    }
    __device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
    {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
    }

    // Agent-coherent forms (cache policy sc1): a store that reaches the point
    // of coherence of all XCDs before it completes, a load served from there.
    __device__ __forceinline__ double buf_load_coherent(__amdgpu_buffer_rsrc_t r, uint32_t off, double)
    {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 16));
    }
    __device__ __forceinline__ float buf_load_coherent(__amdgpu_buffer_rsrc_t r, uint32_t off, float)
    {
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 16));
    }
    __device__ __forceinline__ void buf_store_coherent(__amdgpu_buffer_rsrc_t r, uint32_t off, double v)
    {
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, off, 0, 16);
    }
    __device__ __forceinline__ void buf_store_coherent(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
    {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 16);
    }

    // dst[off] += v as one device-scope atomic (no return value).
    __device__ __forceinline__ void buf_add(__amdgpu_buffer_rsrc_t r, uint32_t off, double v)
    {
      (void)buffer_atomic_fadd_f64(v, r, (int)off, 0, 0);
    }
    __device__ __forceinline__ void buf_add(__amdgpu_buffer_rsrc_t r, uint32_t off, float v)
    {
      (void)__builtin_amdgcn_raw_ptr_buffer_atomic_fadd_f32(v, r, (int)off, 0, 0);
    }

    // on_dirichlet without short-circuit branches
    template <typename Number>
    __device__ __forceinline__ bool dirichlet_node(const CellParams<Number> &prm, int i, int n1, int d)
    {
      const unsigned lo = (prm.faces >> (2 * d)) & 1u;
      const unsigned hi = (prm.faces >> (2 * d + 1)) & 1u;
      return (lo & (unsigned)(i == 0)) | (hi & (unsigned)(i == n1 - 1));
    }

    // Material::compute_parameters for gas_liquid with the blend of
    // LevelSet::Tools::interpolate (material.templates.hpp:515-524,
    // level_set_tools.hpp:40-43); rho*cp as in heat_diffuse_operator.cpp:1167-1210.
    template <typename Number>
    __device__ __forceinline__ void two_phase(const CellParams<Number> &prm, Number H, Number &rc,
                                              Number &k)
    {
      const Number w = prm.sharp ? (H > Number(0.5) ? Number(1) : Number(0)) : H;
      const Number one_w = Number(1) - w;
      k = one_w * prm.g_k + w * prm.l_k;
      if (prm.volume_specific)
        rc = one_w * prm.g_rhocp + w * prm.l_rhocp;
      else
        rc = (one_w * prm.g_rho + w * prm.l_rho) * (one_w * prm.g_cp + w * prm.l_cp);
    }

    // Temperature-dependent Material::compute_parameters_internal for
    // liquid_solid (prm.tmat == 2) and gas_liquid_solid (3) with the
    // temperature derivatives (material.templates.hpp:296-495, 540-756): solid
    // fraction fs = clamp((liquidus - T) / (liquidus - solidus), 0, 1), or
    // T < solidus for the sharp law; each property blends liquid and solid by
    // the cubic smoothstep of fs (exact phase values and zero derivative for a
    // pure phase), c_p gains the apparent capacity of a latent heat of fusion
    // (material.templates.hpp:563-617), then gas_liquid_solid blends the gas
    // value with weight 1 - w
    // (w = H, or H > 0.5 when sharp) and scales the derivative by w; rho_cp and
    // its derivative as get_material_parameters_and_derivatives
    // (heat_diffuse_operator.cpp:1167-1210).
    // C1 ramp from (x1, y1) to (x2, y2), zero slope at both ends over a width
    // delta, linear in between (slope (y2 - y1) / (x2 - x1 - delta)).
    template <typename Number>
    __device__ __forceinline__ void qlq_ramp(Number x, Number x1, Number x2, Number y1, Number y2,
                                             Number delta, Number &v, Number &dv)
    {
      const Number s = (y2 - y1) / (x2 - x1 - delta);
      if (x < x1 + delta)
        {
          const Number d = x - x1;
          v = y1 + Number(0.5) * s * d * d / delta;
          dv = s * d / delta;
        }
      else if (x > x2 - delta)
        {
          const Number d = x2 - x;
          v = y2 - Number(0.5) * s * d * d / delta;
          dv = s * d / delta;
        }
      else
        {
          v = y1 + s * (x - x1 - Number(0.5) * delta);
          dv = s;
        }
    }

    // Heat::ApparentCapacity (source/heat/apparent_capacity.cpp:157-311): the
    // latent heat L as an extra c_p inside the mushy zone (0 < fs < 1) and its
    // temperature derivative: poly4_bell L (2/dT) 15/16 (1 - xi^2)^2 with
    // xi = 1 - 2 fs; constant L / dT; qlq C1 ramps 0 -> 2 L / dT -> 0 over
    // (T_s, T_m, T_l) with transition width dT / 6.
    template <typename Number>
    __device__ __forceinline__ void apparent_capacity(const CellParams<Number> &prm, Number fs,
                                                      Number &c, Number &dc)
    {
      c = dc = Number(0);
      if (!(prm.latent > Number(0)) || fs <= Number(0) || fs >= Number(1))
        return;
      const Number Ts = prm.solidus, Tl = prm.liquidus, dT = Tl - Ts, L = prm.latent;
      if (prm.ac_type == 1)
        c = L / dT;
      else if (prm.ac_type == 0)
        {
          const Number r = Number(2) / dT, xi = Number(1) - Number(2) * fs, q = Number(1) - xi * xi;
          c = L * r * Number(15.0 / 16.0) * q * q;
          dc = L * r * r * Number(15.0 / 4.0) * xi * (xi * xi - Number(1));
        }
      else
        {
          const Number T = Tl - fs * dT, Tm = Number(0.5) * (Ts + Tl), peak = Number(2) * L / dT;
          if (T <= Tm)
            qlq_ramp(T, Ts, Tm, Number(0), peak, dT / Number(6), c, dc);
          else
            qlq_ramp(T, Tm, Tl, peak, Number(0), dT / Number(6), c, dc);
        }
    }

    template <typename Number>
    __device__ __forceinline__ void solid_liquid_material(const CellParams<Number> &prm, Number H,
                                                          Number T, Number &rc, Number &k,
                                                          Number &drc, Number &dk)
    {
      Number fs;
      if (prm.sl_sharp)
        fs = T < prm.solidus ? Number(1) : Number(0);
      else
        {
          fs = (prm.liquidus - T) * prm.inv_mushy;
          fs = fs > Number(1) ? Number(1) : (fs < Number(0) ? Number(0) : fs);
        }
      const bool solid = fs == Number(1);
      const bool pure = solid || fs == Number(0);
      const Number sm = fs * fs * (Number(3) - Number(2) * fs);
      const Number dsm = pure ? Number(0) : -prm.inv_mushy * Number(6) * fs * (Number(1) - fs);
      auto blend = [&](Number l, Number s, Number &v, Number &dv) {
        v = solid ? s : l + (s - l) * sm;
        dv = (s - l) * dsm;
      };
      Number cp, dcp, rho, drho, vc, dvc;
      blend(prm.l_k, prm.s_k, k, dk);
      blend(prm.l_cp, prm.s_cp, cp, dcp);
      Number ac, dac;
      apparent_capacity(prm, fs, ac, dac);
      cp += ac;
      dcp += dac;
      blend(prm.l_rho, prm.s_rho, rho, drho);
      blend(prm.l_rhocp, prm.s_rhocp, vc, dvc);
      if (prm.tmat == 3)
        {
          const Number w = prm.sharp ? (H > Number(0.5) ? Number(1) : Number(0)) : H;
          const Number one_w = Number(1) - w;
          k = one_w * prm.g_k + w * k;
          dk *= w;
          cp = one_w * prm.g_cp + w * cp;
          dcp *= w;
          rho = one_w * prm.g_rho + w * rho;
          drho *= w;
          vc = one_w * prm.g_rhocp + w * vc;
          dvc *= w;
        }
      if (prm.volume_specific)
        {
          rc = vc;
          drc = dvc;
        }
      else
        {
          rc = rho * cp;
          drc = dcp * rho + drho * cp;
        }
    }

    // Deterministic block sum (fixed butterfly + fixed-order wave sum).
    template <int THREADS>
    __device__ __forceinline__ double block_sum(double v, double *sh)
    {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
      const int wave = threadIdx.x >> 6;
      if ((threadIdx.x & 63) == 0)
        sh[wave] = v;
      __syncthreads();
      double s = 0.0;
      if (threadIdx.x == 0)
        {
          for (int w = 0; w < THREADS / 64; ++w)
            s += sh[w];
        }
      return s; // valid in thread 0
    }

    // Sum of partials[0..count) in a fixed order by one block.
    template <int THREADS>
    __device__ __forceinline__ double ordered_sum(const double *partials, int count, double *sh)
    {
      const double v = strided_sum(partials, threadIdx.x, count, THREADS);
      __syncthreads();
      return block_sum<THREADS>(v, sh);
    }

    // p . Ap of one colour pass of a vmult: each workgroup stores its sum in
    // partials[offset + block]; in the last pass (finalize), the workgroup that
    // finishes last adds all `total` slots in index order and sets scal->pap
    // and scal->alpha = rz / pap. Every thread of the workgroup must call it.
    template <int THREADS>
    __device__ __forceinline__ void pass_partials(double acc, double *partials, int offset,
                                                  int total, int finalize, SolverScalars *scal)
    {
      __shared__ double red[THREADS / 64];
      __shared__ bool is_last;
      const double bs = block_sum<THREADS>(acc, red);
      if (threadIdx.x == 0)
        {
          store_partial(partials + offset + blockIdx.x, bs);
          bool lst = false;
          if (finalize)
            {
              wait_vmem();
              const unsigned prev = atomicAdd(&scal->counter, 1u);
              lst = prev == gridDim.x - 1;
            }
          is_last = lst;
        }
      __syncthreads();
      if (is_last)
        {
          const double pap = ordered_sum<THREADS>(partials, total, red);
          if (threadIdx.x == 0)
            {
              scal->pap = pap;
              scal->alpha = scal->rz / pap;
              scal->counter = 0u;
            }
        }
    }


    // Per-cell LDS stride of an N^3 tensor: odd, so lanes working on
    // consecutive cells hit distinct banks.
    template <int N>
    constexpr int lds_stride()
    {
      return (N * N * N) % 2 == 0 ? N * N * N + 1 : N * N * N;
    }

    template <typename Number>
    __device__ __forceinline__ bool on_dirichlet(const CellParams<Number> &prm, int i, int d)
    {
      const int n1 = d == 0 ? prm.n1x : (d == 1 ? prm.n1y : prm.n1z);
      return (((prm.faces >> (2 * d)) & 1u) && i == 0) ||
             (((prm.faces >> (2 * d + 1)) & 1u) && i == n1 - 1);
    }

    struct Segment
    {
      int sx, cy, cz, cx0, ncell;
    };

    // A nodal vector field at the Gauss points of a row segment's cells (the
    // block-synchronous layout: one workgroup per segment of CPB cells, thread
    // = cell x line): component c of node I is v[c * n_dofs + I], read plain
    // (read_dof_values_plain) and interpolated with S in x, y and z into the
    // LDS tensors V0, V1, V2 (cell l at l * STR, index (qz * N + qy) * N + qx).
    // Every thread of the workgroup calls it; it ends with a barrier.
    template <int P, int CPB>
    __device__ __forceinline__ void nodal_vector_at_q(const CellParams<double> &prm,
                                                      const Segment &g, const double *v,
                                                      double *V0, double *V1, double *V2)
    {
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr ct::EO<N> TS = ct::VTab<P>::S;
      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int64_t nd = (int64_t)prm.n1x * prm.n1y * prm.n1z;
      double *W[3] = {V0 + l * STR, V1 + l * STR, V2 + l * STR};
      if (active)
        {
          const int64_t base = (int64_t)(g.cx0 + l) * P +
                               (int64_t)prm.n1x * (g.cy * P + s0 + (int64_t)prm.n1y * (g.cz * P + s1));
#pragma unroll
          for (int c = 0; c < 3; ++c)
            {
              double x[N], o[N];
#pragma unroll
              for (int a = 0; a < N; ++a)
                x[a] = cell_ok ? v[c * nd + base + a] : 0.0;
              eo_apply<N, 1>(TS, x, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                W[c][(s1 * N + s0) * N + q] = o[q];
            }
        }
      __syncthreads();
      if (active)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double x[N], o[N];
#pragma unroll
            for (int b = 0; b < N; ++b)
              x[b] = W[c][(s1 * N + b) * N + s0];
            eo_apply<N, 1>(TS, x, o);
#pragma unroll
            for (int b = 0; b < N; ++b)
              W[c][(s1 * N + b) * N + s0] = o[b];
          }
      __syncthreads();
      if (active)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double x[N], o[N];
#pragma unroll
            for (int z = 0; z < N; ++z)
              x[z] = W[c][(z * N + s1) * N + s0];
            eo_apply<N, 1>(TS, x, o);
#pragma unroll
            for (int z = 0; z < N; ++z)
              W[c][(z * N + s1) * N + s0] = o[z];
          }
      __syncthreads();
    }

    template <typename Number, int CPB>
    __device__ __forceinline__ Segment segment_of(const CellParams<Number> &prm, int colour)
    {
      const int nsx = (prm.nx + CPB - 1) / CPB;
      const int ox = colour & 1, oy = (colour >> 1) & 1, oz = (colour >> 2) & 1;
      const int msx = (nsx - ox + 1) / 2;
      const int my = (prm.ny - oy + 1) / 2;
      int b = blockIdx.x;
      Segment s;
      s.sx = ox + 2 * (b % msx);
      b /= msx;
      s.cy = oy + 2 * (b % my);
      s.cz = oz + 2 * (b / my);
      s.cx0 = s.sx * CPB;
      s.ncell = min(CPB, prm.nx - s.cx0);
      return s;
    }

    // Conductivity at quadrature point (qx, qy, qz) of cell (cx, cy, cz):
    // constant, separable sine, or the two-phase value staged in KBl.
    template <typename Number, int COEF, int N>
    __device__ __forceinline__ Number conductivity(const CellParams<Number> &prm,
                                                   const Number *KBl, int cx, int cy, int cz,
                                                   int qx, int qy, int qz)
    {
      if constexpr (COEF == kCoefConstant)
        return prm.k0;
      else if constexpr (COEF == kCoefSine)
        return prm.k0 + prm.kamp * prm.sin2pi[0][cx * N + qx] * prm.sin2pi[1][cy * N + qy] *
                          prm.sin2pi[2][cz * N + qz];
      else
        return KBl[(qz * N + qy) * N + qx];
    }

    // Coloured first-touch scatter of one cell x-line (b = s0, c = s1) after the
    // in-segment x-face exchange: the first cell in colour order to reach a node
    // stores its value (0 on a constrained node), later cells add; the x-line
    // end a = p is written by the segment's last cell only.
    template <int P, typename Number, int CPB>
    __device__ __forceinline__ void scatter_xline(const CellParams<Number> &prm, const Segment &g,
                                                  int l, int s0, int s1, const Number (&Y)[P + 1],
                                                  Number *__restrict__ dst)
    {
      constexpr int N = P + 1;
      const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
      const bool ysh = (s0 == 0 && g.cy > 0) || (s0 == P && g.cy < prm.ny - 1);
      const bool zsh = (s1 == 0 && g.cz > 0) || (s1 == P && g.cz < prm.nz - 1);
      const bool yz_first = !(ysh && (g.cy & 1)) && !(zsh && (g.cz & 1));
      const bool rowc = on_dirichlet(prm, Yi, 1) || on_dirichlet(prm, Zi, 2);
      const int X0 = (g.cx0 + l) * P;
      const int64_t base = X0 + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);
      const bool last = l == g.ncell - 1;
#pragma unroll
      for (int a = 0; a < N; ++a)
        {
          if (a == N - 1 && !last)
            continue;
          const bool xsh = (a == 0 && l == 0 && g.cx0 > 0) ||
                           (a == P && last && g.cx0 + g.ncell < prm.nx);
          const bool first = yz_first && !(xsh && (g.sx & 1));
          const bool con = rowc || on_dirichlet(prm, X0 + a, 0);
          Number *ptr = dst + base + a;
          if (con)
            {
              if (first)
                *ptr = Number(0);
            }
          else if (first)
            *ptr = Y[a];
          else
            *ptr = *ptr + Y[a];
        }
    }

    // Operator served by this workgroup in a batched launch: the last s with
    // block_off[s] <= blockIdx.x (block_off holds nops + 1 ascending entries).
    __device__ __forceinline__ int find_block_owner(const int *block_off, int nops)
    {
      __shared__ int s_sh;
      if (threadIdx.x == 0)
        {
          int lo = 0, hi = nops;
          while (hi - lo > 1)
            {
              const int mid = (lo + hi) / 2;
              if (block_off[mid] <= (int)blockIdx.x)
                lo = mid;
              else
                hi = mid;
            }
          s_sh = lo;
        }
      __syncthreads();
      return s_sh;
    }
  } // namespace
} // namespace mpdg
