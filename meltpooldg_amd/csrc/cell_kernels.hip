// Cell form of the heat-operator vmult for low degree (p <= 2).
//
// Each lane owns one hex cell and keeps its whole (p+1)^3 tensor in
// registers; a wavefront owns a tile of TX x TY = 64 cells in x-y and sweeps a
// chunk of cell layers along z. The sum factorisation (interpolate, gradient,
// q-point physics, integrate) needs no LDS and no barrier: every 1D
// contraction is a fixed sequence of register FMAs. Shared faces are summed
// deterministically:
//   * x-faces inside a tile row: shuffle from the lane one to the left,
//   * y-faces inside a tile: shuffle from the lane one row down (after the
//     x-exchange, so edge nodes collect all four cells),
//   * z-faces inside a chunk: carried in registers to the next layer,
//   * tile and chunk boundaries: eight ordered colour passes (tile parity in
//     x, y and chunk parity in z); a node is stored by the first pass that
//     touches it and read-modify-written by later ones.
// The p . Ap dot product of the PCG is formed in quadrature space per cell
// (u . A_c u = (S u) . (q-point operator applied to S u)).
//
// Reference behaviour (source/heat/heat_diffuse_operator.cpp): tangent_cell_loop
// :199-241 (read_dof_values with homogeneous constraints as 0, distribute_local_
// to_global dropping constrained rows), tangent_local_cell_operation :955-1086,
// material lookup :1133-1210.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    constexpr int kCellThreads = 256; // four independent wavefronts per workgroup

    // Bit 31 on a byte offset masks a node: vectors are shorter than 2^31 bytes,
    // so a masked offset plus any in-vector displacement stays out of range
    // (reads 0, drops the store) without wrapping.
    constexpr uint32_t kMaskBit = 0x80000000u;

    // y = M x with a compile-time dense matrix (exact zeros skipped).
    template <int N, typename Number>
    __device__ __forceinline__ void dense_apply(const ct::Mat<N> &M, const Number (&x)[N],
                                                Number (&y)[N])
    {
#pragma unroll
      for (int i = 0; i < N; ++i)
        {
          Number s = Number(0);
          bool first = true;
#pragma unroll
          for (int j = 0; j < N; ++j)
            if (M.a[i][j] != 0.0)
              {
                if (first)
                  s = Number(M.a[i][j]) * x[j];
                else
                  s += Number(M.a[i][j]) * x[j];
                first = false;
              }
          y[i] = s;
        }
    }

    // Apply M along direction D (0: x = last index, 1: y, 2: z) of an N^3 tensor.
    template <int D, int N, typename Number>
    __device__ __forceinline__ void tensor_apply(const ct::Mat<N> &M, Number (&t)[N][N][N])
    {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j)
          {
            Number x[N], y[N];
#pragma unroll
            for (int k = 0; k < N; ++k)
              x[k] = D == 0 ? t[i][j][k] : (D == 1 ? t[i][k][j] : t[k][i][j]);
            dense_apply<N>(M, x, y);
#pragma unroll
            for (int k = 0; k < N; ++k)
              {
                if (D == 0)
                  t[i][j][k] = y[k];
                else if (D == 1)
                  t[i][k][j] = y[k];
                else
                  t[k][i][j] = y[k];
              }
          }
    }

    // Keeps the scheduler from interleaving the phases of the cell work (which
    // would keep several gradient tensors alive at once and spill).
    __device__ __forceinline__ void phase_fence() { __builtin_amdgcn_sched_barrier(0); }

    // Cell face bits of node (a, b, c): x-lo, x-hi, y-lo, y-hi, z-lo, z-hi.
    template <int P>
    constexpr unsigned node_faces(int a, int b, int c)
    {
      return (a == 0 ? 1u : 0u) | (a == P ? 2u : 0u) | (b == 0 ? 4u : 0u) | (b == P ? 8u : 0u) |
             (c == 0 ? 16u : 0u) | (c == P ? 32u : 0u);
    }

    // Lane move by one along the wavefront (lane l receives lane l-1) through
    // DPP wave_shr:1; lane 0 receives 0.
    __device__ __forceinline__ float lane_shr1(float v)
    {
      return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                   0x138, 0xF, 0xF, true));
    }
    __device__ __forceinline__ double lane_shr1(double v)
    {
      const u32x2 w = __builtin_bit_cast(u32x2, v);
      u32x2 r;
      r.x = __builtin_amdgcn_update_dpp(0, (int)w.x, 0x138, 0xF, 0xF, true);
      r.y = __builtin_amdgcn_update_dpp(0, (int)w.y, 0x138, 0xF, 0xF, true);
      return __builtin_bit_cast(double, r);
    }

    // Body of the cell-form vmult for the items wave_first, wave_first +
    // wave_stride, ... (< items) of one colour; returns this lane's p . Ap share
    // when dot is set.
    template <int P, typename Number, int COEF, bool VEL>
    __device__ __forceinline__ double cell_body(const Number *__restrict__ src,
                                                Number *__restrict__ dst,
                                                const Number *__restrict__ heav,
                                                const CellParams<Number> &prm, const int colour,
                                                const int layers, const int items,
                                                const int wave_first, const int wave_stride,
                                                const bool dot)
    {
      constexpr int N = P + 1;
      constexpr bool H2 = COEF == kCoefHeaviside;
      constexpr bool SINE = COEF == kCoefSine;
      constexpr ct::Mat<N> MS = ct::CTab<P>::S;
      constexpr ct::Mat<N> MC = ct::CTab<P>::C;
      constexpr ct::Mat<N> MCW = ct::CTab<P>::CW;
      constexpr ct::Mat<N> MSW = ct::CTab<P>::SW;
      constexpr ct::Tab<N> T1 = ct::CTab<P>::t;
      constexpr uint32_t SZ = sizeof(Number);

      const uint32_t vbytes = vector_bytes(prm.n1x, prm.n1y, prm.n1z, (int)SZ);
      const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(src, vbytes);
      const __amdgpu_buffer_rsrc_t rdst = make_rsrc(dst, vbytes);
      const __amdgpu_buffer_rsrc_t rheav = make_rsrc(H2 ? heav : src, vbytes);
      const uint32_t sy = (uint32_t)prm.n1x * SZ;           // byte stride of a y step
      const uint32_t sz = (uint32_t)prm.n1x * prm.n1y * SZ; // byte stride of a z step

      const int lane = threadIdx.x & 63;
      const int tl = prm.tile_log2;
      const int TX = 1 << tl, TY = 64 >> tl;
      const int lx = lane & (TX - 1), ly = lane >> tl;
      const int ox = colour & 1, oy = (colour >> 1) & 1, oz = (colour >> 2) & 1;
      const int ntx = (prm.nx + TX - 1) >> tl;
      const int nty = (prm.ny + TY - 1) / TY;
      const int mtx = (ntx - ox + 1) >> 1, mty = (nty - oy + 1) >> 1;
      const unsigned f = prm.faces;
      // per-direction metric factors; a constant conductivity is folded in
      Number K[3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        K[d] = COEF == kCoefConstant ? prm.k0 * prm.kf[d] : prm.kf[d];

      double acc = 0.0;
      for (int item = wave_first; item < items; item += wave_stride)
        {
          const int tx = ox + 2 * (item % mtx);
          const int rr = item / mtx;
          const int ty = oy + 2 * (rr % mty);
          const int kz = oz + 2 * (rr / mty);
          const int cx = tx * TX + lx, cy = ty * TY + ly;
          const bool active = (cx < prm.nx) & (cy < prm.ny);
          const int cxc = min(cx, prm.nx - 1), cyc = min(cy, prm.ny - 1);
          const bool lastx = (lx == TX - 1) | (cx == prm.nx - 1);
          const bool lasty = (ly == TY - 1) | (cy == prm.ny - 1);
          // Dirichlet faces of the lane's cell (bits as node_faces)
          const unsigned cf_xy = ((f & 1u) && cx == 0 ? 1u : 0u) |
                                 (((f >> 1) & 1u) && cx == prm.nx - 1 ? 2u : 0u) |
                                 (((f >> 2) & 1u) && cy == 0 ? 4u : 0u) |
                                 (((f >> 3) & 1u) && cy == prm.ny - 1 ? 8u : 0u);
          // faces shared with a tile of an earlier colour pass (read-modify-write)
          const bool txo = tx & 1, tyo = ty & 1, kzo = kz & 1;
          const unsigned nf_xy = (txo && lx == 0 && cx > 0 ? 1u : 0u) |
                                 (txo && lastx && cx < prm.nx - 1 ? 2u : 0u) |
                                 (tyo && ly == 0 && cy > 0 ? 4u : 0u) |
                                 (tyo && lasty && cy < prm.ny - 1 ? 8u : 0u);
          const uint32_t m_idle = active ? 0u : kMaskBit;
          const uint32_t m_x = lastx ? 0u : kMaskBit;
          const uint32_t m_y = lasty ? 0u : kMaskBit;
          const int cz0 = kz * layers;
          const int nl = min(layers, prm.nz - cz0);

          // separable coefficient factors of this column of cells: sine
          // conductivity, and the in-plane velocity u_x = ux_x(x) ux_y(y),
          // u_y = uy_x(x) uy_y(y)
          Number s2x[N], s2y[N], ux_x[N], ux_y[N], uy_x[N], uy_y[N];
          if constexpr (SINE)
#pragma unroll
            for (int q = 0; q < N; ++q)
              {
                s2x[q] = prm.sin2pi[0][cxc * N + q];
                s2y[q] = prm.sin2pi[1][cyc * N + q];
              }
          if constexpr (VEL)
#pragma unroll
            for (int q = 0; q < N; ++q)
              {
                ux_x[q] = prm.vel_x[0][cxc * N + q];
                ux_y[q] = prm.vel_y[0][cyc * N + q];
                uy_x[q] = prm.vel_x[1][cxc * N + q];
                uy_y[q] = prm.vel_y[1][cyc * N + q];
              }

          Number carry[N][N]; // upper z-face of the previous layer, after the x/y exchange
#pragma unroll
          for (int b = 0; b < N; ++b)
#pragma unroll
            for (int a = 0; a < N; ++a)
              carry[b][a] = Number(0);

          for (int j = 0; j < nl; ++j)
            {
              const int cz = cz0 + j;
              const bool lastz = j == nl - 1;
              const unsigned cf = cf_xy | (((f >> 4) & 1u) && cz == 0 ? 16u : 0u) |
                                  (((f >> 5) & 1u) && cz == prm.nz - 1 ? 32u : 0u);
              const unsigned nf = nf_xy | (kzo && j == 0 && cz > 0 ? 16u : 0u) |
                                  (kzo && lastz && cz < prm.nz - 1 ? 32u : 0u);
              const uint32_t m_z = lastz ? 0u : kMaskBit;
              const uint32_t base =
                (uint32_t)(cxc * P + prm.n1x * (cyc * P + prm.n1y * (cz * P))) * SZ;
              const uint32_t gbase = base | m_idle;

              // ---- gather; idle lanes read 0, constrained nodes are zeroed below --
              Number v[N][N][N];
              Number hq[H2 ? N : 1][H2 ? N : 1][H2 ? N : 1];
#pragma unroll
              for (int c = 0; c < N; ++c)
#pragma unroll
                for (int b = 0; b < N; ++b)
#pragma unroll
                  for (int a = 0; a < N; ++a)
                    {
                      const uint32_t off = gbase + (b * sy + c * sz) + a * SZ;
                      v[c][b][a] = buf_load(rsrc, off, Number());
                      if constexpr (H2)
                        hq[c][b][a] = buf_load(rheav, off, Number());
                    }
              Number s2z[N];
              if constexpr (SINE)
#pragma unroll
                for (int q = 0; q < N; ++q)
                  s2z[q] = prm.sin2pi[2][cz * N + q];
              if (cf != 0u) // homogeneous constraints read as 0 (rare: boundary cells)
#pragma unroll
                for (int c = 0; c < N; ++c)
#pragma unroll
                  for (int b = 0; b < N; ++b)
#pragma unroll
                    for (int a = 0; a < N; ++a)
                      if (cf & node_faces<P>(a, b, c))
                        v[c][b][a] = Number(0);

              phase_fence();
              // ---- values at the quadrature points ---------------------------------
              tensor_apply<0, N>(MS, v);
              tensor_apply<1, N>(MS, v);
              tensor_apply<2, N>(MS, v);
              if constexpr (H2)
                {
                  tensor_apply<0, N>(MS, hq);
                  tensor_apply<1, N>(MS, hq);
                  tensor_apply<2, N>(MS, hq);
                }

              phase_fence();
              // ---- q-point operator in weight-free form ----------------------------
              // o = mass v + sum_d W^-1 C_d^T W (K_d k C_d v) + rho_cp (u . grad v);
              // the weights sit in the integration matrix S^T W.
              Number o[N][N][N];
              // z-lines: mass term and axial convection
#pragma unroll
              for (int qy = 0; qy < N; ++qy)
#pragma unroll
                for (int qx = 0; qx < N; ++qx)
                  {
                    Number x[N], g[N], fl[N], t[N];
#pragma unroll
                    for (int qz = 0; qz < N; ++qz)
                      x[qz] = v[qz][qy][qx];
                    dense_apply<N>(MC, x, g);
                    Number rcq[N];
                    if constexpr (SINE)
                      {
                        const Number kxy = prm.kamp * s2x[qx] * s2y[qy];
#pragma unroll
                        for (int qz = 0; qz < N; ++qz)
                          fl[qz] = (prm.k0 + kxy * s2z[qz]) * g[qz];
                      }
                    else if constexpr (H2)
                      {
#pragma unroll
                        for (int qz = 0; qz < N; ++qz)
                          {
                            Number k;
                            two_phase(prm, hq[qz][qy][qx], rcq[qz], k);
                            fl[qz] = k * g[qz];
                          }
                      }
                    else
#pragma unroll
                      for (int qz = 0; qz < N; ++qz)
                        fl[qz] = g[qz];
                    dense_apply<N>(MCW, fl, t);
#pragma unroll
                    for (int qz = 0; qz < N; ++qz)
                      {
                        const Number mass = H2 ? rcq[qz] * prm.dtinv_vol : prm.mass;
                        Number val = mass * x[qz] + K[2] * t[qz];
                        if constexpr (VEL)
                          val += (H2 ? rcq[qz] : prm.rho_cp0) * prm.vf[2] * prm.w0 * g[qz];
                        o[qz][qy][qx] = val;
                      }
                  }
              phase_fence();
              // y-lines
#pragma unroll
              for (int qz = 0; qz < N; ++qz)
#pragma unroll
                for (int qx = 0; qx < N; ++qx)
                  {
                    Number x[N], g[N], fl[N], t[N];
#pragma unroll
                    for (int qy = 0; qy < N; ++qy)
                      x[qy] = v[qz][qy][qx];
                    dense_apply<N>(MC, x, g);
                    Number rcq[N];
                    if constexpr (SINE)
                      {
                        const Number kxz = prm.kamp * s2x[qx] * s2z[qz];
#pragma unroll
                        for (int qy = 0; qy < N; ++qy)
                          fl[qy] = (prm.k0 + kxz * s2y[qy]) * g[qy];
                      }
                    else if constexpr (H2)
                      {
#pragma unroll
                        for (int qy = 0; qy < N; ++qy)
                          {
                            Number k;
                            two_phase(prm, hq[qz][qy][qx], rcq[qy], k);
                            fl[qy] = k * g[qy];
                          }
                      }
                    else
#pragma unroll
                      for (int qy = 0; qy < N; ++qy)
                        fl[qy] = g[qy];
                    dense_apply<N>(MCW, fl, t);
#pragma unroll
                    for (int qy = 0; qy < N; ++qy)
                      {
                        Number val = o[qz][qy][qx] + K[1] * t[qy];
                        if constexpr (VEL)
                          val += (H2 ? rcq[qy] : prm.rho_cp0) * prm.vf[1] *
                                 (uy_x[qx] * uy_y[qy]) * g[qy];
                        o[qz][qy][qx] = val;
                      }
                  }
              phase_fence();
              // x-lines, and the cell's p . Ap share (S u) . W (q-point operator)
              Number cellsum = Number(0);
#pragma unroll
              for (int qz = 0; qz < N; ++qz)
#pragma unroll
                for (int qy = 0; qy < N; ++qy)
                  {
                    Number x[N], g[N], fl[N], t[N];
#pragma unroll
                    for (int qx = 0; qx < N; ++qx)
                      x[qx] = v[qz][qy][qx];
                    dense_apply<N>(MC, x, g);
                    Number rcq[N];
                    if constexpr (SINE)
                      {
                        const Number kyz = prm.kamp * s2y[qy] * s2z[qz];
#pragma unroll
                        for (int qx = 0; qx < N; ++qx)
                          fl[qx] = (prm.k0 + kyz * s2x[qx]) * g[qx];
                      }
                    else if constexpr (H2)
                      {
#pragma unroll
                        for (int qx = 0; qx < N; ++qx)
                          {
                            Number k;
                            two_phase(prm, hq[qz][qy][qx], rcq[qx], k);
                            fl[qx] = k * g[qx];
                          }
                      }
                    else
#pragma unroll
                      for (int qx = 0; qx < N; ++qx)
                        fl[qx] = g[qx];
                    dense_apply<N>(MCW, fl, t);
#pragma unroll
                    for (int qx = 0; qx < N; ++qx)
                      {
                        Number val = o[qz][qy][qx] + K[0] * t[qx];
                        if constexpr (VEL)
                          val += (H2 ? rcq[qx] : prm.rho_cp0) * prm.vf[0] *
                                 (ux_x[qx] * ux_y[qy]) * g[qx];
                        o[qz][qy][qx] = val;
                        cellsum += Number(T1.w[qx] * T1.w[qy] * T1.w[qz]) * (x[qx] * val);
                      }
                  }
              if (dot)
                acc += (double)cellsum;

              phase_fence();
              // ---- integrate: (S^T W) along z, y, x ----------------------------------
              tensor_apply<2, N>(MSW, o);
              tensor_apply<1, N>(MSW, o);
              tensor_apply<0, N>(MSW, o);

              phase_fence();
              // ---- shared faces inside the tile and the chunk ------------------------
#pragma unroll
              for (int c = 0; c < N; ++c)
#pragma unroll
                for (int b = 0; b < N; ++b)
                  {
                    const Number t = lane_shr1(o[c][b][P]);
                    if (lx > 0)
                      o[c][b][0] += t;
                  }
#pragma unroll
              for (int c = 0; c < N; ++c)
#pragma unroll
                for (int a = 0; a < N; ++a)
                  {
                    const Number t = __shfl_up(o[c][P][a], TX, 64);
                    if (ly > 0)
                      o[c][0][a] += t;
                  }
#pragma unroll
              for (int b = 0; b < N; ++b)
#pragma unroll
                for (int a = 0; a < N; ++a)
                  {
                    o[0][b][a] += carry[b][a];
                    carry[b][a] = o[P][b][a];
                  }
              if (cf != 0u) // constrained rows are dropped (stored as 0)
#pragma unroll
                for (int c = 0; c < N; ++c)
#pragma unroll
                  for (int b = 0; b < N; ++b)
#pragma unroll
                    for (int a = 0; a < N; ++a)
                      if (cf & node_faces<P>(a, b, c))
                        o[c][b][a] = Number(0);

              phase_fence();
              // ---- coloured scatter --------------------------------------------------
              // Nodes on a face shared with a tile of an earlier pass are added
              // with an atomic: within one pass a node belongs to one tile only
              // and the passes run in order, so every sum is formed in a fixed
              // order. All other written nodes are plain stores. Masked offsets
              // (bit 31) drop the access.
              uint32_t skip[6], keep_store[6], take_add[6];
#pragma unroll
              for (int fi = 0; fi < 6; ++fi)
                {
                  const bool hi = fi & 1;
                  const uint32_t m_skip = !hi ? 0u : (fi == 1 ? m_x : (fi == 3 ? m_y : m_z));
                  const bool shared = (nf >> fi) & 1u;
                  skip[fi] = m_skip;
                  keep_store[fi] = shared ? kMaskBit : 0u;
                  take_add[fi] = shared ? 0u : kMaskBit;
                }
              auto face_or = [&](const uint32_t (&m)[6], int a, int b, int c) -> uint32_t {
                return (a == 0 ? m[0] : 0u) | (a == P ? m[1] : 0u) | (b == 0 ? m[2] : 0u) |
                       (b == P ? m[3] : 0u) | (c == 0 ? m[4] : 0u) | (c == P ? m[5] : 0u);
              };
              auto face_and = [&](const uint32_t (&m)[6], int a, int b, int c) -> uint32_t {
                return (a == 0 ? m[0] : ~0u) & (a == P ? m[1] : ~0u) & (b == 0 ? m[2] : ~0u) &
                       (b == P ? m[3] : ~0u) & (c == 0 ? m[4] : ~0u) & (c == P ? m[5] : ~0u);
              };
              const uint32_t wbase = base | m_idle;
#pragma unroll
              for (int c = 0; c < N; ++c)
#pragma unroll
                for (int b = 0; b < N; ++b)
#pragma unroll
                  for (int a = 0; a < N; ++a)
                    {
                      const uint32_t rel = (b * sy + c * sz) + a * SZ;
                      const uint32_t w = wbase | face_or(skip, a, b, c);
                      buf_store(rdst, (w | face_or(keep_store, a, b, c)) + rel, o[c][b][a]);
                    }
              if (nf != 0u)
#pragma unroll
                for (int c = 0; c < N; ++c)
#pragma unroll
                  for (int b = 0; b < N; ++b)
#pragma unroll
                    for (int a = 0; a < N; ++a)
                      if (node_faces<P>(a, b, c) != 0u)
                        {
                          const uint32_t rel = (b * sy + c * sz) + a * SZ;
                          const uint32_t w = wbase | face_or(skip, a, b, c);
                          buf_add(rdst, (w | face_and(take_add, a, b, c)) + rel, o[c][b][a]);
                        }
            }
        }
      return acc;
    }

    template <int P, typename Number, int COEF, bool VEL, int MINW>
    __global__ void __launch_bounds__(kCellThreads, MINW)
      cell_kernel(const Number *__restrict__ src, Number *__restrict__ dst,
                  const Number *__restrict__ heav, const CellParams<Number> prm, const int colour,
                  const int layers, const int items, double *__restrict__ partials,
                  const int partial_offset, const int partial_total, const int finalize,
                  SolverScalars *__restrict__ scal)
    {
      if (scal != nullptr && scal->done)
        return;
      constexpr int WPB = kCellThreads / 64;
      const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); // wave-uniform
      const double acc = cell_body<P, Number, COEF, VEL>(src, dst, heav, prm, colour, layers, items,
                                                         blockIdx.x * WPB + w, gridDim.x * WPB,
                                                         partials != nullptr);
      if (partials != nullptr)
        pass_partials<kCellThreads>(acc, partials, partial_offset, partial_total, finalize, scal);
    }

    // Batched cell-form vmult: workgroup b serves operator s with
    // block_off[s] <= b < block_off[s+1]; its wavefronts loop over that
    // operator's items of this colour.
    template <int P, typename Number, int COEF, bool VEL, int MINW>
    __global__ void __launch_bounds__(kCellThreads, MINW)
      cell_batch_kernel(const BatchOperator<Number> *__restrict__ ops,
                        const int *__restrict__ block_off, const int nops, const int colour,
                        double *__restrict__ partials, const SolverScalars *__restrict__ scal)
    {
      constexpr int WPB = kCellThreads / 64;
      const int s = find_block_owner(block_off, nops);
      const BatchOperator<Number> &op = ops[s];
      double acc = 0.0;
      const bool skip = scal != nullptr && scal[op.solve].done;
      if (!skip)
        acc = cell_body<P, Number, COEF, VEL>(
          op.src, op.dst, op.heav, op.prm, colour, op.prm.zlayers, op.items[colour],
          (blockIdx.x - block_off[s]) * WPB + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6),
          (block_off[s + 1] - block_off[s]) * WPB, partials != nullptr);
      if (partials != nullptr)
        {
          __shared__ double red[WPB];
          const double bs = block_sum<kCellThreads>(acc, red);
          if (threadIdx.x == 0)
            partials[blockIdx.x] = bs;
        }
    }

    // Register budget per degree and precision (HIP's minimum waves per SIMD):
    // 4 -> 128 VGPRs, 3 -> 168.
    template <int P, typename Number, int COEF>
    constexpr int cell_min_waves()
    {
      return (P == 2 && sizeof(Number) == 8) ? 3 : 4;
    }

    template <typename Number, int P, int COEF, bool VEL>
    hipError_t cell_passes(const CellParams<Number> &prm, const Number *src, Number *dst,
                           const Number *heav, double *partials, SolverScalars *scal,
                           hipStream_t stream)
    {
      auto kernel = cell_kernel<P, Number, COEF, VEL, cell_min_waves<P, Number, COEF>()>;
      static int resident = 0;
      if (resident == 0)
        {
          int nb = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kCellThreads, 0) != hipSuccess || nb < 1)
            nb = 1;
          resident = nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
        }
      const int per_cu = prm.blocks_per_cu > 0 ? prm.blocks_per_cu : resident;
      const VmultGrid vg = cell_grid(prm.nx, prm.ny, prm.nz, prm.tile_log2, prm.zlayers,
                                     per_cu * device_cu_count());
      for (int c = 0; c < 8; ++c)
        {
          if (vg.grid[c] == 0)
            continue;
          kernel<<<vg.grid[c], kCellThreads, 0, stream>>>(src, dst, heav, prm, c, prm.zlayers,
                                                          vg.items[c], partials, vg.offset[c],
                                                          vg.total, c == vg.last ? 1 : 0, scal);
        }
      return hipGetLastError();
    }

    template <typename Number, int P>
    hipError_t cell_deg(int coef, bool vel, const CellParams<Number> &prm, const Number *src,
                        Number *dst, const Number *heav, double *partials, SolverScalars *scal,
                        hipStream_t s)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return cell_passes<Number, P, 0, false>(prm, src, dst, heav, partials, scal, s);
          case 1: return cell_passes<Number, P, 0, true>(prm, src, dst, heav, partials, scal, s);
          case 2: return cell_passes<Number, P, 1, false>(prm, src, dst, heav, partials, scal, s);
          case 3: return cell_passes<Number, P, 1, true>(prm, src, dst, heav, partials, scal, s);
          case 4: return cell_passes<Number, P, 2, false>(prm, src, dst, heav, partials, scal, s);
          case 5: return cell_passes<Number, P, 2, true>(prm, src, dst, heav, partials, scal, s);
          default: return hipErrorInvalidValue;
        }
    }

    template <typename Number, int P, int COEF, bool VEL>
    hipError_t cell_batch_passes(const BatchOperator<Number> *ops, const int *block_off, int nops,
                                 const BatchVmultPlan &plan, double *partials,
                                 const SolverScalars *scal, hipStream_t stream)
    {
      auto kernel = cell_batch_kernel<P, Number, COEF, VEL, cell_min_waves<P, Number, COEF>()>;
      for (int c = 0; c < 8; ++c)
        {
          if (plan.grid[c] == 0)
            continue;
          kernel<<<plan.grid[c], kCellThreads, 0, stream>>>(
            ops, block_off + c * (nops + 1), nops, c, partials ? partials + plan.pofs[c] : nullptr,
            scal);
        }
      return hipGetLastError();
    }

    template <typename Number, int P>
    hipError_t cell_batch_deg(int coef, bool vel, const BatchOperator<Number> *ops,
                              const int *block_off, int nops, const BatchVmultPlan &plan,
                              double *partials, const SolverScalars *scal, hipStream_t s)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return cell_batch_passes<Number, P, 0, false>(ops, block_off, nops, plan, partials, scal, s);
          case 1: return cell_batch_passes<Number, P, 0, true>(ops, block_off, nops, plan, partials, scal, s);
          case 2: return cell_batch_passes<Number, P, 1, false>(ops, block_off, nops, plan, partials, scal, s);
          case 3: return cell_batch_passes<Number, P, 1, true>(ops, block_off, nops, plan, partials, scal, s);
          case 4: return cell_batch_passes<Number, P, 2, false>(ops, block_off, nops, plan, partials, scal, s);
          case 5: return cell_batch_passes<Number, P, 2, true>(ops, block_off, nops, plan, partials, scal, s);
          default: return hipErrorInvalidValue;
        }
    }

    template <typename Number, int P>
    int cell_resident_p()
    {
      auto kernel = cell_batch_kernel<P, Number, 0, false, cell_min_waves<P, Number, 0>()>;
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kCellThreads, 0) != hipSuccess || nb < 1)
        nb = 1;
      return nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
    }
  } // namespace

  VmultGrid cell_grid(int nx, int ny, int nz, int tile_log2, int layers, int max_blocks)
  {
    constexpr int WPB = kCellThreads / 64;
    VmultGrid g{};
    const int TX = 1 << tile_log2, TY = 64 >> tile_log2;
    const int ntx = (nx + TX - 1) / TX, nty = (ny + TY - 1) / TY;
    const int nkz = (nz + layers - 1) / layers;
    int off = 0;
    g.last = -1;
    for (int c = 0; c < 8; ++c)
      {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        const int m = ((ntx - ox + 1) / 2) * ((nty - oy + 1) / 2) * ((nkz - oz + 1) / 2);
        const int blocks = (m + WPB - 1) / WPB;
        g.items[c] = m;
        g.grid[c] = blocks < max_blocks ? blocks : max_blocks;
        g.offset[c] = off;
        off += g.grid[c];
        if (m > 0)
          g.last = c;
      }
    g.total = off;
    return g;
  }

  int cell_items_per_block() { return kCellThreads / 64; }

  int cell_resident_blocks(int degree, bool fp32)
  {
    if (degree == 1)
      return fp32 ? cell_resident_p<float, 1>() : cell_resident_p<double, 1>();
    return fp32 ? cell_resident_p<float, 2>() : cell_resident_p<double, 2>();
  }

  template <typename Number>
  hipError_t launch_vmult_cell(int degree, int coef, bool vel, const CellParams<Number> &prm,
                               const Number *src, Number *dst, const Number *heaviside,
                               double *partials, SolverScalars *scal, hipStream_t stream)
  {
    if (prm.tile_log2 < 3 || prm.tile_log2 > 6)
      return hipErrorInvalidValue;
    switch (degree)
      {
        case 1: return cell_deg<Number, 1>(coef, vel, prm, src, dst, heaviside, partials, scal, stream);
        case 2: return cell_deg<Number, 2>(coef, vel, prm, src, dst, heaviside, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  template <typename Number>
  hipError_t launch_vmult_cell_batch(int degree, int coef, bool vel,
                                     const BatchOperator<Number> *ops, const int *block_off,
                                     int nops, const BatchVmultPlan &plan, double *partials,
                                     const SolverScalars *scal, hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return cell_batch_deg<Number, 1>(coef, vel, ops, block_off, nops, plan, partials, scal, stream);
        case 2: return cell_batch_deg<Number, 2>(coef, vel, ops, block_off, nops, plan, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

#define MPDG_CELL_INST(NUM)                                                                      \
  template hipError_t launch_vmult_cell<NUM>(int, int, bool, const CellParams<NUM> &, const NUM *, \
                                             NUM *, const NUM *, double *, SolverScalars *,       \
                                             hipStream_t);                                        \
  template hipError_t launch_vmult_cell_batch<NUM>(int, int, bool, const BatchOperator<NUM> *,   \
                                                   const int *, int, const BatchVmultPlan &,     \
                                                   double *, const SolverScalars *, hipStream_t);
  MPDG_CELL_INST(double)
  MPDG_CELL_INST(float)
#undef MPDG_CELL_INST
} // namespace mpdg
