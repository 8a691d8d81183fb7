// Element-matrix form of the heat-operator vmult on the matrix cores, for
// operators whose cell matrix is the same in every cell: affine box cells
// with a constant material and a constant (or no) velocity, degree <= 2.
//
// Then the cell matrix A_e (test x trial nodes of the Q(p) cell, the mass,
// Laplace and convection terms of tangent_local_cell_operation,
// source/heat/heat_diffuse_operator.cpp:1029-1042, integrated with the same
// QGauss(p+1) rule, which is exact for these integrands) is built once on the
// host, and a cell's result is y_c = A_e u_c: a dense product, batched over 16
// cells per wavefront with v_mfma_*_16x16x4 (A: 16 rows x 4 columns of A_e,
// one element per lane, row = lane % 16, column = lane / 16; B: 4 nodes x 16
// cells, node = lane / 16, cell = lane % 16). The product of A row i for cell j
// lands in lane j + 16 (i / 4), register i % 4 with the f32 instruction and in
// lane j + 16 (i % 4), register i / 4 with the f64 one (measured on gfx950).
// Lane l = j + 16 g serves cell j of its tile
// and the node slots t of group g: slot t holds node g + 4t of the cell
// (lexicographic a + n b + n^2 c), both as the B operand of k-step t and as
// output register t % 4 of row block t / 4. The rows of A_e are permuted to
// match (the host packs A_e per lane), so a lane ends up with u and y of the
// same nodes: the gather, p . Ap and the scatter need no lane exchange.
//
// Scatter: a wavefront owns a tile of 16 cells (TX x TY in x-y) and a z-chunk
// of cell layers. Its cells' results are summed per layer in the wavefront's
// own LDS block of the tile's nodes, in four steps by the x-y parity of the
// cell (cells of one parity share no node, so each step writes every block
// entry at most once); the top node plane is carried to the next layer.
// Tiles and chunks are coloured by parity in x, y and z and run as eight
// ordered passes: a node shared with an item of an earlier pass is added with
// an atomic, every other node is stored (the first pass to reach a node is
// the one whose item is even in every direction the node is shared in).
// Every sum is formed in a fixed order, so results are bitwise reproducible.
// Dirichlet nodes read as 0 and are written as 0.
#include "device_util.hpp"

namespace mpdg
{
  namespace
  {
    constexpr int kEmThreads = 256; // four independent wavefronts per workgroup
    typedef float em_v4f __attribute__((ext_vector_type(4)));
    typedef double em_v4d __attribute__((ext_vector_type(4)));

    template <typename Number>
    struct EmAcc;
    template <>
    struct EmAcc<float>
    {
      typedef em_v4f type;
    };
    template <>
    struct EmAcc<double>
    {
      typedef em_v4d type;
    };

    __device__ __forceinline__ em_v4f em_mfma(float a, float b, em_v4f c)
    {
      return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ __forceinline__ em_v4d em_mfma(double a, double b, em_v4d c)
    {
      return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }

    __device__ __forceinline__ void em_wave_sync()
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    template <int P, typename Number, int TL>
    __global__ void __launch_bounds__(kEmThreads)
      em_kernel(const Number *__restrict__ src, Number *__restrict__ dst,
                const CellParams<Number> prm, const int colour, const int layers, const int items,
                double *__restrict__ partials, const int partial_offset, const int partial_total,
                const int finalize, SolverScalars *__restrict__ scal)
    {
      constexpr int N = P + 1;
      constexpr int NN = N * N * N;
      constexpr int T = EmShape<P>::T;
      constexpr int RB = EmShape<P>::RB;
      constexpr int TX = 1 << TL, TY = 16 >> TL;
      constexpr int NX = TX * P + 1, NY = TY * P + 1, PLN = NX * NY;
      constexpr int BUF = N * PLN;                 // one parity class's node block
      constexpr int SLOTS = (N * PLN + 63) / 64;   // node-block entries per lane
      constexpr uint32_t SZ = sizeof(Number);
      if (scal != nullptr && scal->done)
        return;
      // per wavefront: the four parity classes' node blocks and the carried plane
      __shared__ Number lds[kEmThreads / 64][4 * BUF + PLN];
      const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
      Number *cls = lds[w];
      Number *carry = lds[w] + 4 * BUF;
      const int lane = threadIdx.x & 63;
      const int j = lane & 15, g = lane >> 4;
      const int jx = j & (TX - 1), jy = j >> TL;
      const int myclass = (jx & 1) | ((jy & 1) << 1);

      Number Aop[RB][T];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int ks = 0; ks < T; ++ks)
          Aop[rb][ks] = prm.em[(rb * T + ks) * 64 + lane];
      // the lane's node slots (node g + 4t of its cell): block offsets
      int toff[T];
      unsigned tvalid = 0;
#pragma unroll
      for (int t = 0; t < T; ++t)
        {
          const int m = g + 4 * t;
          const int mm = m < NN ? m : 0;
          if (m < NN)
            tvalid |= 1u << t;
          toff[t] = myclass * BUF + (mm / (N * N)) * PLN + (jy * P + (mm / N) % N) * NX + jx * P + mm % N;
        }

      const uint32_t vbytes = vector_bytes(prm.n1x, prm.n1y, prm.n1z, (int)SZ);
      const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(src, vbytes);
      const __amdgpu_buffer_rsrc_t rdst = make_rsrc(dst, vbytes);
      const unsigned f = prm.faces;
      const int ox = colour & 1, oy = (colour >> 1) & 1, oz = (colour >> 2) & 1;
      const int ntx = (prm.nx + TX - 1) >> TL;
      const int nty = (prm.ny + TY - 1) / TY;
      const int mtx = (ntx - ox + 1) >> 1, mty = (nty - oy + 1) >> 1;
      const uint32_t sxy = (uint32_t)prm.n1x * SZ, sxyz = (uint32_t)prm.n1x * prm.n1y * SZ;

      double acc = 0.0;
      for (int item = blockIdx.x * (kEmThreads / 64) + w; item < items;
           item += gridDim.x * (kEmThreads / 64))
        {
          const int tx = ox + 2 * (item % mtx);
          const int rr = item / mtx;
          const int ty = oy + 2 * (rr % mty);
          const int kz = oz + 2 * (rr / mty);
          const int cx = tx * TX + jx, cy = ty * TY + jy;
          const bool active = (cx < prm.nx) & (cy < prm.ny);
          const int cz0 = kz * layers;
          const int nl = min(layers, prm.nz - cz0);
          const int ncx = min(TX, prm.nx - tx * TX), ncy = min(TY, prm.ny - ty * TY);
          const int xmax = ncx * P, ymax = ncy * P; // last node column / row of the tile
          const int X0 = tx * TX * P, Y0 = ty * TY * P;
          const bool txo = tx & 1, tyo = ty & 1, kzo = kz & 1;
          // gather plan of the lane's nodes: the x-y part of the byte offset;
          // gskip bit t: node absent, cell idle or on a Dirichlet face in x-y
          uint32_t goff[T];
          unsigned gskip = 0;
#pragma unroll
          for (int t = 0; t < T; ++t)
            {
              const int m = g + 4 * t;
              const int mm = m < NN ? m : 0;
              const int X = cx * P + mm % N, Y = cy * P + (mm / N) % N;
              const bool skip = m >= NN || !active || ((f & 1u) && X == 0) ||
                                ((f & 2u) && X == prm.n1x - 1) || ((f & 4u) && Y == 0) ||
                                ((f & 8u) && Y == prm.n1y - 1);
              gskip |= skip ? 1u << t : 0u;
              goff[t] = (uint32_t)X * SZ + (uint32_t)Y * sxy;
            }
          // store plan of the lane's node-block slots k (entry lane + 64 k =
          // plane c, row yl, column xl): the parity classes holding a cell
          // around the node (only those blocks were written this layer), 4 bits
          // per slot; bits of slots that exist, are Dirichlet in x-y, or are
          // shared in x-y with an item of an earlier pass
          unsigned cmask = 0, sok = 0, scon = 0, sadd = 0;
#pragma unroll
          for (int k = 0; k < SLOTS; ++k)
            {
              const int i = lane + 64 * k;
              const int r = i % PLN;
              const int yl = r / NX, xl = r % NX;
              const bool ok = i < N * PLN && xl <= xmax && yl <= ymax;
              const int ca = min(xl / P, ncx - 1), cb = (xl % P == 0 && xl > 0) ? xl / P - 1 : ca;
              const int ra = min(yl / P, ncy - 1), rb2 = (yl % P == 0 && yl > 0) ? yl / P - 1 : ra;
              unsigned m = (1u << ((ca & 1) | ((ra & 1) << 1))) | (1u << ((cb & 1) | ((ra & 1) << 1))) |
                           (1u << ((ca & 1) | ((rb2 & 1) << 1))) | (1u << ((cb & 1) | ((rb2 & 1) << 1)));
              cmask |= m << (4 * k);
              const int X = X0 + xl, Y = Y0 + yl;
              const bool con = ((f & 1u) && X == 0) || ((f & 2u) && X == prm.n1x - 1) ||
                               ((f & 4u) && Y == 0) || ((f & 8u) && Y == prm.n1y - 1);
              const bool shx = (xl == 0 && X0 > 0) || (xl == xmax && X < prm.n1x - 1);
              const bool shy = (yl == 0 && Y0 > 0) || (yl == ymax && Y < prm.n1y - 1);
              sok |= ok ? 1u << k : 0u;
              scon |= con ? 1u << k : 0u;
              sadd |= ((shx && txo) || (shy && tyo)) ? 1u << k : 0u;
            }

          auto gather = [&](int cz, bool none, Number (&v)[T]) {
#pragma unroll
            for (int t = 0; t < T; ++t)
              {
                const int Z = cz * P + (g + 4 * t) / (N * N);
                const bool skip = none || ((gskip >> t) & 1u) || ((f & 16u) && Z == 0) ||
                                  ((f & 32u) && Z == prm.n1z - 1);
                v[t] = buf_load(rsrc, skip ? kOutOfRange : goff[t] + (uint32_t)Z * sxyz, Number());
              }
          };
          Number u[T];
          gather(cz0, false, u);

          for (int l = 0; l < nl; ++l)
            {
              const int cz = cz0 + l;
              const bool last = l == nl - 1;
              // ---- y = A_e u for the 16 cells of the tile ------------------------
              typename EmAcc<Number>::type d[RB];
#pragma unroll
              for (int rb = 0; rb < RB; ++rb)
                d[rb] = typename EmAcc<Number>::type{0, 0, 0, 0};
#pragma unroll
              for (int ks = 0; ks < T; ++ks)
#pragma unroll
                for (int rb = 0; rb < RB; ++rb)
                  d[rb] = em_mfma(Aop[rb][ks], u[ks], d[rb]);
              // the next layer's gather, in flight during the scatter
              Number un[T];
              gather(cz + 1, last, un);
              // ---- each cell writes its results into its parity class's block ---
#pragma unroll
              for (int t = 0; t < T; ++t)
                {
                  const Number y = d[t / 4][t % 4];
                  acc += (double)u[t] * (double)y;
                  if (((tvalid >> t) & 1u) && active)
                    cls[toff[t]] = y;
                }
              em_wave_sync();
              // ---- sum the classes per node and store the finished planes --------
              const int np = last ? N : P;
#pragma unroll
              for (int k = 0; k < SLOTS; ++k)
                {
                  const int i = lane + 64 * k;
                  const int c = i / PLN;
                  if (!((sok >> k) & 1u))
                    continue;
                  const unsigned m = (cmask >> (4 * k)) & 15u;
                  Number v = Number(0);
#pragma unroll
                  for (int q = 0; q < 4; ++q)
                    if (m & (1u << q))
                      v += cls[q * BUF + i];
                  if (c == 0 && l > 0)
                    v += carry[i];
                  if (c == P && !last)
                    {
                      carry[i - P * PLN] = v; // the next layer's bottom plane
                      continue;
                    }
                  if (c >= np)
                    continue;
                  const int r = i % PLN;
                  const int X = X0 + r % NX, Y = Y0 + r / NX, Z = cz * P + c;
                  const bool con = ((scon >> k) & 1u) || ((f & 16u) && Z == 0) ||
                                   ((f & 32u) && Z == prm.n1z - 1);
                  const bool shz = (c == 0 && l == 0 && cz > 0) || (c == P && last && cz < prm.nz - 1);
                  const bool add = ((sadd >> k) & 1u) || (shz && kzo);
                  const uint32_t off = (uint32_t)X * SZ + (uint32_t)Y * sxy + (uint32_t)Z * sxyz;
                  const Number out = con ? Number(0) : v;
                  if (add)
                    buf_add(rdst, off, out);
                  else
                    buf_store(rdst, off, out);
                }
              em_wave_sync();
#pragma unroll
              for (int t = 0; t < T; ++t)
                u[t] = un[t];
            }
        }
      if (partials != nullptr)
        pass_partials<kEmThreads>(acc, partials, partial_offset, partial_total, finalize, scal);
    }

    template <typename Number, int P, int TL>
    hipError_t em_passes(const CellParams<Number> &prm, const Number *src, Number *dst,
                         double *partials, SolverScalars *scal, hipStream_t stream)
    {
      auto kernel = em_kernel<P, Number, TL>;
      static int resident = 0;
      if (resident == 0)
        {
          int nb = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kEmThreads, 0) != hipSuccess || nb < 1)
            nb = 1;
          resident = nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
        }
      const int per_cu = prm.blocks_per_cu > 0 ? prm.blocks_per_cu : resident;
      const VmultGrid vg = em_grid(prm.nx, prm.ny, prm.nz, TL, prm.zlayers,
                                   per_cu * device_cu_count());
      for (int c = 0; c < 8; ++c)
        {
          if (vg.grid[c] == 0)
            continue;
          kernel<<<vg.grid[c], kEmThreads, 0, stream>>>(src, dst, prm, c, prm.zlayers, vg.items[c],
                                                        partials, vg.offset[c], vg.total,
                                                        c == vg.last ? 1 : 0, scal);
        }
      return hipGetLastError();
    }

    template <typename Number, int P>
    hipError_t em_deg(const CellParams<Number> &prm, const Number *src, Number *dst,
                      double *partials, SolverScalars *scal, hipStream_t stream)
    {
      switch (prm.tile_log2)
        {
          case 1: return em_passes<Number, P, 1>(prm, src, dst, partials, scal, stream);
          case 2: return em_passes<Number, P, 2>(prm, src, dst, partials, scal, stream);
          case 3: return em_passes<Number, P, 3>(prm, src, dst, partials, scal, stream);
          case 4: return em_passes<Number, P, 4>(prm, src, dst, partials, scal, stream);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  VmultGrid em_grid(int nx, int ny, int nz, int tile_log2, int layers, int max_blocks)
  {
    constexpr int WPB = kEmThreads / 64;
    VmultGrid g{};
    const int TX = 1 << tile_log2, TY = 16 >> tile_log2;
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

  template <typename Number>
  hipError_t launch_vmult_em(int degree, const CellParams<Number> &prm, const Number *src,
                             Number *dst, double *partials, SolverScalars *scal,
                             hipStream_t stream)
  {
    if (prm.em == nullptr || prm.tile_log2 < 1 || prm.tile_log2 > 4)
      return hipErrorInvalidValue;
    switch (degree)
      {
        case 1: return em_deg<Number, 1>(prm, src, dst, partials, scal, stream);
        case 2: return em_deg<Number, 2>(prm, src, dst, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  template hipError_t launch_vmult_em<double>(int, const CellParams<double> &, const double *,
                                              double *, double *, SolverScalars *, hipStream_t);
  template hipError_t launch_vmult_em<float>(int, const CellParams<float> &, const float *, float *,
                                             double *, SolverScalars *, hipStream_t);
} // namespace mpdg
