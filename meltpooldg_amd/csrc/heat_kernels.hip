// CDNA4 cell kernels of the matrix-free heat operator: the vmult and the
// closed-form Jacobi diagonal.
//
// vmult_kernel (wavefront-segment form, see its comment) sweeps x-lines,
// y-lines and z-lines of each cell through a wavefront's own LDS region with
// the 1D contractions done per lane in registers (even-odd split, compile-time
// tables). Shared faces are summed deterministically: x-faces inside a
// workgroup, z-faces inside a work item's z-chunk, and everything else by eight
// ordered colour passes in which a lattice node is stored by the first pass
// that touches it and read-modify-written by the later ones.
//
// diag_kernel is one workgroup per row segment (thread = cell x line) with
// block barriers between the line orientations; it runs once per solve.
//
// Reference behaviour reproduced (source/heat/heat_diffuse_operator.cpp):
//   vmult / tangent_cell_loop :179-241 (gather with homogeneous constraints
//   resolved to 0, distribute_local_to_global dropping constrained rows),
//   tangent_local_cell_operation :955-1086 (evaluate values|gradients,
//   val = dt^-1 rho_cp u, val_grad = k grad u, + rho_cp grad u . v,
//   integrate values|gradients), material lookup :1133-1210,
//   compute_inverse_diagonal_from_matrixfree :271-284.
#include "ctables.hpp"
#include "device_util.hpp"
#include "kernels.hpp"

namespace mpdg
{
  namespace
  {
    template <int P, bool FP32>
    constexpr int cpb()
    {
      return cells_per_block_for(P, FP32);
    }

    template <int P, bool FP32>
    constexpr int block_threads()
    {
      return ((cpb<P, FP32>() * (P + 1) * (P + 1) + 63) / 64) * 64;
    }

    // Minimum wavefronts per SIMD the register allocator must leave room for
    // (HIP's second __launch_bounds__ argument): 4 caps the kernel at 128 VGPRs.
    constexpr int kMinWavesPerSimd = 4;

    // Orders this wavefront's LDS accesses across lanes: LDS instructions of one
    // wavefront execute in issue order, so a compiler barrier is enough.
    __device__ __forceinline__ void wave_sync()
    {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }

    // Orders the LDS phases of a vmult layer: the wavefront's own (its lines
    // cover whole cells) or the workgroup's (spread form).
    template <bool WG>
    __device__ __forceinline__ void phase_sync()
    {
      if constexpr (WG)
        __syncthreads();
      else
        wave_sync();
    }

    // The 1D matrices of the line contractions.
    enum LineMatrix
    {
      kLmS = 0,  // interpolate to Gauss points
      kLmC = 1,  // collocation derivative on Gauss points
      kLmCt = 2, // its transpose (flux back to values)
      kLmSt = 3, // S^T (integrate)
      kLmM = 4,  // reference-cell 1D mass matrix (tensor-product form)
      kLmK = 5   // reference-cell 1D stiffness matrix (tensor-product form)
    };

    template <int P, int M>
    constexpr ct::Mat<P + 1> line_mat()
    {
      return M == kLmS ? ct::CTab<P>::S :
             M == kLmC ? ct::CTab<P>::C :
             M == kLmCt ? ct::CTab<P>::Ct : ct::CTab<P>::St;
    }

    template <int P, int M>
    constexpr ct::EO<P + 1> line_eo()
    {
      return M == kLmS ? ct::VTab<P>::S :
             M == kLmC ? ct::VTab<P>::C :
             M == kLmCt ? ct::VTab<P>::Ct :
             M == kLmSt ? ct::VTab<P>::St :
             M == kLmM ? ct::VTab<P>::Mk : ct::VTab<P>::Kk;
    }

    // MFMA form of a line contraction (FP32, 4 or more points): lane l owns
    // line l. v_mfma_f32_4x4x1_16b treats lanes 4b..4b+3 as block b and forms
    // D[i][j] += A[i] B[j] with A[i] from lane 4b+i and B[j] from lane 4b+j,
    // leaving D[i][j] in register i of lane 4b+j. With A[i] = M[i][k] and
    // B[j] = x_k of lane 4b+j, K = n steps give lane l the outputs y_0..y_3 of
    // its own line; outputs 4.. (n > 4) are left to the vector ALU. col[m][k]
    // holds M_m[lane % 4][k], the A operand of step k of matrix m.
    template <int N, bool MF>
    struct LineCols
    {
      float col[4][MF ? N : 1];
    };

    template <int P, int M>
    __device__ __forceinline__ float line_col(int r, int k)
    {
      constexpr ct::Mat<P + 1> A = line_mat<P, M>();
      return (float)(r == 0 ? A.a[0][k] : r == 1 ? A.a[1][k] : r == 2 ? A.a[2][k] : A.a[3][k]);
    }

    template <int P, bool MF>
    __device__ __forceinline__ LineCols<P + 1, MF> line_cols(int lane)
    {
      LineCols<P + 1, MF> c{};
      if constexpr (MF)
        {
          const int r = lane & 3;
#pragma unroll
          for (int k = 0; k < P + 1; ++k)
            {
              c.col[kLmS][k] = line_col<P, kLmS>(r, k);
              c.col[kLmC][k] = line_col<P, kLmC>(r, k);
              c.col[kLmCt][k] = line_col<P, kLmCt>(r, k);
              c.col[kLmSt][k] = line_col<P, kLmSt>(r, k);
            }
        }
      return c;
    }

    // y = M x along one line: MFMA form (MF) or even-odd vector-ALU form.
    template <int P, int M, bool MF, typename Number>
    __device__ __forceinline__ void line_apply(const LineCols<P + 1, MF> &lc,
                                               const Number (&x)[P + 1], Number (&y)[P + 1])
    {
      constexpr int N = P + 1;
      if constexpr (MF && M <= kLmSt)
        {
          static_assert(sizeof(Number) == 4 && N >= 4, "MFMA line form: FP32, 4+ points");
          typedef float v4f __attribute__((ext_vector_type(4)));
          v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int k = 0; k < N; ++k)
            acc = __builtin_amdgcn_mfma_f32_4x4x1f32(lc.col[M][k], x[k], acc, 0, 0, 0);
          constexpr ct::Mat<N> A = line_mat<P, M>();
#pragma unroll
          for (int q = 4; q < N; ++q)
            {
              Number s = Number(A.a[q][0]) * x[0];
#pragma unroll
              for (int k = 1; k < N; ++k)
                s += Number(A.a[q][k]) * x[k];
              y[q] = s;
            }
#pragma unroll
          for (int q = 0; q < 4; ++q)
            y[q] = acc[q];
        }
      else
        {
          constexpr ct::EO<N> T = line_eo<P, M>();
          eo_apply<N, (M == kLmC || M == kLmCt) ? -1 : 1>(T, x, y);
        }
    }

    // Work item of the vmult: x segment sx (CPB cells), row cy, z chunk kz of
    // up to `layers` cell layers.
    struct Item
    {
      int sx, cy, kz, cx0, ncell, cz0, nl;
    };

    template <typename Number, int CPB>
    __device__ __forceinline__ Item decode_item(const CellParams<Number> &prm, int colour,
                                                int layers, int item)
    {
      const int nsx = (prm.nx + CPB - 1) / CPB;
      const int ox = colour & 1, oy = (colour >> 1) & 1, oz = (colour >> 2) & 1;
      const int msx = (nsx - ox + 1) / 2;
      const int my = (prm.ny - oy + 1) / 2;
      Item it;
      it.sx = ox + 2 * (item % msx);
      const int r = item / msx;
      it.cy = oy + 2 * (r % my);
      it.kz = oz + 2 * (r / my);
      it.cx0 = it.sx * CPB;
      it.ncell = min(CPB, prm.nx - it.cx0);
      it.cz0 = it.kz * layers;
      it.nl = min(layers, prm.nz - it.cz0);
      return it;
    }

    // Ordered schedule: ticket t of a launch is item t - off[c] of colour c,
    // off[c] <= t < off[c + 1] (colours in pass order).
    struct OrderGrid
    {
      int off[9];
    };

    // og.off[c] for a run-time c, as a select chain over the (scalar) kernel
    // arguments: a dynamic index would copy the array to private memory
    __device__ __forceinline__ int order_off(const OrderGrid &og, int c)
    {
      int r = 0;
#pragma unroll
      for (int k = 0; k < 9; ++k)
        r = k == c ? og.off[k] : r;
      return r;
    }

    template <typename Number, int CPB>
    __device__ __forceinline__ Item decode_ticket(const CellParams<Number> &prm,
                                                  const OrderGrid &og, int layers, int t)
    {
      int c = 0;
#pragma unroll
      for (int k = 1; k < 8; ++k)
        c += t >= og.off[k] ? 1 : 0;
      return decode_item<Number, CPB>(prm, c, layers, t - order_off(og, c));
    }

    // Ordered schedule, cross-workgroup visibility of dst: the read-modify-write
    // loads and the stores of dst use the agent-coherent cache policy, and a
    // workgroup publishes an item only after its stores have completed
    // (vmcnt(0)), so neighbours on other XCDs read the sums without whole-cache
    // write-back / invalidate fences (which cost more than the whole vmult).

    template <bool C, typename Number>
    __device__ __forceinline__ Number dst_load(__amdgpu_buffer_rsrc_t r, uint32_t off)
    {
      if constexpr (C)
        return buf_load_coherent(r, off, Number());
      else
        return buf_load(r, off, Number());
    }
    template <bool C, typename Number>
    __device__ __forceinline__ void dst_store(__amdgpu_buffer_rsrc_t r, uint32_t off, Number v)
    {
      if constexpr (C)
        buf_store_coherent(r, off, v);
      else
        buf_store(r, off, v);
    }

    // Next ticket of this workgroup. Q = min(gridDim, kOrderQueues) queues, one
    // per XCD (workgroups are dispatched to XCDs round-robin): workgroup b takes
    // the tickets q, q + Q, q + 2Q, ... of queue q = b mod Q in increasing order,
    // each queue counter in its own cache line so the XCDs do not contend on one
    // atomic. Workgroups 0..Q-1 are dispatched first and leave only after their
    // queue is exhausted, so every queue is served while it has tickets.
    __device__ __forceinline__ int order_grab(OrderState *os)
    {
      const int Q = min((int)gridDim.x, kOrderQueues);
      const int q = (int)blockIdx.x % Q;
      const unsigned k = atomicAdd(&os->ticket[q][0], 1u);
      return (int)k * Q + q;
    }

    // Bound of one wait for a neighbouring item (2^20 polls of a few hundred
    // cycles each: about a second, far above any item's run time, so only a
    // defect can reach it; the wait then gives up and sets order->error).
    constexpr unsigned kOrderSpinLimit = 1u << 20;

    // Ordered schedule: before item `it` reads back (read-modify-write) any node
    // it shares with its neighbours, every neighbouring item of an earlier
    // colour must have stored. Items touch nodes of their 26 neighbours in the
    // (segment, row, z-chunk) grid; a neighbour's colour is its parity, and its
    // ticket inverts decode_item. Lanes 0..26 of wave 0 poll one neighbour
    // each; the workgroup then synchronises, and its coherent dst loads see
    // the neighbours' completed coherent stores.
    // Every awaited ticket is lower than this one, and the lowest unfinished
    // ticket never waits: its queue serves tickets in order, so it is held or
    // next to be taken by a running workgroup. Hence the waits end.
    template <typename Number, int CPB>
    __device__ __forceinline__ void order_wait(const CellParams<Number> &prm, const OrderGrid &og,
                                               const Item &it, int layers, unsigned epoch)
    {
      const int k = threadIdx.x;
      if (k < 27 && k != 13)
        {
          const int nsx = (prm.nx + CPB - 1) / CPB;
          const int nkz = (prm.nz + layers - 1) / layers;
          const int x = it.sx + k % 3 - 1, y = it.cy + (k / 3) % 3 - 1, z = it.kz + k / 9 - 1;
          const int c = (it.sx & 1) | ((it.cy & 1) << 1) | ((it.kz & 1) << 2);
          const int cn = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
          if (x >= 0 && x < nsx && y >= 0 && y < prm.ny && z >= 0 && z < nkz && cn < c)
            {
              const int msx = (nsx - (cn & 1) + 1) / 2;
              const int my = (prm.ny - ((cn >> 1) & 1) + 1) / 2;
              const int t = order_off(og, cn) + (x >> 1) + msx * ((y >> 1) + my * (z >> 1));
              unsigned spins = 0;
              while (__hip_atomic_load(prm.order_flags + t, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT) != epoch)
                {
                  __builtin_amdgcn_s_sleep(2);
                  if (++spins > kOrderSpinLimit)
                    {
                      __hip_atomic_fetch_or(&prm.order->error, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
                      break;
                    }
                }
            }
        }
      __syncthreads();
    }

    // Gather of one x-line (y = s0, z = s1 of cell layer cz, block cell cb) with
    // homogeneous constraints read as 0 (FEEvaluation::read_dof_values).
    template <int P, typename Number>
    __device__ __forceinline__ void gather_xline(const CellParams<Number> &prm,
                                                 __amdgpu_buffer_rsrc_t src, const Item &it,
                                                 int cz, int cb, int s0, int s1, bool ok,
                                                 Number (&u)[P + 1])
    {
      constexpr int N = P + 1;
      constexpr uint32_t SZ = sizeof(Number);
      const int Yi = it.cy * P + s0, Zi = cz * P + s1;
      const int X0 = (it.cx0 + cb) * P;
      const bool take = ok & (cb < it.ncell) & !dirichlet_node(prm, Yi, prm.n1y, 1) &
                        !dirichlet_node(prm, Zi, prm.n1z, 2);
      // one masked offset per line; only the end nodes can sit on an x face
      const uint32_t lb = take ? (uint32_t)(X0 + prm.n1x * (Yi + prm.n1y * Zi)) * SZ : kOutOfRange;
      const bool x0d = (prm.faces & 1u) & (X0 == 0);
      const bool x1d = ((prm.faces >> 1) & 1u) & (X0 + P == prm.n1x - 1);
      u[0] = buf_load(src, x0d ? kOutOfRange : lb, Number());
#pragma unroll
      for (int a = 1; a < N - 1; ++a)
        u[a] = buf_load(src, lb + a * SZ, Number());
      u[N - 1] = buf_load(src, x1d ? kOutOfRange : lb + (N - 1) * SZ, Number());
    }

    template <int P, typename Number>
    __device__ __forceinline__ void gather_plain(const CellParams<Number> &prm,
                                                 __amdgpu_buffer_rsrc_t v, const Item &it,
                                                 int cz, int cb, int s0, int s1, bool ok,
                                                 Number (&u)[P + 1])
    {
      constexpr int N = P + 1;
      const uint32_t base =
        (uint32_t)((it.cx0 + cb) * P + prm.n1x * (it.cy * P + s0 + prm.n1y * (cz * P + s1))) *
        sizeof(Number);
      const bool take = ok & (cb < it.ncell);
#pragma unroll
      for (int a = 0; a < N; ++a)
        u[a] = buf_load(v, take ? base + a * (uint32_t)sizeof(Number) : kOutOfRange, Number());
    }

    // Per-lane constants of one line slot: offset of the line's first node from
    // the layer's base node, and where the line sits in its cell.
    enum LaneBit : unsigned
    {
      kS0Lo = 1u,  // s0 == 0   (y face of the cell, or qx == 0)
      kS0Hi = 2u,  // s0 == p
      kS1Lo = 4u,  // s1 == 0   (z face)
      kS1Hi = 8u,  // s1 == p
      kCb0 = 16u   // first cell of the segment
    };

    struct LaneConst
    {
      int loff;      // cb*p + n1x*(s0 + n1y*s1)
      unsigned bits; // LaneBit flags
    };

    // Uniform facts of one (work item, cell layer): masks over LaneBit.
    struct LayerU
    {
      int base;           // lattice index of the layer's first node (cx0*p, cy*p, cz*p)
      unsigned rowc;      // lines on a Dirichlet y/z face
      unsigned notfirst;  // lines shared in y/z with an earlier colour pass
      unsigned defer;     // top-face lines carried to the next layer
      bool x0d, x1d;      // segment's first / last cell touches a Dirichlet x face
      bool x0nf, xPnf;    // segment's first / last x end shared with an earlier pass
    };

    template <typename Number, int P>
    __device__ __forceinline__ LayerU layer_uniforms(const CellParams<Number> &prm, const Item &it,
                                                     int j)
    {
      const int cz = it.cz0 + j;
      const unsigned f = prm.faces;
      LayerU u;
      u.base = it.cx0 * P + prm.n1x * (it.cy * P + prm.n1y * (cz * P));
      u.rowc = (((f >> 2) & 1u) && it.cy == 0 ? kS0Lo : 0u) |
               (((f >> 3) & 1u) && it.cy == prm.ny - 1 ? kS0Hi : 0u) |
               (((f >> 4) & 1u) && cz == 0 ? kS1Lo : 0u) |
               (((f >> 5) & 1u) && cz == prm.nz - 1 ? kS1Hi : 0u);
      const bool cy_odd = it.cy & 1, kz_odd = it.kz & 1;
      u.notfirst = (cy_odd && it.cy > 0 ? kS0Lo : 0u) | (cy_odd && it.cy < prm.ny - 1 ? kS0Hi : 0u) |
                   (kz_odd && j == 0 && it.cz0 > 0 ? kS1Lo : 0u) |
                   (kz_odd && j == it.nl - 1 && it.cz0 + it.nl < prm.nz ? kS1Hi : 0u);
      u.defer = j < it.nl - 1 ? kS1Hi : 0u;
      u.x0d = (f & 1u) && it.cx0 == 0;
      u.x1d = ((f >> 1) & 1u) && it.cx0 + it.ncell == prm.nx;
      const bool sx_odd = it.sx & 1;
      u.x0nf = sx_odd && it.cx0 > 0;
      u.xPnf = sx_odd && it.cx0 + it.ncell < prm.nx;
      return u;
    }

    // Gather of one x-line with homogeneous constraints read as 0
    // (FEEvaluation::read_dof_values); only the end nodes can sit on an x face.
    template <int P, typename Number>
    __device__ __forceinline__ void gather_line(__amdgpu_buffer_rsrc_t src, const LaneConst &lc,
                                                const LayerU &U, int cb, int ncell, bool ok,
                                                Number (&u)[P + 1])
    {
      constexpr int N = P + 1;
      constexpr uint32_t SZ = sizeof(Number);
      const bool take = ok & (cb < ncell) & ((lc.bits & U.rowc) == 0u);
      const uint32_t lb = take ? (uint32_t)(U.base + lc.loff) * SZ : kOutOfRange;
      const bool x0 = U.x0d & ((lc.bits & kCb0) != 0u);
      const bool x1 = U.x1d & (cb == ncell - 1);
      u[0] = buf_load(src, x0 ? kOutOfRange : lb, Number());
#pragma unroll
      for (int a = 1; a < N - 1; ++a)
        u[a] = buf_load(src, lb + a * SZ, Number());
      u[N - 1] = buf_load(src, x1 ? kOutOfRange : lb + (N - 1) * SZ, Number());
    }

    template <int P, typename Number>
    __device__ __forceinline__ void gather_line_plain(__amdgpu_buffer_rsrc_t src,
                                                      const LaneConst &lc, const LayerU &U, int cb,
                                                      int ncell, bool ok, Number (&u)[P + 1])
    {
      constexpr int N = P + 1;
      constexpr uint32_t SZ = sizeof(Number);
      const uint32_t lb = (ok & (cb < ncell)) ? (uint32_t)(U.base + lc.loff) * SZ : kOutOfRange;
#pragma unroll
      for (int a = 0; a < N; ++a)
        u[a] = buf_load(src, lb + a * SZ, Number());
    }

    // Cell-loop vmult, wavefront-segment form.
    //
    // A workgroup of WPB wavefronts owns a row segment of CPB = WPB*CW cells
    // along x; wavefront w owns cells [w*CW, (w+1)*CW) of it. The CW*N^2 lines
    // of those cells are spread over the 64 lanes x R slots of the wavefront,
    // cell index fastest. Lanes beyond the last line repeat a real line: they
    // load and compute exactly what its owner does (so their LDS writes carry
    // the same values) and are masked off global stores and the dot product. Every 1D contraction is one lane working on a
    // whole line in registers; moving between x-, y- and z-lines goes through
    // the wavefront's own LDS region, ordered by wave_sync() only. The single
    // workgroup barrier per cell layer is the exchange of shared x-faces between
    // neighbouring wavefronts.
    //
    // Each workgroup loops over the work items of one colour and over the cell
    // layers of each item (z-faces between layers carried through LDS, so only
    // the chunk's outer z-faces take part in the colouring). The next layer's
    // x-lines and the current layer's read-modify-write values are loaded into
    // registers early so their latency overlaps the sum-factorisation work.
    // Body of the vmult: this workgroup processes items item_first,
    // item_first + item_stride, ... (< items) of one colour of one operator and
    // returns this thread's share of p . Ap (sum over its cells of u . y).
    //
    // Ordered schedule (ORD): one launch covers all colours. `item` is then the
    // launch's ticket (items of colour 0 first), taken from prm.order->ticket
    // one item ahead; each item waits for its earlier-colour neighbours before
    // its read-modify-write loads and, after its stores, publishes its flag
    // (and, with ipart, its p . Ap share in ipart[ticket]).
    template <int P, typename Number, int COEF, bool VEL, int CW, int R, int WPB, bool MF,
              bool ORD = false>
    __device__ __forceinline__ double vmult_body(const Number *__restrict__ src,
                                                 Number *__restrict__ dst,
                                                 const Number *__restrict__ heav,
                                                 const CellParams<Number> &prm, const int colour,
                                                 const int layers, const int items,
                                                 const int item_first, const int item_stride,
                                                 const OrderGrid og = OrderGrid{},
                                                 double *__restrict__ ipart = nullptr,
                                                 unsigned epoch = 0)
    {
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      // R = 0: spread form, CW = cells of the workgroup, one line per thread
      constexpr bool SPREAD = R == 0;
      constexpr int RR = SPREAD ? 1 : R;
      constexpr int CPB = SPREAD ? CW : CW * WPB;
      constexpr int LT = SPREAD ? CPB * NL : CW * NL;
      constexpr int THREADS = 64 * WPB;
      constexpr bool H2 = COEF == kCoefHeaviside;
      constexpr bool SINE = COEF == kCoefSine;
      // Constant coefficients and no velocity: the cell matrix is
      //   mass Mh(x)Mh(x)Mh + k sum_d kf_d (Kh in direction d, Mh in the others)
      // (Mh, Kh: reference 1D mass and stiffness on the Gauss rule, which is
      // exact for them), applied as 7 nodal line contractions in 4 line phases
      // instead of 12 contractions and the quadrature-point work in 9.
      constexpr bool KRON = COEF == kCoefConstant && !VEL && !MF;
      constexpr bool CD = ORD; // agent-coherent dst accesses
      static_assert(SPREAD ? (LT <= THREADS && LT > THREADS - NL)
                           : (LT <= 64 * R && LT > 64 * (R - 1)),
                    "lines must fill the last lane slot");
      constexpr ct::Tab<N> T1 = ct::VTab<P>::t;

      __shared__ Number A[CPB * STR];
      __shared__ Number B[CPB * STR];
      __shared__ Number KB[H2 ? CPB * STR : 1];
      __shared__ Number RB[(H2 && VEL) ? CPB * STR : 1];
      __shared__ Number FZ[2][CPB * NL]; // z-face carry [cell][b][a], by layer parity
      __shared__ Number XE[2][CPB * NL]; // right x-face of each cell line [cell][line]
      __shared__ Number KW[SINE ? CPB * STR : 1]; // k w3 at the q-points [cell][qz][qy][qx]

      const uint32_t vbytes = vector_bytes(prm.n1x, prm.n1y, prm.n1z, (int)sizeof(Number));
      const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(src, vbytes);
      const __amdgpu_buffer_rsrc_t rdst = make_rsrc(dst, vbytes);
      const __amdgpu_buffer_rsrc_t rheav = make_rsrc(H2 ? heav : src, vbytes);
      const int w = threadIdx.x >> 6;
      const int lane = threadIdx.x & 63;
      const LineCols<N, MF> LC = line_cols<P, MF>(lane);
      int cb[RR], s0v[RR], s1v[RR], sv[RR];
      bool lv[RR];
#pragma unroll
      for (int r = 0; r < RR; ++r)
        {
          int g, l;
          if constexpr (SPREAD)
            {
              // thread t owns line t / CPB of cell t % CPB; the few spare
              // threads (last wavefront) repeat lines of that wavefront
              const int g0 = threadIdx.x;
              lv[r] = g0 < LT;
              g = lv[r] ? g0 : g0 - (THREADS - LT);
              l = g % CPB;
              sv[r] = g / CPB;
              cb[r] = l;
            }
          else
            {
              // a spare lane (only in the last slot) repeats a line of that same slot,
              // so both run in one instruction and see the same LDS contents
              const int g0 = lane + 64 * r;
              lv[r] = g0 < LT;
              g = lv[r] ? g0 : 64 * (R - 1) + (g0 - LT) % (LT - 64 * (R - 1));
              l = g % CW;
              sv[r] = g / CW;
              cb[r] = w * CW + l;
            }
          s0v[r] = sv[r] % N;
          s1v[r] = sv[r] / N;
        }
      LaneConst lc[RR];
#pragma unroll
      for (int r = 0; r < RR; ++r)
        {
          lc[r].loff = cb[r] * P + prm.n1x * (s0v[r] + prm.n1y * s1v[r]);
          lc[r].bits = (s0v[r] == 0 ? kS0Lo : 0u) | (s0v[r] == P ? kS0Hi : 0u) |
                       (s1v[r] == 0 ? kS1Lo : 0u) | (s1v[r] == P ? kS1Hi : 0u) |
                       (cb[r] == 0 ? kCb0 : 0u);
        }

      __shared__ int s_ticket[1];
      __shared__ double s_red[ORD ? WPB : 1];
      int item = item_first;
      if constexpr (ORD)
        {
          if (threadIdx.x == 0)
            s_ticket[0] = order_grab(prm.order);
          __syncthreads();
          // uniform (scalar) from here on, like the passes' item index
          item = __builtin_amdgcn_readfirstlane(s_ticket[0]);
        }
      int j = 0;
      Item it{};
      if constexpr (ORD)
        {
          if (item < items)
            it = decode_ticket<Number, CPB>(prm, og, layers, item);
        }
      else
        it = decode_item<Number, CPB>(prm, colour, layers, item);
      LayerU LU = layer_uniforms<Number, P>(prm, it, 0);
      Number u[RR][N], hl[RR][N];
#pragma unroll
      for (int r = 0; r < RR; ++r)
        {
          gather_line<P>(rsrc, lc[r], LU, cb[r], it.ncell, item < items, u[r]);
          if constexpr (H2)
            gather_line_plain<P>(rheav, lc[r], LU, cb[r], it.ncell, item < items, hl[r]);
        }
      double acc = 0.0;
      int step = 0; // layers done by this workgroup; selects the x-exchange buffer

      while (item < items)
        {
          const int cz = it.cz0 + j;
          // separable-sine conductivity: sin(2 pi x) sin(2 pi y) at the lane's
          // z-line of the z-line phase (qx = s0, qy = s1), loaded early
          Number ssx[RR], ssy[RR];
          if constexpr (SINE)
#pragma unroll
            for (int r = 0; r < RR; ++r)
              {
                const int cx = min(it.cx0 + cb[r], prm.nx - 1);
                ssx[r] = prm.sin2pi[0][cx * N + s0v[r]];
                ssy[r] = prm.sin2pi[1][it.cy * N + s1v[r]];
              }
          int item_n = item, j_n = j + 1;
          Item it_n = it;
          if (j_n >= it.nl)
            {
              j_n = 0;
              if constexpr (ORD)
                {
                  if (threadIdx.x == 0)
                    s_ticket[0] = order_grab(prm.order);
                  __syncthreads();
                  item_n = __builtin_amdgcn_readfirstlane(s_ticket[0]);
                  if (item_n < items)
                    it_n = decode_ticket<Number, CPB>(prm, og, layers, item_n);
                }
              else
                {
                  item_n = item + item_stride;
                  if (item_n < items)
                    it_n = decode_item<Number, CPB>(prm, colour, layers, item_n);
                }
            }

          // ---- x-lines (y = s0, z = s1): interpolate in x ------------------------
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              Number *Al = A + cb[r] * STR;
              Number t1[N];
              if constexpr (KRON)
                {
                  // Mh u into A, Kh u into B (x-lines)
                  Number *Bl = B + cb[r] * STR;
                  Number t2[N];
                  line_apply<P, kLmM, MF>(LC, u[r], t1);
                  line_apply<P, kLmK, MF>(LC, u[r], t2);
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    {
                      Al[(s1v[r] * N + s0v[r]) * N + q] = t1[q];
                      Bl[(s1v[r] * N + s0v[r]) * N + q] = t2[q];
                    }
                }
              else
                {
                  line_apply<P, kLmS, MF>(LC, u[r], t1);
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    Al[(s1v[r] * N + s0v[r]) * N + q] = t1[q];
                }
              if constexpr (H2)
                {
                  Number *KBl = KB + cb[r] * STR;
                  line_apply<P, kLmS, MF>(LC, hl[r], t1);
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    KBl[(s1v[r] * N + s0v[r]) * N + q] = t1[q];
                }
            }

          // prefetch of the next layer's x-lines
          const LayerU LUn = layer_uniforms<Number, P>(prm, it_n, j_n);
          Number un[RR][N], hn[RR][N];
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              gather_line<P>(rsrc, lc[r], LUn, cb[r], it_n.ncell, item_n < items, un[r]);
              if constexpr (H2)
                gather_line_plain<P>(rheav, lc[r], LUn, cb[r], it_n.ncell, item_n < items, hn[r]);
            }

          if constexpr (ORD)
            if (j == 0)
              order_wait<Number, CPB>(prm, og, it, layers, epoch);

          // scatter plan of each output x-line (b = s0, c = s1). Interior nodes
          // (0 < a < p) share the line's mode; only the end nodes can be on an x
          // face or shared with the neighbouring segment. Offsets are out of
          // range where nothing is written / read; nodes stored first get 0
          // from the early read-modify-write load.
          uint32_t wl[RR], w0o[RR], wPo[RR];
          Number dv[RR][N];
          bool zl[RR], z0[RR], zP[RR];
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              constexpr uint32_t SZ = sizeof(Number);
              const uint32_t obase = (uint32_t)(LU.base + lc[r].loff) * SZ;
              const unsigned bits = lc[r].bits;
              const bool last_cell = cb[r] == it.ncell - 1;
              const bool lw = lv[r] & (cb[r] < it.ncell) & ((bits & LU.defer) == 0u);
              const bool rowc = (bits & LU.rowc) != 0u;
              const bool yz_first = (bits & LU.notfirst) == 0u;
              // interior nodes
              wl[r] = (lw & (!rowc | yz_first)) ? obase : kOutOfRange;
              zl[r] = rowc;
              const uint32_t li = (lw & !yz_first & !rowc) ? obase : kOutOfRange;
#pragma unroll
              for (int a = 1; a < N - 1; ++a)
                dv[r][a] = dst_load<CD, Number>(rdst, li + a * SZ);
              // node a = 0
              const bool cb0 = (bits & kCb0) != 0u;
              const bool first0 = yz_first & !(cb0 & LU.x0nf);
              const bool con0 = rowc | (cb0 & LU.x0d);
              const bool wr0 = lw & (!con0 | first0);
              w0o[r] = wr0 ? obase : kOutOfRange;
              z0[r] = con0;
              dv[r][0] = dst_load<CD, Number>(rdst, (wr0 & !first0 & !con0) ? obase : kOutOfRange);
              // node a = p (written by the segment's last cell only)
              const bool firstP = yz_first & !(last_cell & LU.xPnf);
              const bool conP = rowc | (last_cell & LU.x1d);
              const bool wrP = lw & last_cell & (!conP | firstP);
              wPo[r] = wrP ? obase + P * SZ : kOutOfRange;
              zP[r] = conP;
              dv[r][N - 1] = dst_load<CD, Number>(rdst, (wrP & !firstP & !conP) ? obase + P * SZ : kOutOfRange);
            }
          phase_sync<SPREAD>();

          // ---- y-lines (qx = s0, z = s1): interpolate in y ------------------------
          if constexpr (KRON)
            {
              // y-lines (x = s0, z = s1): A <- mass Mh t + k kf_x Mh s + k kf_y Kh t,
              // B <- Mh t (t = Mh_x u, s = Kh_x u)
              const Number c0 = prm.k0 * prm.kf[0], c1 = prm.k0 * prm.kf[1];
#pragma unroll
              for (int r = 0; r < RR; ++r)
                {
                  Number *Al = A + cb[r] * STR;
                  Number *Bl = B + cb[r] * STR;
                  Number tv[N], sv2[N], mt[N], ms[N], kt[N];
#pragma unroll
                  for (int b = 0; b < N; ++b)
                    {
                      tv[b] = Al[(s1v[r] * N + b) * N + s0v[r]];
                      sv2[b] = Bl[(s1v[r] * N + b) * N + s0v[r]];
                    }
                  line_apply<P, kLmM, MF>(LC, tv, mt);
                  line_apply<P, kLmM, MF>(LC, sv2, ms);
                  line_apply<P, kLmK, MF>(LC, tv, kt);
#pragma unroll
                  for (int b = 0; b < N; ++b)
                    {
                      Al[(s1v[r] * N + b) * N + s0v[r]] = prm.mass * mt[b] + c0 * ms[b] + c1 * kt[b];
                      Bl[(s1v[r] * N + b) * N + s0v[r]] = mt[b];
                    }
                }
            }
          else
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              Number *Al = A + cb[r] * STR;
              Number v[N], o[N];
#pragma unroll
              for (int b = 0; b < N; ++b)
                v[b] = Al[(s1v[r] * N + b) * N + s0v[r]];
              line_apply<P, kLmS, MF>(LC, v, o);
#pragma unroll
              for (int b = 0; b < N; ++b)
                Al[(s1v[r] * N + b) * N + s0v[r]] = o[b];
              if constexpr (H2)
                {
                  Number *KBl = KB + cb[r] * STR;
#pragma unroll
                  for (int b = 0; b < N; ++b)
                    v[b] = KBl[(s1v[r] * N + b) * N + s0v[r]];
                  line_apply<P, kLmS, MF>(LC, v, o);
#pragma unroll
                  for (int b = 0; b < N; ++b)
                    KBl[(s1v[r] * N + b) * N + s0v[r]] = o[b];
                }
            }
          phase_sync<SPREAD>();

          // ---- z-lines (qx = s0, qy = s1): values, z-gradient and z-flux ------
          // The z-line's own part of the q-point residual is completed here,
          //   Vz = rho_cp/dt w3 U + C^T (k w3 kf_z dU/dz) (+ rho_cp w3 u_z dU/dz),
          // so only Vz stays in registers until the x- and y-fluxes are added.
          Number Vz[RR][N];
          if constexpr (KRON)
            {
              // z-lines (x = s0, y = s1): A <- Mh_z A + k kf_z Kh_z B, the cell's result
              const Number c2 = prm.k0 * prm.kf[2];
#pragma unroll
              for (int r = 0; r < RR; ++r)
                {
                  Number *Al = A + cb[r] * STR;
                  const Number *Bl = B + cb[r] * STR;
                  Number p1[N], p2[N], m1[N], k2[N];
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    {
                      p1[c] = Al[(c * N + s1v[r]) * N + s0v[r]];
                      p2[c] = Bl[(c * N + s1v[r]) * N + s0v[r]];
                    }
                  line_apply<P, kLmM, MF>(LC, p1, m1);
                  line_apply<P, kLmK, MF>(LC, p2, k2);
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    Al[(c * N + s1v[r]) * N + s0v[r]] = m1[c] + c2 * k2[c];
                }
            }
          else
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const int s0 = s0v[r], s1 = s1v[r];
              Number *Al = A + cb[r] * STR;
              Number v[N], U[N], Gz[N], kq[N], rcq[N], f[N], o[N];
#pragma unroll
              for (int c = 0; c < N; ++c)
                v[c] = Al[(c * N + s1) * N + s0];
              line_apply<P, kLmS, MF>(LC, v, U);
              line_apply<P, kLmC, MF>(LC, U, Gz);
#pragma unroll
              for (int q = 0; q < N; ++q)
                Al[(q * N + s1) * N + s0] = U[q];
              if constexpr (H2)
                {
                  Number *KBl = KB + cb[r] * STR;
                  Number *RBl = RB + ((H2 && VEL) ? cb[r] * STR : 0);
                  Number Hq[N];
#pragma unroll
                  for (int c = 0; c < N; ++c)
                    v[c] = KBl[(c * N + s1) * N + s0];
                  line_apply<P, kLmS, MF>(LC, v, Hq);
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    {
                      two_phase(prm, Hq[q], rcq[q], kq[q]);
                      KBl[(q * N + s1) * N + s0] = kq[q];
                      if constexpr (VEL)
                        RBl[(q * N + s1) * N + s0] = rcq[q];
                    }
                }
              else
                {
                  const Number c = SINE ? prm.kamp * ssx[r] * ssy[r] : Number(0);
#pragma unroll
                  for (int q = 0; q < N; ++q)
                    {
                      kq[q] = SINE ? prm.k0 + c * prm.sin2pi[2][cz * N + q] : prm.k0;
                      rcq[q] = prm.rho_cp0;
                    }
                }
              const Number wxy = (Number)T1.w[s0] * (Number)T1.w[s1];
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  const Number kw = kq[q] * ((Number)T1.w[q] * wxy);
                  if constexpr (SINE)
                    KW[cb[r] * STR + (q * N + s1) * N + s0] = kw;
                  f[q] = kw * prm.kf[2] * Gz[q];
                }
              line_apply<P, kLmCt, MF>(LC, f, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  const Number w3 = (Number)T1.w[q] * wxy;
                  const Number mass = H2 ? rcq[q] * prm.dtinv_vol : prm.mass;
                  Number val = mass * w3 * U[q] + o[q];
                  if constexpr (VEL)
                    val += rcq[q] * w3 * prm.vf[2] * prm.w0 * Gz[q];
                  Vz[r][q] = val;
                }
            }
          phase_sync<SPREAD>();

          if constexpr (!KRON)
            {
          // ---- x-lines (qy = s0, qz = s1): x-gradient flux --------------------
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const int s0 = s0v[r], s1 = s1v[r];
              const int cx = min(it.cx0 + cb[r], prm.nx - 1);
              Number *Al = A + cb[r] * STR;
              Number *Bl = B + cb[r] * STR;
              const Number *KBl = KB + (H2 ? cb[r] * STR : 0);
              const Number *RBl = RB + ((H2 && VEL) ? cb[r] * STR : 0);
              Number v[N], gx[N], f[N], o[N], cv[N];
#pragma unroll
              for (int q = 0; q < N; ++q)
                v[q] = Al[(s1 * N + s0) * N + q];
              line_apply<P, kLmC, MF>(LC, v, gx);
              const Number wyz = (Number)T1.w[s0] * (Number)T1.w[s1];
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  const Number w3 = (Number)T1.w[q] * wyz;
                  const Number kw = SINE ? KW[cb[r] * STR + (s1 * N + s0) * N + q]
                                         : conductivity<Number, COEF, N>(prm, KBl, cx, it.cy, cz, q, s0, s1) * w3;
                  f[q] = kw * prm.kf[0] * gx[q];
                  if constexpr (VEL)
                    {
                      const Number rc = H2 ? RBl[(s1 * N + s0) * N + q] : prm.rho_cp0;
                      const Number vx = prm.vel_x[0][cx * N + q] * prm.vel_y[0][it.cy * N + s0];
                      cv[q] = rc * w3 * prm.vf[0] * vx * gx[q];
                    }
                }
              line_apply<P, kLmCt, MF>(LC, f, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                Bl[(s1 * N + s0) * N + q] = VEL ? o[q] + cv[q] : o[q];
            }
          phase_sync<SPREAD>();

          // ---- y-lines (qx = s0, qz = s1): y-gradient flux ----------------------
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const int s0 = s0v[r], s1 = s1v[r];
              const int cx = min(it.cx0 + cb[r], prm.nx - 1);
              Number *Al = A + cb[r] * STR;
              Number *Bl = B + cb[r] * STR;
              const Number *KBl = KB + (H2 ? cb[r] * STR : 0);
              const Number *RBl = RB + ((H2 && VEL) ? cb[r] * STR : 0);
              Number v[N], gy[N], f[N], o[N], cv[N];
#pragma unroll
              for (int q = 0; q < N; ++q)
                v[q] = Al[(s1 * N + q) * N + s0];
              line_apply<P, kLmC, MF>(LC, v, gy);
              const Number wxz = (Number)T1.w[s0] * (Number)T1.w[s1];
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  const Number w3 = (Number)T1.w[q] * wxz;
                  const Number kw = SINE ? KW[cb[r] * STR + (s1 * N + q) * N + s0]
                                         : conductivity<Number, COEF, N>(prm, KBl, cx, it.cy, cz, s0, q, s1) * w3;
                  f[q] = kw * prm.kf[1] * gy[q];
                  if constexpr (VEL)
                    {
                      const Number rc = H2 ? RBl[(s1 * N + q) * N + s0] : prm.rho_cp0;
                      const Number vy = prm.vel_x[1][cx * N + s0] * prm.vel_y[1][it.cy * N + q];
                      cv[q] = rc * w3 * prm.vf[1] * vy * gy[q];
                    }
                }
              line_apply<P, kLmCt, MF>(LC, f, o);
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  const int idx = (s1 * N + q) * N + s0;
                  Bl[idx] = Bl[idx] + (VEL ? o[q] + cv[q] : o[q]);
                }
            }
          phase_sync<SPREAD>();

          // ---- z-lines (qx = s0, qy = s1): add the x- and y-fluxes, integrate z
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const int s0 = s0v[r], s1 = s1v[r];
              Number *Al = A + cb[r] * STR;
              const Number *Bl = B + cb[r] * STR;
              Number V[N], y1[N];
#pragma unroll
              for (int q = 0; q < N; ++q)
                V[q] = Vz[r][q] + Bl[(q * N + s1) * N + s0];
              line_apply<P, kLmSt, MF>(LC, V, y1);
#pragma unroll
              for (int c = 0; c < N; ++c)
                Al[(c * N + s1) * N + s0] = y1[c];
            }
          phase_sync<SPREAD>();

          // ---- y-lines (qx = s0, c = s1): integrate y ------------------------------
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              Number *Al = A + cb[r] * STR;
              Number v[N], o[N];
#pragma unroll
              for (int q = 0; q < N; ++q)
                v[q] = Al[(s1v[r] * N + q) * N + s0v[r]];
              line_apply<P, kLmSt, MF>(LC, v, o);
#pragma unroll
              for (int b = 0; b < N; ++b)
                Al[(s1v[r] * N + b) * N + s0v[r]] = o[b];
            }
          phase_sync<SPREAD>();
            }

          // ---- x-lines (b = s0, c = s1): integrate x, publish right faces ---------
          Number Y[RR][N];
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const Number *Al = A + cb[r] * STR;
              Number v[N];
#pragma unroll
              for (int q = 0; q < N; ++q)
                v[q] = Al[(s1v[r] * N + s0v[r]) * N + q];
              if constexpr (KRON)
                {
#pragma unroll
                  for (int a = 0; a < N; ++a)
                    Y[r][a] = v[a];
                }
              else
                line_apply<P, kLmSt, MF>(LC, v, Y[r]);
              if (lv[r])
#pragma unroll
                for (int a = 0; a < N; ++a)
                  acc += (double)u[r][a] * (double)Y[r][a];
              XE[step & 1][cb[r] * NL + sv[r]] = Y[r][N - 1];
            }
          __syncthreads();

          // ---- sum shared x- and z-faces, scatter -----------------------------------
#pragma unroll
          for (int r = 0; r < RR; ++r)
            {
              const int b = s0v[r], c = s1v[r];
              if (cb[r] > 0)
                Y[r][0] += XE[step & 1][(cb[r] - 1) * NL + sv[r]];
              if (c == 0 && j > 0)
#pragma unroll
                for (int a = 0; a < N; ++a)
                  Y[r][a] += FZ[(j - 1) & 1][(cb[r] * N + b) * N + a];
              if (c == P && j < it.nl - 1)
#pragma unroll
                for (int a = 0; a < N; ++a)
                  FZ[j & 1][(cb[r] * N + b) * N + a] = Y[r][a];
              dst_store<CD>(rdst, w0o[r], z0[r] ? Number(0) : dv[r][0] + Y[r][0]);
#pragma unroll
              for (int a = 1; a < N - 1; ++a)
                dst_store<CD>(rdst, wl[r] + a * (uint32_t)sizeof(Number), zl[r] ? Number(0) : dv[r][a] + Y[r][a]);
              dst_store<CD>(rdst, wPo[r], zP[r] ? Number(0) : dv[r][N - 1] + Y[r][N - 1]);
            }

          if constexpr (ORD)
            if (j == it.nl - 1)
              {
                // the item's stores reach agent scope before its flag does
                wait_vmem();
                double bs = 0.0;
                if (ipart != nullptr)
                  bs = block_sum<THREADS>(acc, s_red);
                else
                  __syncthreads();
                acc = 0.0;
                if (threadIdx.x == 0)
                  {
                    if (ipart != nullptr)
                      __hip_atomic_store(ipart + item, bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    // every wave of the workgroup released its stores before the barrier
                    __hip_atomic_store(prm.order_flags + item, epoch, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                  }
              }

#pragma unroll
          for (int r = 0; r < RR; ++r)
#pragma unroll
            for (int a = 0; a < N; ++a)
              {
                u[r][a] = un[r][a];
                if constexpr (H2)
                  hl[r][a] = hn[r][a];
              }
          item = item_n;
          j = j_n;
          it = it_n;
          LU = LUn;
          ++step;
          phase_sync<SPREAD>();
        }

      return acc;
    }

    template <int P, typename Number, int COEF, bool VEL, int CW, int R, int WPB, int MINW, bool MF>
    __global__ void __launch_bounds__(64 * WPB, MINW)
      vmult_kernel(const Number *__restrict__ src, Number *__restrict__ dst,
                   const Number *__restrict__ heav, const CellParams<Number> prm,
                   const int colour, const int layers, const int items, const int item_base,
                   double *__restrict__ partials, const int partial_offset,
                   const int partial_total, const int finalize, SolverScalars *__restrict__ scal)
    {
      constexpr int THREADS = 64 * WPB;
      if (scal != nullptr && scal->done)
        return;
      // items [item_base, item_base + items) of the colour
      const double acc = vmult_body<P, Number, COEF, VEL, CW, R, WPB, MF>(
        src, dst, heav, prm, colour, layers, item_base + items, item_base + (int)blockIdx.x,
        gridDim.x);
      // p . Ap: sum over cells of u_cell . y_cell (constrained entries are 0)
      if (partials != nullptr)
        pass_partials<THREADS>(acc, partials, partial_offset, partial_total, finalize, scal);
    }

    // p . Ap of the colour passes: partials[0..total) summed in index order,
    // scal->pap and scal->alpha = rz / pap (skipped once scal->done is set).
    __global__ void __launch_bounds__(256) pap_finalize_kernel(const double *__restrict__ partials,
                                                               int total, SolverScalars *scal)
    {
      __shared__ double red[4];
      if (scal->done)
        return;
      const double pap = ordered_sum<256>(partials, total, red);
      if (threadIdx.x == 0)
        {
          scal->pap = pap;
          scal->alpha = scal->rz / pap;
        }
    }

    // Ordered schedule: the whole vmult in one launch (see vmult_body). The
    // workgroup that leaves last sums the per-item p . Ap shares in ticket order
    // (with partials and scal: scal->pap, scal->alpha = rz / pap, as the last
    // colour pass does), then resets the ticket and exit counters and advances
    // the epoch for the next launch on the stream.
    template <int P, typename Number, int COEF, bool VEL, int CW, int R, int WPB, int MINW, bool MF>
    __global__ void __launch_bounds__(64 * WPB, MINW)
      vmult_ordered_kernel(const Number *__restrict__ src, Number *__restrict__ dst,
                           const Number *__restrict__ heav, const CellParams<Number> prm,
                           const int layers, const OrderGrid og, double *__restrict__ partials,
                           SolverScalars *__restrict__ scal)
    {
      constexpr int THREADS = 64 * WPB;
      if (scal != nullptr && scal->done)
        return;
      __shared__ unsigned s_epoch;
      __shared__ bool s_last;
      __shared__ double red[WPB];
      if (threadIdx.x == 0)
        s_epoch = __hip_atomic_load(&prm.order->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
      __syncthreads();
      const unsigned epoch = __builtin_amdgcn_readfirstlane(s_epoch);
      const int total = og.off[8];
      double *ipart = (partials != nullptr && scal != nullptr) ? partials : nullptr;
      vmult_body<P, Number, COEF, VEL, CW, R, WPB, MF, true>(src, dst, heav, prm, 0, layers, total,
                                                             0, 0, og, ipart, epoch);
      // the per-item shares are agent-coherent stores: completing them orders
      // them before the exit count
      wait_vmem();
      __syncthreads();
      if (threadIdx.x == 0)
        s_last = atomicAdd(&prm.order->done, 1u) == gridDim.x - 1;
      __syncthreads();
      if (!s_last)
        return;
      if (ipart != nullptr)
        {
          const double pap = ordered_sum<THREADS>(ipart, total, red);
          if (threadIdx.x == 0)
            {
              scal->pap = pap;
              scal->alpha = scal->rz / pap;
            }
        }
      if (threadIdx.x < kOrderQueues)
        __hip_atomic_store(&prm.order->ticket[threadIdx.x][0], 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      if (threadIdx.x == 0)
        {
          __hip_atomic_store(&prm.order->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&prm.order->epoch, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }

    // Batched vmult: the workgroups of one launch are split among the
    // operators of a batch (one colour, one degree); workgroup b serves
    // operator s with block_off[s] <= b < block_off[s+1] and loops over that
    // operator's items of this colour. Each workgroup writes its p . Ap share
    // to partials[b] (per-operator sums are formed by a separate kernel).
    template <int P, typename Number, int COEF, bool VEL, int CW, int R, int WPB, int MINW, bool MF>
    __global__ void __launch_bounds__(64 * WPB, MINW)
      vmult_batch_kernel(const BatchOperator<Number> *__restrict__ ops, const int *__restrict__ block_off,
                         const int nops, const int colour, double *__restrict__ partials,
                         const SolverScalars *__restrict__ scal)
    {
      constexpr int THREADS = 64 * WPB;
      const int s = find_block_owner(block_off, nops);
      const BatchOperator<Number> &op = ops[s];
      double acc = 0.0;
      const bool skip = scal != nullptr && scal[op.solve].done;
      if (!skip)
        acc = vmult_body<P, Number, COEF, VEL, CW, R, WPB, MF>(
          op.src, op.dst, op.heav, op.prm, colour, op.prm.zlayers, op.items[colour],
          blockIdx.x - block_off[s], block_off[s + 1] - block_off[s]);
      if (partials != nullptr)
        {
          __shared__ double red[WPB];
          const double bs = block_sum<THREADS>(acc, red);
          if (threadIdx.x == 0)
            partials[blockIdx.x] = bs;
        }
    }

    // Closed-form diagonal of the cell matrix, summed like vmult's scatter.
    // diag_i = sum_q JxW [ c phi_i^2 + k sum_d (d_d phi_i)^2 + rho_cp phi_i v . grad phi_i ]
    // evaluated with squared 1D matrices: S o S, D o D, S o D.
    template <int P, typename Number, int COEF, bool VEL, int CPB, int THREADS>
    __global__ void __launch_bounds__(THREADS)
      diag_kernel(Number *__restrict__ dst, const Number *__restrict__ heav,
                  const DiagTables<Number> tab, const CellParams<Number> prm, const int colour)
    {
      constexpr int N = P + 1;
      constexpr int NL = N * N;
      constexpr int STR = lds_stride<N>();
      constexpr bool H2 = COEF == kCoefHeaviside;
      // KB/RB (material at q) alias B/C: the x-line phase reads them for the
      // same positions it then overwrites, within one thread.
      __shared__ Number A[CPB * STR];
      __shared__ Number B[CPB * STR];
      __shared__ Number C[CPB * STR];
      __shared__ Number D3[VEL ? CPB * STR : 1];
      __shared__ Number E3[VEL ? CPB * STR : 1];
      Number *KB = B;
      Number *RB = C;

      const Segment g = segment_of<Number, CPB>(prm, colour);
      const int t = threadIdx.x;
      const bool active = t < CPB * NL;
      const int l = t % CPB;
      const int s = t / CPB;
      const int s0 = s % N, s1 = s / N;
      const bool cell_ok = active && l < g.ncell;
      const int cx = min(g.cx0 + l, prm.nx - 1);
      Number *Al = A + l * STR;
      Number *Bl = B + l * STR;
      Number *Cl = C + l * STR;
      Number *Dl = D3 + (VEL ? l * STR : 0);
      Number *El = E3 + (VEL ? l * STR : 0);
      Number *KBl = KB + (H2 ? l * STR : 0);
      Number *RBl = RB + (H2 ? l * STR : 0);

      if constexpr (H2)
        {
          // heaviside at q: plain gather + three interpolation sweeps
          if (active)
            {
              const int Yi = g.cy * P + s0, Zi = g.cz * P + s1;
              const int64_t base = (g.cx0 + l) * P + (int64_t)prm.n1x * (Yi + (int64_t)prm.n1y * Zi);
              Number hl[N], h1[N];
#pragma unroll
              for (int a = 0; a < N; ++a)
                hl[a] = cell_ok ? heav[base + a] : Number(0);
              eo_apply<N, 1>(tab.S, hl, h1);
#pragma unroll
              for (int q = 0; q < N; ++q)
                Al[(s1 * N + s0) * N + q] = h1[q];
            }
          __syncthreads();
          if (active)
            {
              Number v[N], w[N];
#pragma unroll
              for (int b = 0; b < N; ++b)
                v[b] = Al[(s1 * N + b) * N + s0];
              eo_apply<N, 1>(tab.S, v, w);
#pragma unroll
              for (int b = 0; b < N; ++b)
                Al[(s1 * N + b) * N + s0] = w[b];
            }
          __syncthreads();
          if (active)
            {
              Number v[N], Hq[N];
#pragma unroll
              for (int c = 0; c < N; ++c)
                v[c] = Al[(c * N + s1) * N + s0];
              eo_apply<N, 1>(tab.S, v, Hq);
#pragma unroll
              for (int q = 0; q < N; ++q)
                {
                  Number rc, k;
                  two_phase(prm, Hq[q], rc, k);
                  KBl[(q * N + s1) * N + s0] = k;
                  RBl[(q * N + s1) * N + s0] = rc;
                }
            }
          __syncthreads();
        }

      // x-lines (qy = s0, qz = s1): coefficient fields, contract x
      if (active)
        {
          Number Fm[N], Fkx[N], Fky[N], Fkz[N], Fvx[N], Fvy[N], Fvz[N];
          const Number wyz = tab.W[s0] * tab.W[s1];
#pragma unroll
          for (int q = 0; q < N; ++q)
            {
              const Number w3 = tab.W[q] * wyz;
              const Number k = conductivity<Number, COEF, N>(prm, KBl, cx, g.cy, g.cz, q, s0, s1);
              const Number rc = H2 ? RBl[(s1 * N + s0) * N + q] : prm.rho_cp0;
              Fm[q] = (H2 ? rc * prm.dtinv_vol : prm.mass) * w3;
              Fkx[q] = k * w3 * prm.kf[0];
              Fky[q] = k * w3 * prm.kf[1];
              Fkz[q] = k * w3 * prm.kf[2];
              if constexpr (VEL)
                {
                  const Number vx = prm.vel_x[0][cx * N + q] * prm.vel_y[0][g.cy * N + s0];
                  const Number vy = prm.vel_x[1][cx * N + q] * prm.vel_y[1][g.cy * N + s0];
                  Fvx[q] = rc * w3 * prm.vf[0] * vx;
                  Fvy[q] = rc * w3 * prm.vf[1] * vy;
                  Fvz[q] = rc * w3 * prm.vf[2] * prm.w0;
                }
            }
          Number g1[N], t1[N], g2[N], g4[N];
          eo_apply<N, 1>(tab.S2t, Fm, g1);
          eo_apply<N, 1>(tab.D2t, Fkx, t1);
#pragma unroll
          for (int a = 0; a < N; ++a)
            g1[a] += t1[a];
          if constexpr (VEL)
            {
              eo_apply<N, -1>(tab.SDt, Fvx, t1);
#pragma unroll
              for (int a = 0; a < N; ++a)
                g1[a] += t1[a];
            }
          eo_apply<N, 1>(tab.S2t, Fky, g2);
          eo_apply<N, 1>(tab.S2t, Fkz, g4);
#pragma unroll
          for (int a = 0; a < N; ++a)
            {
              const int idx = (s1 * N + s0) * N + a;
              Al[idx] = g1[a];
              Bl[idx] = g2[a];
              Cl[idx] = g4[a];
            }
          if constexpr (VEL)
            {
              Number g3[N], g5[N];
              eo_apply<N, 1>(tab.S2t, Fvy, g3);
              eo_apply<N, 1>(tab.S2t, Fvz, g5);
#pragma unroll
              for (int a = 0; a < N; ++a)
                {
                  const int idx = (s1 * N + s0) * N + a;
                  Dl[idx] = g3[a];
                  El[idx] = g5[a];
                }
            }
        }
      __syncthreads();

      // y-lines (a = s0, qz = s1)
      if (active)
        {
          Number v1[N], v2[N], v4[N], h1[N], t1[N], h2[N];
#pragma unroll
          for (int r = 0; r < N; ++r)
            {
              const int idx = (s1 * N + r) * N + s0;
              v1[r] = Al[idx];
              v2[r] = Bl[idx];
              v4[r] = Cl[idx];
            }
          eo_apply<N, 1>(tab.S2t, v1, h1);
          eo_apply<N, 1>(tab.D2t, v2, t1);
#pragma unroll
          for (int b = 0; b < N; ++b)
            h1[b] += t1[b];
          eo_apply<N, 1>(tab.S2t, v4, h2);
          Number h3[N];
          if constexpr (VEL)
            {
              Number v3[N], v5[N];
#pragma unroll
              for (int r = 0; r < N; ++r)
                {
                  const int idx = (s1 * N + r) * N + s0;
                  v3[r] = Dl[idx];
                  v5[r] = El[idx];
                }
              eo_apply<N, -1>(tab.SDt, v3, t1);
#pragma unroll
              for (int b = 0; b < N; ++b)
                h1[b] += t1[b];
              eo_apply<N, 1>(tab.S2t, v5, h3);
            }
#pragma unroll
          for (int b = 0; b < N; ++b)
            {
              const int idx = (s1 * N + b) * N + s0;
              Al[idx] = h1[b];
              Bl[idx] = h2[b];
              if constexpr (VEL)
                Cl[idx] = h3[b];
            }
        }
      __syncthreads();

      // z-lines (a = s0, b = s1)
      if (active)
        {
          Number v1[N], v2[N], d[N], t1[N];
#pragma unroll
          for (int r = 0; r < N; ++r)
            {
              const int idx = (r * N + s1) * N + s0;
              v1[r] = Al[idx];
              v2[r] = Bl[idx];
            }
          eo_apply<N, 1>(tab.S2t, v1, d);
          eo_apply<N, 1>(tab.D2t, v2, t1);
#pragma unroll
          for (int c = 0; c < N; ++c)
            d[c] += t1[c];
          if constexpr (VEL)
            {
              Number v3[N];
#pragma unroll
              for (int r = 0; r < N; ++r)
                v3[r] = Cl[(r * N + s1) * N + s0];
              eo_apply<N, -1>(tab.SDt, v3, t1);
#pragma unroll
              for (int c = 0; c < N; ++c)
                d[c] += t1[c];
            }
#pragma unroll
          for (int c = 0; c < N; ++c)
            Al[(c * N + s1) * N + s0] = d[c];
        }
      __syncthreads();

      // x-lines (b = s0, c = s1): exchange shared x-face, scatter
      Number Y[N];
      if (active)
        {
#pragma unroll
          for (int a = 0; a < N; ++a)
            Y[a] = Al[(s1 * N + s0) * N + a];
          B[s * CPB + l] = Y[N - 1];
        }
      __syncthreads();
      if (cell_ok)
        {
          if (l > 0)
            Y[0] += B[s * CPB + l - 1];
          scatter_xline<P, Number, CPB>(prm, g, l, s0, s1, Y, dst);
        }
    }

    // Launch shape of the line-form vmult. Wavefront form (degrees 1-3 and the
    // MFMA form): CW cells per wavefront, R line slots per lane, 4 wavefronts
    // per workgroup (two slots per lane for the colour-pass launches of FP64
    // degree 3 and FP32 degrees 1 and 2 without a velocity or a nodal
    // heaviside: 85.3 vs 71.7 GDoF/s at 96^3 Q3, 45.2 vs 26.8 at 128^3 Q1;
    // FP32 degree 2 in two-wavefront workgroups, so that a workgroup keeps
    // its 28-cell row segment: 72.9 vs 66.7 GDoF/s at 96^3 and 108.6 vs 97.4
    // at 192^3, constant k; with those fields the second slot would cost
    // occupancy); the MFMA form holds the matrix columns in 4(P+1) more
    // VGPRs, so it takes fewer cells per wavefront at degree 4 and 6 and a
    // 256-VGPR budget at degree 5. Spread form (FMA, from degree 4, whose cells
    // do not tile a wavefront: 25, 36, 49 lines): the CPB cells of a workgroup
    // spread one line per thread over all its wavefronts (250, 252, 245 of 256
    // threads), the phases then ordered by workgroup barriers; the kernel takes
    // CW = CPB and R = 0. Register budgets: 128 VGPRs (4 waves per SIMD); FP64
    // from degree 5 256; FP32 degree 4 of an operator without a velocity or a
    // nodal heaviside (LEAN, the mixed CG's inner operator) 80 (6 waves), which
    // measured 130.6 against 128.2 GDoF/s at 4 waves (64^3, sine k) and 115
    // for the wavefront form (2 cells per wavefront, 6 waves).
    template <int P, bool FP32, bool MF = false, bool LEAN = false>
    struct VShape
    {
      static constexpr bool SPREAD = !MF && P >= 4;
      static constexpr int WPB = (FP32 && LEAN && !MF && P == 2) ? 2 : 4;
      static constexpr int CWW = MF ? (P == 4 ? 2 : P == 5 ? 3 : P == 6 ? 1 : 4)
                                    : P == 1 ? (FP32 && LEAN ? 32 : 16)
                                    : P == 2 ? (FP32 && LEAN ? 14 : 7)
                                    : P == 3 ? (!FP32 && LEAN ? 8 : 4) : 1;
      static constexpr int CPB = SPREAD ? (64 * WPB) / ((P + 1) * (P + 1)) : CWW * WPB;
      static constexpr int CW = SPREAD ? CPB : CWW;
      static constexpr int R = SPREAD ? 0 : (CWW * (P + 1) * (P + 1) + 63) / 64;
      static constexpr int MINW = (LEAN && FP32 && !MF && P == 4) ? 6
                                  : ((P >= 5 && !FP32) || (MF && P == 5)) ? 2 : kMinWavesPerSimd;
    };

    // Slab schedule of the colour passes (prm.slabs == 2): the z-chunks of the
    // work items are split into slab 0 = [0, ks) and slab 1 = [ks, nkz), and
    // each slab runs its eight passes on its own stream. Items of the two
    // slabs share only the lattice plane between chunks ks - 1 and ks. Those
    // two chunks differ in parity, so the even one's items (colours 0..3)
    // store into that plane before the odd one's items (colours 4..7)
    // read-modify-write it, as in the single-stream passes: the odd side's
    // pass 4 waits for the even side's pass 3 (event mid). Every lattice node
    // therefore receives its contributions in the same order as with the
    // single-stream passes and dst is bitwise the same; the p . Ap partials
    // are grouped per slab launch.
    template <typename Kernel, typename Number, int CPB, int THREADS>
    hipError_t vmult_slabs(Kernel kernel, const CellParams<Number> &prm, const Number *src,
                           Number *dst, const Number *heav, double *partials, SolverScalars *scal,
                           hipStream_t stream, const SlabStreams &ss, int max_blocks)
    {
      const int nsx = (prm.nx + CPB - 1) / CPB;
      const int nkz = (prm.nz + prm.zlayers - 1) / prm.zlayers;
      const int ks = nkz / 2;
      const hipStream_t st[2] = {stream, ss.side};
      const int even = ((ks - 1) & 1) == 0 ? 0 : 1; // slab whose boundary chunk is even
      hipError_t e = hipEventRecord(ss.fork, stream);
      if (e == hipSuccess)
        e = hipStreamWaitEvent(ss.side, ss.fork, 0);
      int off = 0;
      for (int c = 0; c < 8 && e == hipSuccess; ++c)
        {
          const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
          const int per_kz = ((nsx - ox + 1) / 2) * ((prm.ny - oy + 1) / 2);
          // items of colour c in chunks below kz (kz is the slowest item index)
          auto below = [&](int kz) { return per_kz * ((kz - oz + 1) / 2); };
          const int base[2] = {0, below(ks)};
          const int cnt[2] = {below(ks), below(nkz) - below(ks)};
          if (c == 4)
            e = hipStreamWaitEvent(st[1 - even], ss.mid, 0);
          for (int sl = 0; sl < 2 && e == hipSuccess; ++sl)
            {
              if (cnt[sl] == 0)
                continue;
              const int grid = std::min(cnt[sl], max_blocks);
              kernel<<<grid, THREADS, 0, st[sl]>>>(src, dst, heav, prm, c, prm.zlayers, cnt[sl],
                                                   base[sl], partials, off, 0, 0, scal);
              off += grid;
              e = hipGetLastError();
            }
          if (c == 3 && e == hipSuccess)
            e = hipEventRecord(ss.mid, st[even]);
        }
      if (e == hipSuccess)
        e = hipEventRecord(ss.join, ss.side);
      if (e == hipSuccess)
        e = hipStreamWaitEvent(stream, ss.join, 0);
      if (e == hipSuccess && partials != nullptr && scal != nullptr && off > 0)
        {
          pap_finalize_kernel<<<1, 256, 0, stream>>>(partials, off, scal);
          e = hipGetLastError();
        }
      return e;
    }

    template <typename Number, int P, int COEF, bool VEL, bool MF>
    hipError_t vmult_passes(const CellParams<Number> &prm, const VmultTables<Number> &tab,
                            const Number *src, Number *dst, const Number *heav, double *partials,
                            SolverScalars *scal, hipStream_t stream, const SlabStreams *slabs)
    {
      using Shape = VShape<P, sizeof(Number) == 4, MF, COEF != kCoefHeaviside && !VEL>;
      constexpr int THREADS = 64 * Shape::WPB;
      constexpr int CPB = Shape::CPB;
      auto kernel = vmult_kernel<P, Number, COEF, VEL, Shape::CW, Shape::R, Shape::WPB, Shape::MINW, MF>;
      static int resident = 0; // workgroups per CU, from the occupancy calculator
      if (resident == 0)
        {
          int nb = 0;
          if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, THREADS, 0) != hipSuccess || nb < 1)
            nb = 1;
          resident = nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
        }
      const int per_cu = prm.blocks_per_cu > 0 ? prm.blocks_per_cu : resident;
      const VmultGrid vg = vmult_grid(prm.nx, prm.ny, prm.nz, CPB, prm.zlayers,
                                      per_cu * device_cu_count());
      if (prm.order != nullptr)
        {
          // the ordered kernel holds more state: the non-lean shape
          using OShape = VShape<P, sizeof(Number) == 4, MF>;
          constexpr int OTHREADS = 64 * OShape::WPB;
          auto okernel =
            vmult_ordered_kernel<P, Number, COEF, VEL, OShape::CW, OShape::R, OShape::WPB, OShape::MINW, MF>;
          static int oresident = 0;
          if (oresident == 0)
            {
              int nb = 0;
              if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, okernel, OTHREADS, 0) != hipSuccess ||
                  nb < 1)
                nb = 1;
              oresident = nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
            }
          // every workgroup of an ordered launch must be resident at once: a
          // tuned count above the kernel's own occupancy is capped, so the
          // waits' forward progress does not rest on the dispatch order
          const int opc = prm.blocks_per_cu > 0 ? std::min(prm.blocks_per_cu, oresident) : oresident;
          const VmultGrid og8 = vmult_grid(prm.nx, prm.ny, prm.nz, OShape::CPB,
                                           prm.zlayers, opc * device_cu_count());
          OrderGrid og{};
          for (int c = 0; c < 8; ++c)
            og.off[c + 1] = og.off[c] + og8.items[c];
          const int grid = std::max(1, std::min(og.off[8], opc * device_cu_count()));
          okernel<<<grid, OTHREADS, 0, stream>>>(src, dst, heav, prm, prm.zlayers, og, partials, scal);
          return hipGetLastError();
        }
      if (slabs != nullptr && prm.slabs == 2 && (prm.nz + prm.zlayers - 1) / prm.zlayers >= 2)
        return vmult_slabs<decltype(kernel), Number, CPB, THREADS>(
          kernel, prm, src, dst, heav, partials, scal, stream, *slabs, per_cu * device_cu_count());
      for (int c = 0; c < 8; ++c)
        {
          if (vg.grid[c] == 0)
            continue;
          // the passes only store their p . Ap partials; one workgroup sums
          // them after the last pass (the launch boundary orders the stores),
          // so no workgroup of the last pass waits for its stores to complete
          kernel<<<vg.grid[c], THREADS, 0, stream>>>(src, dst, heav, prm, c, prm.zlayers,
                                                     vg.items[c], 0, partials, vg.offset[c],
                                                     vg.total, 0, scal);
        }
      if (partials != nullptr && scal != nullptr && vg.last >= 0)
        pap_finalize_kernel<<<1, 256, 0, stream>>>(partials, vg.total, scal);
      return hipGetLastError();
    }

    template <typename Number, int P, int COEF, bool VEL>
    hipError_t diag_passes(const CellParams<Number> &prm, const DiagTables<Number> &tab,
                           Number *dst, const Number *heav, hipStream_t stream)
    {
      constexpr bool F32 = sizeof(Number) == 4;
      constexpr int CPB = cpb<P, F32>();
      constexpr int THREADS = block_threads<P, F32>();
      const ColourGrid cg = colour_grid(prm.nx, prm.ny, prm.nz, CPB);
      for (int c = 0; c < 8; ++c)
        {
          if (cg.blocks[c] == 0)
            continue;
          diag_kernel<P, Number, COEF, VEL, CPB, THREADS><<<cg.blocks[c], THREADS, 0, stream>>>(
            dst, heav, tab, prm, c);
        }
      return hipGetLastError();
    }

    // MFMA line form exists for FP32 from degree 3 (4+ points per line)
    template <typename Number, int P>
    constexpr bool has_mfma_form()
    {
      return sizeof(Number) == 4 && P >= 3;
    }

    template <typename Number, int P, bool MF>
    hipError_t vmult_deg_mf(int coef, bool vel, const CellParams<Number> &prm,
                            const VmultTables<Number> &tab, const Number *src, Number *dst,
                            const Number *heav, double *partials, SolverScalars *scal, hipStream_t s,
                            const SlabStreams *slabs)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return vmult_passes<Number, P, 0, false, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          case 1: return vmult_passes<Number, P, 0, true, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          case 2: return vmult_passes<Number, P, 1, false, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          case 3: return vmult_passes<Number, P, 1, true, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          case 4: return vmult_passes<Number, P, 2, false, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          case 5: return vmult_passes<Number, P, 2, true, MF>(prm, tab, src, dst, heav, partials, scal, s, slabs);
          default: return hipErrorInvalidValue;
        }
    }

    template <typename Number, int P>
    hipError_t vmult_deg(int coef, bool vel, const CellParams<Number> &prm,
                         const VmultTables<Number> &tab, const Number *src, Number *dst,
                         const Number *heav, double *partials, SolverScalars *scal, hipStream_t s,
                         const SlabStreams *slabs)
    {
      if constexpr (has_mfma_form<Number, P>())
        if (prm.variant == kVariantLineMfma)
          return vmult_deg_mf<Number, P, true>(coef, vel, prm, tab, src, dst, heav, partials, scal, s, slabs);
      return vmult_deg_mf<Number, P, false>(coef, vel, prm, tab, src, dst, heav, partials, scal, s, slabs);
    }

    template <typename Number, int P>
    hipError_t diag_deg(int coef, bool vel, const CellParams<Number> &prm,
                        const DiagTables<Number> &tab, Number *dst, const Number *heav,
                        hipStream_t s)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return diag_passes<Number, P, 0, false>(prm, tab, dst, heav, s);
          case 1: return diag_passes<Number, P, 0, true>(prm, tab, dst, heav, s);
          case 2: return diag_passes<Number, P, 1, false>(prm, tab, dst, heav, s);
          case 3: return diag_passes<Number, P, 1, true>(prm, tab, dst, heav, s);
          case 4: return diag_passes<Number, P, 2, false>(prm, tab, dst, heav, s);
          case 5: return diag_passes<Number, P, 2, true>(prm, tab, dst, heav, s);
          default: return hipErrorInvalidValue;
        }
    }
  } // namespace

  namespace
  {
    template <int P>
    double ctable_defect_p(const Tables1D &t)
    {
      constexpr ct::Tab<P + 1> c = ct::make_tab<P>();
      double d = 0.0;
      for (int i = 0; i <= P; ++i)
        {
          d = fmax(d, fabs(c.gll[i] - t.gll[i]));
          d = fmax(d, fabs(c.x[i] - t.xq[i]));
          d = fmax(d, fabs(c.w[i] - t.wq[i]));
          for (int j = 0; j <= P; ++j)
            {
              d = fmax(d, fabs(c.S[i][j] - t.S[i][j]));
              d = fmax(d, fabs(c.C[i][j] - t.C[i][j]) / (1.0 + fabs(t.C[i][j])));
              d = fmax(d, fabs(c.D[i][j] - t.D[i][j]) / (1.0 + fabs(t.D[i][j])));
            }
        }
      return d;
    }
  } // namespace

  double ctable_defect(int degree, const Tables1D &t)
  {
    switch (degree)
      {
        case 1: return ctable_defect_p<1>(t);
        case 2: return ctable_defect_p<2>(t);
        case 3: return ctable_defect_p<3>(t);
        case 4: return ctable_defect_p<4>(t);
        case 5: return ctable_defect_p<5>(t);
        case 6: return ctable_defect_p<6>(t);
        default: return 1.0;
      }
  }

  int cells_per_block(int degree, bool fp32) { return cells_per_block_for(degree, fp32); }

  int threads_per_block(int degree, bool fp32)
  {
    const int n = degree + 1;
    return ((cells_per_block(degree, fp32) * n * n + 63) / 64) * 64;
  }

  ColourGrid colour_grid(int nx, int ny, int nz, int cpb_)
  {
    ColourGrid g{};
    const int nsx = (nx + cpb_ - 1) / cpb_;
    int off = 0;
    g.last = -1;
    for (int c = 0; c < 8; ++c)
      {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        const int m = ((nsx - ox + 1) / 2) * ((ny - oy + 1) / 2) * ((nz - oz + 1) / 2);
        g.blocks[c] = m;
        g.offset[c] = off;
        off += m;
        if (m > 0)
          g.last = c;
      }
    g.total = off;
    return g;
  }

  int device_cu_count()
  {
    static int cus = 0;
    if (cus == 0)
      {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1)
          n = 256;
        cus = n;
      }
    return cus;
  }

  VmultGrid vmult_grid(int nx, int ny, int nz, int cpb_, int layers, int max_blocks)
  {
    VmultGrid g{};
    const int nsx = (nx + cpb_ - 1) / cpb_;
    const int nkz = (nz + layers - 1) / layers;
    int off = 0;
    g.last = -1;
    for (int c = 0; c < 8; ++c)
      {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        const int m = ((nsx - ox + 1) / 2) * ((ny - oy + 1) / 2) * ((nkz - oz + 1) / 2);
        g.items[c] = m;
        g.grid[c] = m < max_blocks ? m : max_blocks;
        g.offset[c] = off;
        off += g.grid[c];
        if (m > 0)
          g.last = c;
      }
    g.total = off;
    return g;
  }

  template <typename Number>
  hipError_t launch_vmult(int degree, int coef, bool vel, const CellParams<Number> &prm,
                          const VmultTables<Number> &tab, const Number *src, Number *dst,
                          const Number *heaviside, double *partials, SolverScalars *scal,
                          hipStream_t stream, const SlabStreams *slabs)
  {
    if (prm.variant == kVariantCell && degree <= kCellMaxDegree)
      return launch_vmult_cell<Number>(degree, coef, vel, prm, src, dst, heaviside, partials, scal,
                                       stream);
    if (prm.variant == kVariantCellMfma && degree <= kEmMaxDegree)
      return launch_vmult_em<Number>(degree, prm, src, dst, partials, scal, stream);
    switch (degree)
      {
        case 1: return vmult_deg<Number, 1>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        case 2: return vmult_deg<Number, 2>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        case 3: return vmult_deg<Number, 3>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        case 4: return vmult_deg<Number, 4>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        case 5: return vmult_deg<Number, 5>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        case 6: return vmult_deg<Number, 6>(coef, vel, prm, tab, src, dst, heaviside, partials, scal, stream, slabs);
        default: return hipErrorInvalidValue;
      }
  }

  namespace
  {
    template <typename Number, int P, int COEF, bool VEL, bool MF>
    hipError_t vmult_batch_passes(const BatchOperator<Number> *ops, const int *block_off, int nops,
                                  const BatchVmultPlan &plan, double *partials,
                                  const SolverScalars *scal, hipStream_t stream)
    {
      using Shape = VShape<P, sizeof(Number) == 4, MF>;
      constexpr int THREADS = 64 * Shape::WPB;
      auto kernel = vmult_batch_kernel<P, Number, COEF, VEL, Shape::CW, Shape::R, Shape::WPB, Shape::MINW, MF>;
      for (int c = 0; c < 8; ++c)
        {
          if (plan.grid[c] == 0)
            continue;
          kernel<<<plan.grid[c], THREADS, 0, stream>>>(ops, block_off + c * (nops + 1), nops, c,
                                                       partials ? partials + plan.pofs[c] : nullptr,
                                                       scal);
        }
      return hipGetLastError();
    }

    template <typename Number, int P, bool MF>
    hipError_t vmult_batch_deg_mf(int coef, bool vel, const BatchOperator<Number> *ops,
                                  const int *block_off, int nops, const BatchVmultPlan &plan,
                                  double *partials, const SolverScalars *scal, hipStream_t s)
    {
      switch (coef * 2 + (vel ? 1 : 0))
        {
          case 0: return vmult_batch_passes<Number, P, 0, false, MF>(ops, block_off, nops, plan, partials, scal, s);
          case 1: return vmult_batch_passes<Number, P, 0, true, MF>(ops, block_off, nops, plan, partials, scal, s);
          case 2: return vmult_batch_passes<Number, P, 1, false, MF>(ops, block_off, nops, plan, partials, scal, s);
          case 3: return vmult_batch_passes<Number, P, 1, true, MF>(ops, block_off, nops, plan, partials, scal, s);
          case 4: return vmult_batch_passes<Number, P, 2, false, MF>(ops, block_off, nops, plan, partials, scal, s);
          case 5: return vmult_batch_passes<Number, P, 2, true, MF>(ops, block_off, nops, plan, partials, scal, s);
          default: return hipErrorInvalidValue;
        }
    }

    template <typename Number, int P>
    hipError_t vmult_batch_deg(int coef, bool vel, int variant, const BatchOperator<Number> *ops,
                               const int *block_off, int nops, const BatchVmultPlan &plan,
                               double *partials, const SolverScalars *scal, hipStream_t s)
    {
      if constexpr (has_mfma_form<Number, P>())
        if (variant == kVariantLineMfma)
          return vmult_batch_deg_mf<Number, P, true>(coef, vel, ops, block_off, nops, plan, partials, scal, s);
      return vmult_batch_deg_mf<Number, P, false>(coef, vel, ops, block_off, nops, plan, partials, scal, s);
    }

    template <typename Number, int P, int COEF, bool VEL, bool MF>
    int resident_of()
    {
      using Shape = VShape<P, sizeof(Number) == 4, MF>;
      auto kernel = vmult_batch_kernel<P, Number, COEF, VEL, Shape::CW, Shape::R, Shape::WPB, Shape::MINW, MF>;
      int nb = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64 * Shape::WPB, 0) != hipSuccess || nb < 1)
        nb = 1;
      return nb < kMaxResidentPerCu ? nb : kMaxResidentPerCu;
    }

    // One workgroup per operator: thread t adds the operator's blocks t,
    // t + 256, ... of each colour (colours in order, loads sixteen at a time),
    // then the workgroup adds the threads' sums in a fixed tree, so the sum has
    // a fixed order. A member with hundreds of blocks per colour took ~45 us
    // with one thread walking its blocks; this is one pass of loads.
    __global__ void __launch_bounds__(256)
      batch_pap_finalize_kernel(const int *block_off, const int *ops_solve, int nops,
                                BatchVmultPlan plan, const double *partials, SolverScalars *scal)
    {
      __shared__ double red[4];
      const int o = blockIdx.x;
      SolverScalars &sc = scal[ops_solve[o]];
      if (sc.done)
        return;
      double v = 0.0;
      for (int c = 0; c < 8; ++c)
        {
          const int *bo = block_off + c * (nops + 1);
          v += strided_sum(partials + plan.pofs[c] + bo[o], threadIdx.x, bo[o + 1] - bo[o], 256);
        }
      const double pap = block_sum<256>(v, red);
      if (threadIdx.x == 0)
        {
          sc.pap = pap;
          sc.alpha = sc.rz / pap;
        }
    }
  } // namespace

  BatchVmultPlan batch_vmult_plan(const int *items, int nops, int max_blocks, int *block_off)
  {
    BatchVmultPlan plan{};
    int slot = 0;
    for (int c = 0; c < 8; ++c)
      {
        long total = 0;
        for (int o = 0; o < nops; ++o)
          total += items[o * 8 + c];
        // items per workgroup so that the colour's launch stays within max_blocks
        const long per = total > max_blocks ? (total + max_blocks - 1) / max_blocks : 1;
        int *bo = block_off + c * (nops + 1);
        bo[0] = 0;
        for (int o = 0; o < nops; ++o)
          {
            const int it = items[o * 8 + c];
            const int g = it == 0 ? 0 : (int)((it + per - 1) / per);
            bo[o + 1] = bo[o] + g;
          }
        plan.grid[c] = bo[nops];
        plan.pofs[c] = slot;
        slot += plan.grid[c];
      }
    plan.total_slots = slot;
    return plan;
  }

  int vmult_cells_per_block(int degree, bool fp32, bool mf)
  {
#define MPDG_CPB(P)                                                                               \
  case P:                                                                                         \
    return mf ? VShape<P, true, true>::CPB : fp32 ? VShape<P, true>::CPB : VShape<P, false>::CPB;
    switch (degree)
      {
        MPDG_CPB(1)
        MPDG_CPB(2)
        MPDG_CPB(3)
        MPDG_CPB(4)
        MPDG_CPB(5)
        MPDG_CPB(6)
        default: return 1;
      }
#undef MPDG_CPB
  }

  template <typename Number>
  VmultGrid vmult_items(int degree, const CellParams<Number> &prm)
  {
    if (prm.variant == kVariantCell && degree <= kCellMaxDegree)
      return cell_grid(prm.nx, prm.ny, prm.nz, prm.tile_log2, prm.zlayers, 1 << 30);
    const bool mf = prm.variant == kVariantLineMfma && sizeof(Number) == 4 && degree >= kMfmaMinDegree;
    return vmult_grid(prm.nx, prm.ny, prm.nz, vmult_cells_per_block(degree, sizeof(Number) == 4, mf),
                      prm.zlayers, 1 << 30);
  }
  template VmultGrid vmult_items<double>(int, const CellParams<double> &);
  template VmultGrid vmult_items<float>(int, const CellParams<float> &);

  int vmult_resident_blocks(int degree, bool fp32, int coef, bool vel, int variant)
  {
    (void)coef;
    (void)vel;
    if (variant == kVariantCell && degree <= kCellMaxDegree)
      return cell_resident_blocks(degree, fp32);
#define MPDG_RES(P)                                                                               \
  case P:                                                                                         \
    return mf && P >= kMfmaMinDegree ? resident_of<float, P, 0, false, P >= kMfmaMinDegree>()     \
           : fp32 ? resident_of<float, P, 0, false, false>() : resident_of<double, P, 0, false, false>();
    const bool mf = fp32 && variant == kVariantLineMfma;
    switch (degree)
      {
        MPDG_RES(1)
        MPDG_RES(2)
        MPDG_RES(3)
        MPDG_RES(4)
        MPDG_RES(5)
        MPDG_RES(6)
        default: return 1;
      }
#undef MPDG_RES
  }

  template <typename Number>
  hipError_t launch_vmult_batch(int degree, int coef, bool vel, int variant,
                                const BatchOperator<Number> *ops, const int *block_off, int nops,
                                const BatchVmultPlan &plan, double *partials,
                                const SolverScalars *scal, hipStream_t stream)
  {
    if (variant == kVariantCell && degree <= kCellMaxDegree)
      return launch_vmult_cell_batch<Number>(degree, coef, vel, ops, block_off, nops, plan,
                                             partials, scal, stream);
    switch (degree)
      {
        case 1: return vmult_batch_deg<Number, 1>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        case 2: return vmult_batch_deg<Number, 2>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        case 3: return vmult_batch_deg<Number, 3>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        case 4: return vmult_batch_deg<Number, 4>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        case 5: return vmult_batch_deg<Number, 5>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        case 6: return vmult_batch_deg<Number, 6>(coef, vel, variant, ops, block_off, nops, plan, partials, scal, stream);
        default: return hipErrorInvalidValue;
      }
  }

  hipError_t launch_batch_pap_finalize(const int *block_off, const int *ops_solve, int nops,
                                       const BatchVmultPlan &plan, const double *partials,
                                       SolverScalars *scal, hipStream_t stream)
  {
    batch_pap_finalize_kernel<<<nops, 256, 0, stream>>>(block_off, ops_solve, nops, plan, partials, scal);
    return hipGetLastError();
  }

  template hipError_t launch_vmult_batch<double>(int, int, bool, int, const BatchOperator<double> *,
                                                 const int *, int, const BatchVmultPlan &, double *,
                                                 const SolverScalars *, hipStream_t);
  template hipError_t launch_vmult_batch<float>(int, int, bool, int, const BatchOperator<float> *,
                                                const int *, int, const BatchVmultPlan &, double *,
                                                const SolverScalars *, hipStream_t);

  template <typename Number>
  hipError_t launch_diagonal(int degree, int coef, bool vel, const CellParams<Number> &prm,
                             const DiagTables<Number> &tab, Number *diag,
                             const Number *heaviside, hipStream_t stream)
  {
    switch (degree)
      {
        case 1: return diag_deg<Number, 1>(coef, vel, prm, tab, diag, heaviside, stream);
        case 2: return diag_deg<Number, 2>(coef, vel, prm, tab, diag, heaviside, stream);
        case 3: return diag_deg<Number, 3>(coef, vel, prm, tab, diag, heaviside, stream);
        case 4: return diag_deg<Number, 4>(coef, vel, prm, tab, diag, heaviside, stream);
        case 5: return diag_deg<Number, 5>(coef, vel, prm, tab, diag, heaviside, stream);
        case 6: return diag_deg<Number, 6>(coef, vel, prm, tab, diag, heaviside, stream);
        default: return hipErrorInvalidValue;
      }
  }

  template hipError_t launch_vmult<double>(int, int, bool, const CellParams<double> &,
                                           const VmultTables<double> &, const double *, double *,
                                           const double *, double *, SolverScalars *, hipStream_t,
                                           const SlabStreams *);
  template hipError_t launch_vmult<float>(int, int, bool, const CellParams<float> &,
                                          const VmultTables<float> &, const float *, float *,
                                          const float *, double *, SolverScalars *, hipStream_t,
                                          const SlabStreams *);
  template hipError_t launch_diagonal<double>(int, int, bool, const CellParams<double> &,
                                              const DiagTables<double> &, double *, const double *,
                                              hipStream_t);
  template hipError_t launch_diagonal<float>(int, int, bool, const CellParams<float> &,
                                             const DiagTables<float> &, float *, const float *,
                                             hipStream_t);
} // namespace mpdg
