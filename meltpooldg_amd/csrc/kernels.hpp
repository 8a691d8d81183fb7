// Shared declarations between the C-ABI layer and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tables.hpp"

namespace mpdg
{
  // Largest vector the cell kernels address (bytes, exclusive): their raw
  // buffer resources carry 32-bit sizes and offsets, and the cell form marks a
  // skipped node by bit 31 of its byte offset. mpdg_heat_create rejects meshes
  // whose FP64 vectors reach it.
  constexpr int64_t kMaxVectorBytes = (int64_t)1 << 31;

  // Byte size of a lattice vector for a raw buffer resource (the host has
  // checked that it is below kMaxVectorBytes, so the 32-bit result is exact).
  __host__ __device__ inline uint32_t vector_bytes(int n1x, int n1y, int n1z, int size)
  {
    return (uint32_t)((int64_t)n1x * n1y * n1z * size);
  }

  // Per-handle operator constants passed by value to the cell kernels.
  template <typename Number>
  struct CellParams
  {
    int nx, ny, nz;       // cells per direction
    int n1x, n1y, n1z;    // lattice nodes per direction
    unsigned faces;       // Dirichlet face mask (bit 2d: lo, 2d+1: hi)
    Number mass;          // rho_cp / dt * |J| for a constant material
    Number dtinv_vol;     // |J| / dt (two-phase: multiplied by rho_cp(H) at q)
    Number vol;           // |J| (cell volume, affine Cartesian cells)
    Number kf[3];         // |J| / h_d^2 (Laplace metric, affine Cartesian cells)
    Number vf[3];         // |J| / h_d (convection metric)
    Number k0;            // material conductivity (constant material)
    Number kamp;          // separable-sine conductivity amplitude
    Number rho_cp0;       // constant rho*cp (convection term)
    Number w0;            // z velocity (constant over the box)
    Number g_rho, g_cp, g_k, l_rho, l_cp, l_k, g_rhocp, l_rhocp;
    int sharp, volume_specific;
    // temperature-dependent material: tmat = 2 liquid_solid, 3 gas_liquid_solid
    // (0 otherwise); solid phase, solid-fraction law (mushy zone or sharp at the
    // solidus) with inv_mushy = 1 / (liquidus - solidus)
    int tmat, sl_sharp;
    Number s_rho, s_cp, s_k, s_rhocp;
    Number solidus, liquidus, inv_mushy;
    // apparent capacity of the latent heat of fusion (latent = 0: none):
    // ac_type 0 poly4_bell, 1 constant, 2 qlq
    int ac_type;
    Number latent;
    int zlayers;              // cell layers per z-chunk of a vmult work item
    int blocks_per_cu;        // persistent workgroups per CU (0: occupancy limit)
    int variant;              // vmult form: kVariantLine or kVariantCell
    int tile_log2;            // cell form: log2 of the tile's cells along x (3..6)
    // line forms, ordered schedule (order != nullptr): one launch per vmult,
    // work items taken in colour order from order->ticket, each waiting for the
    // completion flags (order_flags, one per item, >= nx*ny*nz entries) of its
    // neighbouring items of earlier colours
    struct OrderState *order;
    unsigned *order_flags;
    // line forms, slab schedule (slabs == 2, with the handle's SlabStreams):
    // the colour passes split into two z-slabs of work items on two streams
    int slabs;
    const Number *sin2pi[3]; // sin(2 pi x) at the quadrature coordinates of each direction
    // in-plane velocity at q, separable: u_x = vel_x[0](x) vel_y[0](y), u_y = vel_x[1](x) vel_y[1](y)
    // (tables over the x resp. y quadrature coordinates); u_z = w0
    const Number *vel_x[2];
    const Number *vel_y[2];
    // element-matrix form (kVariantCellMfma): the cell matrix packed per lane
    // as the MFMA A operands, em[(rb * T + ks) * 64 + lane] (EmShape)
    const Number *em;
  };

  // Element-matrix form of degree P: T node slots per lane (k-steps of the
  // 16x16x4 products, node g + 4t for lane group g), RB blocks of 16 rows.
  template <int P>
  struct EmShape
  {
    static constexpr int NN = (P + 1) * (P + 1) * (P + 1);
    static constexpr int T = (NN + 3) / 4;
    static constexpr int RB = (T + 3) / 4;
  };

  // Device state of the ordered vmult schedule, one per operator precision
  // (zero-initialised once). ticket: next work item of each queue; done: workgroups that have
  // left; epoch: number of finished launches (an item's flag holds the epoch of
  // the launch that completed it, so flags never need clearing); error: set if
  // a bounded wait for a neighbour ran out (results then unreliable).
  constexpr int kOrderQueues = 8; // one ticket queue per XCD
  struct OrderState
  {
    unsigned ticket[kOrderQueues][32]; // queue q's counter in its own 128-byte line
    unsigned done, epoch, error;
  };

  // 1D tables of the vmult kernel in even-odd form (device copy by value).
  template <typename Number>
  struct VmultTables
  {
    EvenOdd<Number> S, St, C, Ct;
    Number W[kMaxN];
  };

  // 1D tables of the closed-form diagonal kernel.
  template <typename Number>
  struct DiagTables
  {
    EvenOdd<Number> S, S2t, D2t, SDt;
    Number W[kMaxN];
  };

  // Device-resident Krylov scalars; the kernels read and update them so the
  // host only polls for convergence every few iterations.
  struct SolverScalars
  {
    double rz, rz_old, pap, alpha, beta, rr, res, tol;
    double res0;
    double alpha_prev; // alpha of the first iteration of a pair (paired x update)
    int iteration, done, status, max_it;
    unsigned counter;
    unsigned pad;
    unsigned long long t_done; // device wall clock when done was set (batched solves)
  };

  // Velocity of the advection term: none, the separable tables of CellParams
  // (vortex or constant), or a nodal DoF-vector field
  enum VelocityMode
  {
    kVelNone = 0,
    kVelTables = 1,
    kVelNodal = 2
  };

  enum CoefMode
  {
    kCoefConstant = 0,
    kCoefSine = 1,
    kCoefHeaviside = 2,
    kCoefTemperature = 3 // material evaluated at T(q) (and H(q)): liquid_solid, gas_liquid_solid
  };

  // Form of the vmult kernel. Line: the lanes of a wavefront own the 1D lines
  // of a few cells and move between line directions through LDS (any degree).
  // Cell: each lane owns one whole cell in registers, a wavefront a tile of
  // 64 cells; shared faces are summed with cross-lane shuffles (degree <= 2).
  enum VmultVariant
  {
    kVariantLine = 0,
    kVariantCell = 1,
    kVariantLineMfma = 2, // line form with the 1D contractions on FP32 MFMA (degree >= 3)
    kVariantCellMfma = 3  // element-matrix form on MFMA (constant coefficients, degree <= 2)
  };
  constexpr int kMfmaMinDegree = 3;
  constexpr int kCellMaxDegree = 2;
  constexpr int kEmMaxDegree = 2;

  // Launch shape of one colour pass.
  struct ColourGrid
  {
    int blocks[8];
    int offset[8];  // prefix sum of blocks (index of the pass's first partial)
    int total;
    int last;       // last colour with blocks (finalises the partial dot)
  };

  // Cells per workgroup (one row segment along x). Sized so that 64-lane
  // wavefronts are mostly full and the LDS tensors fit the 160 KiB of a CU.
  constexpr int cells_per_block_for(int degree, bool fp32)
  {
    return degree <= 2 ? 32 : (degree <= 4 ? 16 : (fp32 ? 16 : 8));
  }
  int cells_per_block(int degree, bool fp32);
  // max difference between the compile-time and the long-double host tables
  double ctable_defect(int degree, const Tables1D &t);

  // Persistent launch shape of the vmult: per colour the number of work items
  // (segment x row x z-chunk) and of workgroups looping over them.
  // Two-slab schedule of the line forms' colour passes: the handle's second
  // stream and the events that fork it from, and join it to, the launch
  // stream, plus the event that orders the two slabs at their shared plane.
  struct SlabStreams
  {
    hipStream_t side;
    hipEvent_t fork, join, mid;
  };

  struct VmultGrid
  {
    int items[8];
    int grid[8];
    int offset[8]; // index of the colour's first partial dot
    int total;
    int last;
  };
  VmultGrid vmult_grid(int nx, int ny, int nz, int cpb, int layers, int max_blocks);
  int device_cu_count();
  // Upper bound on the partial-dot slots of one vmult.
  constexpr int kMaxResidentPerCu = 16;
  int threads_per_block(int degree, bool fp32);
  ColourGrid colour_grid(int nx, int ny, int nz, int cpb);

  // Cell-loop vmult: dst = A src over all 8 colour passes. With partials (and
  // scal) it also reduces p.Ap into scal->pap and sets scal->alpha = rz / pap;
  // with scal it does nothing once scal->done is set. slabs: the handle's
  // second stream and events, used when prm.slabs == 2 (line forms).
  template <typename Number>
  hipError_t launch_vmult(int degree, int coef, bool vel, const CellParams<Number> &prm,
                          const VmultTables<Number> &tab, const Number *src, Number *dst,
                          const Number *heaviside, double *partials, SolverScalars *scal,
                          hipStream_t stream, const SlabStreams *slabs = nullptr);

  // Cell form of the vmult (cell_kernels.hip), same contract as launch_vmult,
  // degree <= kCellMaxDegree; launch_vmult forwards to it when
  // prm.variant == kVariantCell.
  template <typename Number>
  hipError_t launch_vmult_cell(int degree, int coef, bool vel, const CellParams<Number> &prm,
                               const Number *src, Number *dst, const Number *heaviside,
                               double *partials, SolverScalars *scal, hipStream_t stream);
  // Per colour: tiles of 2^tile_log2 x 2^(6-tile_log2) cells times z-chunks,
  // four tiles in flight per workgroup.
  VmultGrid cell_grid(int nx, int ny, int nz, int tile_log2, int layers, int max_blocks);
  int cell_items_per_block();
  int cell_resident_blocks(int degree, bool fp32);

  // Element-matrix form of the vmult (em_kernels.hip), same contract as
  // launch_vmult for degree <= kEmMaxDegree with prm.em set; launch_vmult
  // forwards to it when prm.variant == kVariantCellMfma.
  template <typename Number>
  hipError_t launch_vmult_em(int degree, const CellParams<Number> &prm, const Number *src,
                             Number *dst, double *partials, SolverScalars *scal,
                             hipStream_t stream);
  // Per colour: tiles of 2^tile_log2 x 2^(4-tile_log2) cells (tile_log2 1..4)
  // times z-chunks, four tiles in flight per workgroup.
  VmultGrid em_grid(int nx, int ny, int nz, int tile_log2, int layers, int max_blocks);

  // Residual right-hand side (rhs_kernels.hip): dst += -(rho_cp/dt M (T - T_old)
  // - M_1 f + K T + rho_cp C_u T) on the free rows, T, T_old, f read plain
  // (constrained values included); heat_source may be null.
  // vmode: kVelNone, kVelTables (prm.vel_* tables and w0) or kVelNodal
  // (velocity: three component-blocked nodal vectors).
  hipError_t launch_create_rhs(int degree, int coef, int vmode, const CellParams<double> &prm,
                               double *dst, const double *temperature,
                               const double *temperature_old, const double *heat_source,
                               const double *heaviside, const double *velocity,
                               hipStream_t stream);

  // *out = sum over cells and Gauss points of JxW u^2 (FE L2 norm squared);
  // partials needs l2_norm_blocks() slots, *ticket must start at 0 (reset on exit).
  int l2_norm_blocks(int degree, int nx, int ny, int nz);
  // Tangent with the coefficients evaluated per quadrature point from nodal
  // fields (material_kernels.hip), FP64: a temperature-dependent material
  // (prm.tmat >= 2: the material and its temperature derivatives at T(q), T
  // and T_old read plain) or a constant-density one, the heaviside when given,
  // and the velocity of vmode (kVelNodal: three component-blocked nodal
  // vectors). dst = J src over 8 colour passes (coloured first-touch scatter).
  // Same p.Ap / stop contract as launch_vmult.
  hipError_t launch_tmat_vmult(int degree, int vmode, const CellParams<double> &prm,
                               const double *src, double *dst, const double *temperature,
                               const double *temperature_old, const double *heaviside,
                               const double *velocity, double *partials, SolverScalars *scal,
                               hipStream_t stream);
  // Diagonal of the same tangent (not inverted), dst overwritten.
  hipError_t launch_tmat_diagonal(int degree, int vmode, const CellParams<double> &prm,
                                  const DiagTables<double> &tab, double *dst,
                                  const double *temperature, const double *temperature_old,
                                  const double *heaviside, const double *velocity,
                                  hipStream_t stream);
  // Workgroup partial slots of launch_tmat_vmult.
  int tmat_blocks(int degree, int nx, int ny, int nz);

  hipError_t launch_l2_norm_sq(int degree, const CellParams<double> &prm, const double *u,
                               double *partials, unsigned *ticket, double *out, hipStream_t s);

  template <typename Number>
  hipError_t launch_diagonal(int degree, int coef, bool vel, const CellParams<Number> &prm,
                             const DiagTables<Number> &tab, Number *diag,
                             const Number *heaviside, hipStream_t stream);

  // Robin boundary faces of one operator (face_kernels.hip): convection
  // alpha (T - T_inf) and radiation eps sigma (T^4 - T_inf^4) on box faces
  // f = 2d + side. A wavefront per face cell integrates it into a scratch
  // row of (p+1)^2 values; then one thread per node of a Robin face adds the
  // rows of the face cells around it. A node on several Robin faces belongs to
  // the lowest of them, which sums all their contributions in ascending face
  // order, so every row is written once, in a fixed order.
  struct RobinParams
  {
    int nfaces;           // Robin faces of the launch, ascending face id
    int face[6];
    int64_t node_off[7];  // prefix sums of the faces' node lattices (thread ranges)
    int64_t cell_off[7];  // prefix sums of the faces' face cells (scratch rows)
    unsigned robin;       // mask of the Robin faces
    unsigned conv, rad;   // convection / radiation face masks
    unsigned dirichlet;   // Dirichlet face mask
    int n[3], n1[3];      // cells and lattice nodes per direction
    double h[3];          // cell extents
    double alpha, t_inf_conv, eps_sigma, t_inf_rad;
    // residual only: Neumann faces (in `robin` too) and their -(q . n) per face
    unsigned neumann;
    double qn[6];
  };
  enum RobinMode
  {
    kRobinTangent = 0,  // dst += A_face src (src and T constraint-resolved)
    kRobinDiagonal = 1, // dst += diag(A_face)
    kRobinRhs = 2       // dst += -(alpha (T - T_inf) + eps sigma (T^4 - T_inf^4) + q . n) faces, T plain
  };
  constexpr int kRobinThreads = 256;
  // workgroups of the node launch (partial-dot slots it needs)
  int robin_blocks(const RobinParams &rp);
  // doubles of the face-cell scratch rows
  int64_t robin_scratch(const RobinParams &rp, int degree);
  // Tangent with scal != nullptr: also adds p.(A_face p) to the p.Ap that the
  // cell vmult has just stored in scal->pap and resets scal->alpha; does
  // nothing once scal->done is set. partials needs robin_blocks() slots,
  // scratch robin_scratch() doubles.
  template <typename Number>
  hipError_t launch_robin(int degree, int mode, const RobinParams &rp, Number *dst,
                          const Number *src, const double *temperature, double *scratch,
                          double *partials, SolverScalars *scal, hipStream_t stream);

  // One operator of a batch: its parameters, the vectors its vmult reads and
  // writes, its work items per colour and the solve it belongs to.
  template <typename Number>
  struct BatchOperator
  {
    CellParams<Number> prm;
    const Number *src;
    Number *dst;
    const Number *heav;
    int items[8];
    int solve;
  };

  // Launch plan of a batched vmult over the operators of one degree.
  struct BatchVmultPlan
  {
    int grid[8];      // workgroups per colour
    int pofs[8];      // first partial slot of each colour
    int total_slots;  // partial slots over all colours
  };

  // Builds block_off (8 x (nops+1), host) and the plan from per-operator items.
  BatchVmultPlan batch_vmult_plan(const int *items /* nops x 8 */, int nops, int max_blocks,
                                  int *block_off);
  int vmult_resident_blocks(int degree, bool fp32, int coef, bool vel, int variant);
  int vmult_cells_per_block(int degree, bool fp32, bool mfma);
  // Work items per colour of one operator's vmult in its kernel form.
  template <typename Number>
  VmultGrid vmult_items(int degree, const CellParams<Number> &prm);

  template <typename Number>
  hipError_t launch_vmult_batch(int degree, int coef, bool vel, int variant,
                                const BatchOperator<Number> *ops,
                                const int *block_off, int nops, const BatchVmultPlan &plan,
                                double *partials, const SolverScalars *scal, hipStream_t stream);
  template <typename Number>
  hipError_t launch_vmult_cell_batch(int degree, int coef, bool vel,
                                     const BatchOperator<Number> *ops, const int *block_off,
                                     int nops, const BatchVmultPlan &plan, double *partials,
                                     const SolverScalars *scal, hipStream_t stream);
  // alpha_s = rz_s / sum of solve s's vmult partials (one block per operator).
  hipError_t launch_batch_pap_finalize(const int *block_off, const int *ops_solve, int nops,
                                       const BatchVmultPlan &plan, const double *partials,
                                       SolverScalars *scal, hipStream_t stream);

  // Vector kernels (solver_kernels.hip).
  constexpr int kVecBlocks = 1024;
  constexpr int kVecThreads = 256;

  hipError_t launch_invert_diagonal(double *d, int64_t n, hipStream_t s);
  hipError_t launch_jacobi(double *dst, const double *src, const double *inv, int64_t n, hipStream_t s);

  template <typename Number, typename InvT = Number>
  hipError_t launch_cg_init(const double *b, const Number *ax, Number *r, Number *p,
                            const InvT *inv_diag, int64_t n, double rel_tol, double abs_tol,
                            int max_it, double *partials, SolverScalars *scal, hipStream_t s);
  // x update of a CG iteration: every iteration (x += alpha p); or in pairs:
  // none on the first iteration of a pair (its alpha kept), both on the
  // second, x = (x + alpha_prev p_prev) + alpha p, bitwise the same x.
  constexpr int kXEvery = 0, kXDefer = 1, kXPair = 2;

  template <typename Number, typename InvT = Number>
  hipError_t launch_cg_update(Number *x, Number *r, const Number *p, const Number *q,
                              const InvT *inv_diag, int64_t n, double *partials,
                              SolverScalars *scal, double *history, int history_len,
                              hipStream_t s, int x_mode = kXEvery,
                              const Number *p_prev = nullptr);
  // p = D r + beta p_in (p_in = nullptr: in place)
  template <typename Number, typename InvT = Number>
  hipError_t launch_cg_direction(Number *p, const Number *r, const InvT *inv_diag, int64_t n,
                                 const SolverScalars *scal, hipStream_t s,
                                 const Number *p_in = nullptr);
  // x += alpha p with scal->alpha, run after a solve that stopped on the
  // first iteration of a pair
  template <typename Number>
  hipError_t launch_cg_flush(Number *x, const Number *p, int64_t n, const SolverScalars *scal,
                             hipStream_t s);

  // Mixed precision helpers.
  hipError_t launch_residual(double *r, const double *b, const double *ax, int64_t n,
                             double *partials, SolverScalars *scal, hipStream_t s);
  hipError_t launch_zero_constrained(double *v, int n1x, int n1y, int n1z, unsigned faces,
                                     hipStream_t s);
  // dst = a x + b y (dst may alias x or y)
  hipError_t launch_axpby(double *dst, double a, const double *x, double b, const double *y,
                          int64_t n, hipStream_t s);
  hipError_t launch_to_f32(float *dst, const double *src, int64_t n, hipStream_t s);
  // FP64 -> bfloat16 (the FP32 bit pattern's upper half, rounded to nearest even)
  hipError_t launch_to_bf16(uint16_t *dst, const double *src, int64_t n, hipStream_t s);
  hipError_t launch_add_f32(double *x, const float *d, int64_t n, hipStream_t s);
  hipError_t launch_fill(double *x, double v, int64_t n, hipStream_t s);
  // lattice[index[i]] = dofs[i] (to_lattice) or dofs[i] = lattice[index[i]]
  hipError_t launch_permute(double *dst, const double *src, const uint32_t *index, int64_t n,
                            bool to_lattice, hipStream_t s);
  hipError_t launch_fill_f32(float *x, float v, int64_t n, hipStream_t s);
  // FE_Q_iso_Q1(p) nodal values (equidistant lattice) -> FE_Q(p) nodal values
  // (GLL lattice): each GLL support point a of a cell lies in the iso element's
  // sub-interval s[a] at fraction t[a] per direction; the value there is the
  // iso field's trilinear interpolant in that sub-cell.
  struct IsoQ1Map
  {
    int s[kMaxN];
    double t[kMaxN];
  };
  hipError_t launch_iso_q1_to_gll(double *dst, const double *src, int n1x, int n1y, int n1z,
                                  int p, const IsoQ1Map &m, hipStream_t s);
  hipError_t launch_dot(const double *a, const double *b, int64_t n, double *partials,
                        double *out, unsigned *counter, hipStream_t s);

  // Batched (segmented) vector kernels, batch_kernels.hip. Solve s owns entries
  // [off, off+n) of the concatenated vectors and workgroups [blk0, blk0+nblk).
  struct BatchSeg
  {
    int64_t off, n;
    int blk0, nblk;
  };
  constexpr int kBatchThreads = 256;
  template <typename Number>
  hipError_t launch_batch_cg_init(const double *b, const Number *ax, Number *r, Number *p,
                                  const Number *inv, const BatchSeg *seg, int nsolve, int nblocks,
                                  double *partials, SolverScalars *scal, const double *rel_tol,
                                  double abs_tol, int max_it, const int *active, hipStream_t s);
  template <typename Number>
  hipError_t launch_batch_cg_update(Number *x, Number *r, const Number *p, const Number *q,
                                    const Number *inv, const BatchSeg *seg, int nsolve,
                                    int nblocks, double *partials, SolverScalars *scal,
                                    hipStream_t s);
  template <typename Number>
  hipError_t launch_batch_cg_direction(Number *p, const Number *r, const Number *inv,
                                       const BatchSeg *seg, int nsolve, int nblocks,
                                       const SolverScalars *scal, hipStream_t s);
  hipError_t launch_batch_residual(double *r, const double *b, const double *ax,
                                   const BatchSeg *seg, int nsolve, int nblocks, double *partials,
                                   double *res, hipStream_t s);

  // Per-member defect correction of the batched mixed-precision solve, run on
  // the device so that every member moves through its own outer steps:
  // Outer (r = b - A64 x, stop test) -> InnerStart (r32, d = 0, PCG start) ->
  // Inner (FP32 PCG chunks until its own stop) -> Outer (x += d) ... -> Done.
  enum BatchPhase
  {
    kPhaseOuter = 0,
    kPhaseInnerStart = 1,
    kPhaseInner = 2,
    kPhaseDone = 3
  };
  struct BatchOuter
  {
    double tol, res0, res, inner_rel;
    unsigned long long t_start, t_done; // device wall clock (wall_clock64 ticks)
    int phase, outer, inner_total, status;
    unsigned counter;
    unsigned pad[3];
  };
  // Every member to Outer (mixed) with its inner PCG stopped; t_start stamped.
  // ogate[s].done = 1 keeps a member out of the FP64 residual vmult.
  hipError_t launch_batch_begin(BatchOuter *st, SolverScalars *scal, SolverScalars *ogate,
                                int nsolve, hipStream_t s);
  // Outer members: r = b - ax, ||r||, then Done or InnerStart.
  hipError_t launch_batch_outer_step(double *r, const double *b, const double *ax,
                                     const BatchSeg *seg, int nsolve, int nblocks,
                                     double *partials, BatchOuter *st, SolverScalars *ogate,
                                     double abs_tol, double rel_tol, double inner_rel, int max_it,
                                     hipStream_t s);
  // InnerStart members: r32 = r, d = 0, p = D r32, their FP32 PCG scalars; -> Inner.
  hipError_t launch_batch_inner_start(const double *r, float *r32, float *p32, float *x32,
                                      const float *inv, const BatchSeg *seg, int nsolve,
                                      int nblocks, double *partials, BatchOuter *st,
                                      SolverScalars *scal, int max_it, hipStream_t s);
  // Inner members whose PCG stopped: x += d, iterations counted; -> Outer.
  hipError_t launch_batch_inner_finish(double *x, const float *x32, const BatchSeg *seg,
                                       int nsolve, int nblocks, BatchOuter *st,
                                       const SolverScalars *scal, SolverScalars *ogate,
                                       hipStream_t s);

  // GMRES building blocks (classical Gram-Schmidt over a basis of at most
  // kMaxBasis vectors stored with leading dimension ld). A non-null `stop`
  // makes a launch do nothing when *stop != 0 (a restart cycle that has
  // stopped), uniformly over the grid: the flag is not written while it runs.
  constexpr int kMaxBasis = 32;
  template <typename Number>
  hipError_t launch_multi_dot(const Number *V, int64_t ld, int k, const Number *w, int64_t n,
                              double *partials, double *out, unsigned *counter, hipStream_t s,
                              const int *stop = nullptr);
  template <typename Number>
  hipError_t launch_multi_axpy(Number *w, const Number *V, int64_t ld, int k, const double *h,
                               int64_t n, double *partials, double *nrm2, unsigned *counter,
                               hipStream_t s, const int *stop = nullptr);
  // w -= sum_c h[c] V_c and out[c] = V_c . w (the updated w) in one sweep
  template <typename Number>
  hipError_t launch_multi_axpy_dot(Number *w, const Number *V, int64_t ld, int k, const double *h,
                                   int64_t n, double *partials, double *out, unsigned *counter,
                                   hipStream_t s, const int *stop = nullptr);
  // v = w / sqrt(nrm2) and z = d .* v in one pass
  template <typename Number>
  hipError_t launch_scale_precondition(Number *v, Number *z, const Number *w, const Number *d,
                                       const double *nrm2, int64_t n, hipStream_t s,
                                       const int *stop = nullptr);
  template <typename Number>
  hipError_t launch_scale(Number *v, const Number *w, const double *nrm2, int64_t n, hipStream_t s);
  template <typename Number>
  hipError_t launch_diag_scale(Number *z, const Number *v, const Number *d, int64_t n, hipStream_t s);
  // x += D V y over the first k basis vectors; a non-null kdev supplies k from
  // device memory (the cycle's basis size)
  template <typename Number, typename XNumber>
  hipError_t launch_gmres_update(XNumber *x, const Number *V, int64_t ld, int k, const double *y,
                                 const Number *d, int64_t n, hipStream_t s,
                                 const int *kdev = nullptr);

  // Device state of right-preconditioned GMRES(m): the Hessenberg matrix, the
  // Givens rotations and the least-squares right-hand side live in device
  // memory, so a whole restart cycle runs without the host. sc.done is the
  // cycle's stop flag (the vmult and basis kernels honour it).
  struct GmresState
  {
    SolverScalars sc;
    double tol;        // max(abs_tol, rel_tol ||r_0||)
    double res;        // |g_{j+1}| after the last step
    int it, max_it;    // iterations so far, limit
    int k;             // basis vectors of the cycle's update
    int inner_conv;    // the cycle stopped on the residual estimate
    double cs[kMaxBasis], sn[kMaxBasis], g[kMaxBasis + 1], y[kMaxBasis];
    double hm[kMaxBasis * kMaxBasis]; // Hessenberg column j of row l at l * kMaxBasis + j
  };
  // it = 0, tol, max_it (once per solve)
  hipError_t launch_gmres_init(GmresState *st, double tol, int max_it, hipStream_t s);
  // a cycle starts from the residual norm rs->res: g = (res, 0, ...), stop cleared
  hipError_t launch_gmres_start(GmresState *st, const SolverScalars *rs, int restart,
                                hipStream_t s);
  // step j: Hessenberg column from the two Gram-Schmidt passes (h1 + h2) and
  // ||w|| = sqrt(*nrm2), the earlier rotations, a new rotation, the residual
  // estimate |g_{j+1}| and the stop test (converged, limit, breakdown, NaN)
  hipError_t launch_gmres_givens(GmresState *st, const double *h1, const double *h2,
                                 const double *nrm2, int j, int restart, hipStream_t s);
  // y from the k x k triangular system (back substitution)
  hipError_t launch_gmres_solve(GmresState *st, hipStream_t s);
} // namespace mpdg
