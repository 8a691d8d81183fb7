// C ABI of the heat-operator engine (include/mpdg_heat.h): handle set-up
// (the MatrixFree "plan" of a box mesh), vmult, inverse diagonal and the
// device-resident Krylov drivers.
#include "../../include/mpdg_heat.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "tables.hpp"

using namespace mpdg;

namespace
{
  thread_local std::string g_last_error;

  mpdg_status fail(mpdg_status s, const std::string &msg)
  {
    g_last_error = msg;
    return s;
  }

#define MPDG_HIP(expr)                                                                          \
  do                                                                                            \
    {                                                                                           \
      const hipError_t e_ = (expr);                                                             \
      if (e_ != hipSuccess)                                                                     \
        return fail(MPDG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    }                                                                                           \
  while (0)

  template <typename T>
  struct DevBuf
  {
    T *ptr = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count)
    {
      if (count <= n && ptr)
        return hipSuccess;
      if (ptr)
        (void)hipFree(ptr);
      ptr = nullptr;
      n = 0;
      const hipError_t e = hipMalloc(&ptr, count * sizeof(T) + 64);
      if (e == hipSuccess)
        n = count;
      return e;
    }
    void release()
    {
      if (ptr)
        (void)hipFree(ptr);
      ptr = nullptr;
      n = 0;
    }
  };
} // namespace

struct mpdg_heat
{
  mpdg_heat_desc desc{};
  hipStream_t stream = nullptr;
  int device = 0;
  Tables1D t1d;
  int64_t N = 0, N_free = 0;
  double h[3]{};
  double dt = 0.0;
  bool dt_set = false;
  mpdg_tuning tune[2]{}; // vmult launch tuning per precision (MPDG_FP64, MPDG_FP32)

  CellParams<double> prm64{};
  CellParams<float> prm32{};
  // ordered vmult schedule: device state and per-item flags per precision
  DevBuf<OrderState> order[2];
  DevBuf<unsigned> order_flags[2];
  // sticky report of a timed-out ordered wait for calls that do not
  // synchronise (the vmult entry points): each such launch is followed by an
  // asynchronous copy of the error bit into pinned memory and an event; the
  // next call on the handle reads the copy once its event has completed
  // slab schedule of the colour passes: the second stream and its events
  SlabStreams slabs{};
  unsigned *host_order_err = nullptr; // pinned, one per precision
  hipEvent_t order_ev[2]{};
  bool order_pending[2]{};
  bool order_failed = false;
  VmultTables<double> vt64{};
  VmultTables<float> vt32{};
  DiagTables<double> dt64{};

  DevBuf<double> sin_tab64, sin_tab32_src; // sin(2 pi x), sin^2(pi x) per direction (double)
  DevBuf<float> sin_tab32;
  DevBuf<double> heav64;
  DevBuf<float> heav32;
  bool has_heaviside = false;
  DevBuf<double> em64;  // element-matrix form: the cell matrix packed per lane (EmShape)
  DevBuf<float> em32;
  std::vector<double> em_host64; // host copies the device tables are uploaded from
  std::vector<float> em_host32;
  DevBuf<double> vel64; // nodal velocity, 3 N (MPDG_VELOCITY_NODAL)
  bool has_velocity = false;

  // Robin faces and the temperature the tangent is linearised at (a caller's
  // device vector, not a copy: the reference operator holds a reference)
  RobinParams robin{};
  RobinParams robin_rhs{}; // the residual's faces: Robin and Neumann
  const double *temperature = nullptr;
  // old temperature the tangent of a temperature-dependent material reads
  // (the operator's reference to temperature_old), also a caller's vector
  const double *temperature_old = nullptr;
  DevBuf<double> robin_rows; // one row of (p+1)^2 face integrals per Robin face cell

  // solver workspace
  DevBuf<double> r, p, q, inv_diag, partials;
  DevBuf<double> p2; // second direction of the paired CG iterations
  // SolverControl::enable_history_data: a caller's device buffer of
  // history_len entries that the FP64 CG fills with ||r_k|| (k = 0 initial)
  double *history = nullptr;
  int history_len = 0;
  DevBuf<float> r32, p32, q32, x32, inv32;
  DevBuf<uint16_t> inv16; // the FP32 inner CG's inverse diagonal in bfloat16
  DevBuf<double> V64, gm;  // GMRES basis (FP64) and small coefficient block
  DevBuf<float> V32;       // GMRES basis of the FP32 inner solves
  DevBuf<SolverScalars> scal;
  SolverScalars *host_scal = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // Krylov iteration chunks captured as hipGraphs: a capture stream of the
  // handle's own (the caller's stream may be the legacy NULL stream, which
  // cannot be captured), fork / join events, one cached graph per precision.
  // state_gen changes with every operator-parameter update (fill_params), so a
  // graph never replays stale kernel arguments.
  struct ChunkGraph
  {
    hipGraphExec_t exec = nullptr;
    uint64_t gen = 0;
    int steps = 0;
    const void *key[16]{};
  };
  bool use_graphs = true;
  uint64_t state_gen = 1;
  hipStream_t gstream = nullptr;
  hipEvent_t gfork = nullptr, gjoin = nullptr;
  ChunkGraph pcg_graph[2];
  std::vector<ChunkGraph> gmres_graphs[2]; // per precision: the cycle's step chunks, then its end
  DevBuf<GmresState> gst;            // device GMRES state
  GmresState *host_gst = nullptr;    // pinned copy of its scalar head

  // Newton time step: residual, increment, reduction scratch and results
  DevBuf<double> nrhs, ndu, npart, nres;
  DevBuf<unsigned> nticket;
  double *host_nres = nullptr;

  ~mpdg_heat()
  {
    for (auto *b : {&sin_tab64, &sin_tab32_src, &heav64, &vel64, &r, &p, &p2, &q, &inv_diag,
                    &partials, &V64, &gm, &robin_rows})
      b->release();
    for (auto *b : {&sin_tab32, &heav32, &r32, &p32, &q32, &x32, &inv32, &V32})
      b->release();
    inv16.release();
    for (auto *b : {&nrhs, &ndu, &npart, &nres})
      b->release();
    nticket.release();
    for (int k = 0; k < 2; ++k)
      {
        order[k].release();
        order_flags[k].release();
      }
    if (host_nres)
      (void)hipHostFree(host_nres);
    if (host_order_err)
      (void)hipHostFree(host_order_err);
    for (hipEvent_t ev : {slabs.fork, slabs.join, slabs.mid})
      if (ev)
        (void)hipEventDestroy(ev);
    if (slabs.side)
      (void)hipStreamDestroy(slabs.side);
    for (hipEvent_t ev : order_ev)
      if (ev)
        (void)hipEventDestroy(ev);
    scal.release();
    if (host_scal)
      (void)hipHostFree(host_scal);
    if (ev0)
      (void)hipEventDestroy(ev0);
    if (ev1)
      (void)hipEventDestroy(ev1);
    for (auto &g : pcg_graph)
      if (g.exec)
        (void)hipGraphExecDestroy(g.exec);
    for (auto &v : gmres_graphs)
      for (auto &g : v)
        if (g.exec)
          (void)hipGraphExecDestroy(g.exec);
    gst.release();
    if (host_gst)
      (void)hipHostFree(host_gst);
    if (gfork)
      (void)hipEventDestroy(gfork);
    if (gjoin)
      (void)hipEventDestroy(gjoin);
    if (gstream)
      (void)hipStreamDestroy(gstream);
  }
};

namespace
{
  // Slab schedule resources, made when a tuning first asks for it: a second
  // stream (default priority) and its events.
  hipError_t ensure_slabs(mpdg_heat *H)
  {
    if (H->slabs.side)
      return hipSuccess;
    hipError_t e = hipSuccess;
    for (hipEvent_t *ev : {&H->slabs.fork, &H->slabs.join, &H->slabs.mid})
      if (e == hipSuccess && !*ev)
        e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e == hipSuccess)
      e = hipStreamCreateWithFlags(&H->slabs.side, hipStreamNonBlocking);
    return e;
  }

  // Default launch tuning per degree and precision: the autotuner's picks on
  // MI355X at the benchmark-like sizes (unit cube, sine k, default materials;
  // measured GDoF/s in the comments); mpdg_heat_autotune refines it for a
  // given operator.
  mpdg_tuning default_tuning(int degree, bool fp32, int64_t cells)
  {
    mpdg_tuning t{};
    t.variant = MPDG_VMULT_LINE;
    t.tile_log2 = 4;
    // the cell form needs about a million cells to fill the chip in each of
    // its eight colour passes (64 cells per wavefront). The lean FP32 line
    // form (128-cell row segments) beats it only where the rows fill whole
    // segments (128^3: 41.8 vs 37.7 GDoF/s; 160^3: 31.5 vs 45.0; 200^3,
    // constant k: 48.8 vs 56.3), so the autotuner decides that case.
    if (degree == 1 && cells >= (int64_t(1) << 20))
      {
        t.variant = MPDG_VMULT_CELL;
        t.tile_log2 = 5;
        t.z_layers = 2;
        t.blocks_per_cu = 0;
        return t;
      }
    switch (degree)
      {
        case 1: t.z_layers = fp32 ? 4 : 2; t.blocks_per_cu = fp32 ? 4 : 0; break;
        case 2:
          // FP32 (two-slot shape): one-layer chunks measured 72.6 against 68.3
          // GDoF/s at 96^3, two-layer ones 108.6 against 100.1 at 192^3
          t.z_layers = fp32 ? (cells >= (int64_t(1) << 22) ? 2 : 1) : 2;
          t.blocks_per_cu = fp32 ? 0 : 4;
          break;
        case 3: t.z_layers = fp32 ? 2 : 4; t.blocks_per_cu = fp32 ? 8 : 4; break;
        case 4: t.z_layers = fp32 ? 2 : 4; t.blocks_per_cu = fp32 ? 0 : 4; break;
        case 5:
          // FP32: the spread FMA form measured 154 against 111 GDoF/s for the
          // matrix-core form (64^3, sine k)
          t.z_layers = fp32 ? 2 : 4;
          t.blocks_per_cu = 0;
          break;
        default: t.z_layers = 2; t.blocks_per_cu = 4; break;
      }
    // Slab schedule (two streams, two-layer chunks) on large meshes. FP64
    // degrees 2..4 from 2^18 cells: 1.05x (Q4 64^3), 1.09x (Q3 96^3), 1.15x
    // (Q2 96^3) the passes; degrees 5..6 from 2^16 cells: Q5 64^3 89.1 vs
    // 83.9, Q6 48^3 83.5 vs 70.3 (8 workgroups per CU). FP32 degree 3 (96^3:
    // 118.2 vs 103.5), 5 (64^3: 172.0 vs 156.2) and 6 (48^3, one-layer chunks:
    // 158.1 vs 144.3); FP32 degrees 2 and 4 did not gain (Q4 126 vs 132).
    const bool big = cells >= (int64_t(1) << (degree >= 5 ? 16 : 18));
    if (big && !fp32 && degree >= 2)
      {
        t.schedule = MPDG_SCHEDULE_SLABS;
        t.z_layers = 2;
        t.blocks_per_cu = degree == 6 ? 8 : 0;
      }
    if (big && fp32 && (degree == 3 || degree >= 5))
      {
        t.schedule = MPDG_SCHEDULE_SLABS;
        t.z_layers = degree == 6 ? 1 : 2;
        t.blocks_per_cu = degree == 3 ? 0 : 8;
      }
    return t;
  }

  bool tuning_valid(const mpdg_tuning &t, int degree, int precision)
  {
    return t.z_layers >= 1 && t.blocks_per_cu >= 0 && t.blocks_per_cu <= kMaxResidentPerCu &&
           (t.schedule == MPDG_SCHEDULE_PASSES || t.schedule == MPDG_SCHEDULE_ORDERED ||
            t.schedule == MPDG_SCHEDULE_SLABS) &&
           (t.variant == MPDG_VMULT_LINE ||
            (t.variant == MPDG_VMULT_CELL && degree <= kCellMaxDegree && t.tile_log2 >= 3 &&
             t.tile_log2 <= 6) ||
            (t.variant == MPDG_VMULT_LINE_MFMA && degree >= kMfmaMinDegree &&
             precision == MPDG_FP32) ||
            (t.variant == MPDG_VMULT_CELL_MFMA && degree <= kEmMaxDegree && t.tile_log2 >= 1 &&
             t.tile_log2 <= 4));
  }

  // Robin faces of a descriptor (the tangent's and diagonal's; with_neumann:
  // the residual's, the Neumann faces with their -(q . n) included): faces
  // ascending, each face's node lattice a contiguous thread range (tangential
  // directions e1 < e2, e1 fastest).
  RobinParams robin_params(const mpdg_heat_desc &d, const double *h, bool with_neumann = false)
  {
    constexpr double kStefanBoltzmann = 5.670374419e-8; // physical_constants.hpp:6
    RobinParams rp{};
    rp.conv = d.convection_faces & 0x3Fu;
    rp.rad = d.radiation_faces & 0x3Fu;
    rp.neumann = with_neumann ? d.neumann_faces & 0x3Fu : 0u;
    rp.robin = rp.conv | rp.rad | rp.neumann;
    for (int f = 0; f < 6; ++f)
      rp.qn[f] = (f & 1 ? 1.0 : -1.0) * d.neumann_flux[f][f >> 1]; // q . n, n outward
    rp.dirichlet = d.dirichlet_faces & 0x3Fu;
    for (int k = 0; k < 3; ++k)
      {
        rp.n[k] = d.n_cells[k];
        rp.n1[k] = d.n_cells[k] * d.degree + 1;
        rp.h[k] = h[k];
      }
    rp.alpha = d.convection_coefficient;
    rp.t_inf_conv = d.convection_temperature_infinity;
    rp.eps_sigma = d.emissivity * kStefanBoltzmann;
    rp.t_inf_rad = d.radiation_temperature_infinity;
    rp.node_off[0] = 0;
    rp.cell_off[0] = 0;
    for (int f = 0; f < 6; ++f)
      if ((rp.robin >> f) & 1u)
        {
          const int dd = f >> 1;
          const int e1 = dd == 0 ? 1 : 0, e2 = dd == 2 ? 1 : 2;
          rp.face[rp.nfaces] = f;
          rp.node_off[rp.nfaces + 1] = rp.node_off[rp.nfaces] + (int64_t)rp.n1[e1] * rp.n1[e2];
          rp.cell_off[rp.nfaces + 1] = rp.cell_off[rp.nfaces] + (int64_t)rp.n[e1] * rp.n[e2];
          ++rp.nfaces;
        }
    return rp;
  }

  // Analytic coefficient tables at the Gauss coordinates of each cell row:
  // sin(2 pi x_d) per direction d (separable conductivity), then the separable
  // in-plane velocity, u_x = X_x(x) Y_x(y) and u_y = X_y(x) Y_y(y), as
  // [X_x | Y_x | X_y | Y_y]:
  //   vortex   u = (-sin^2(pi x) sin(2 pi y), sin(2 pi x) sin^2(pi y), w0);
  //   constant u = (v0, v1, v2).
  std::vector<double> coefficient_tables(const mpdg_heat_desc &d, const Tables1D &t,
                                         const double *h)
  {
    const int n = t.n;
    auto coords = [&](int k, auto f) {
      std::vector<double> out;
      for (int c = 0; c < d.n_cells[k]; ++c)
        for (int q = 0; q < n; ++q)
          out.push_back(f(d.box_lo[k] + (c + t.xq[q]) * h[k]));
      return out;
    };
    const double two_pi = 2.0 * M_PI;
    auto sin2pi = [&](double x) { return std::sin(two_pi * x); };
    auto sinpi2 = [&](double x) {
      const double s = std::sin(M_PI * x);
      return s * s;
    };
    std::vector<double> tab;
    for (int k = 0; k < 3; ++k)
      {
        const std::vector<double> v = coords(k, sin2pi);
        tab.insert(tab.end(), v.begin(), v.end());
      }
    std::vector<double> X[2], Y[2];
    if (d.velocity == MPDG_VELOCITY_VORTEX)
      {
        X[0] = coords(0, [&](double x) { return -sinpi2(x); });
        Y[0] = coords(1, sin2pi);
        X[1] = coords(0, sin2pi);
        Y[1] = coords(1, sinpi2);
      }
    else
      {
        const bool cst = d.velocity == MPDG_VELOCITY_CONSTANT;
        const double v0 = cst ? d.velocity_value[0] : 0.0, v1 = cst ? d.velocity_value[1] : 0.0;
        X[0] = coords(0, [&](double) { return v0; });
        Y[0] = coords(1, [&](double) { return 1.0; });
        X[1] = coords(0, [&](double) { return 1.0; });
        Y[1] = coords(1, [&](double) { return v1; });
      }
    for (int c = 0; c < 2; ++c)
      {
        tab.insert(tab.end(), X[c].begin(), X[c].end());
        tab.insert(tab.end(), Y[c].begin(), Y[c].end());
      }
    return tab;
  }

  bool em_eligible(const mpdg_heat_desc &d);

  template <typename Number>
  void fill_params(mpdg_heat *H, CellParams<Number> &prm, const Number *tab)
  {
    ++H->state_gen; // kernel arguments change: captured graphs are stale
    const mpdg_heat_desc &d = H->desc;
    const int n = H->t1d.n;
    prm.nx = d.n_cells[0];
    prm.ny = d.n_cells[1];
    prm.nz = d.n_cells[2];
    prm.n1x = d.n_cells[0] * d.degree + 1;
    prm.n1y = d.n_cells[1] * d.degree + 1;
    prm.n1z = d.n_cells[2] * d.degree + 1;
    prm.faces = d.dirichlet_faces;
    const double vol = H->h[0] * H->h[1] * H->h[2];
    const double g_rhocp = d.gas.density * d.gas.specific_heat_capacity;
    const double l_rhocp = d.liquid.density * d.liquid.specific_heat_capacity;
    const double dt_inv = H->dt_set ? 1.0 / H->dt : 0.0; // OperatorBase::reset_time_increment
    prm.mass = (Number)(dt_inv * g_rhocp * vol);
    prm.dtinv_vol = (Number)(dt_inv * vol);
    prm.vol = (Number)vol;
    for (int k = 0; k < 3; ++k)
      {
        prm.kf[k] = (Number)(vol / (H->h[k] * H->h[k]));
        prm.vf[k] = (Number)(vol / H->h[k]);
      }
    prm.k0 = (Number)d.gas.thermal_conductivity;
    prm.kamp = (Number)(d.conductivity_field == MPDG_CONDUCTIVITY_SEPARABLE_SINE ?
                          d.conductivity_amplitude :
                          0.0);
    prm.rho_cp0 = (Number)g_rhocp;
    prm.w0 = (Number)(d.velocity == MPDG_VELOCITY_CONSTANT ? d.velocity_value[2] : d.velocity_w0);
    prm.g_rho = (Number)d.gas.density;
    prm.g_cp = (Number)d.gas.specific_heat_capacity;
    prm.g_k = (Number)d.gas.thermal_conductivity;
    prm.l_rho = (Number)d.liquid.density;
    prm.l_cp = (Number)d.liquid.specific_heat_capacity;
    prm.l_k = (Number)d.liquid.thermal_conductivity;
    prm.g_rhocp = (Number)g_rhocp;
    prm.l_rhocp = (Number)l_rhocp;
    prm.sharp = d.two_phase_sharp;
    prm.volume_specific = d.volume_specific_heat;
    prm.tmat = (d.material == MPDG_MATERIAL_LIQUID_SOLID || d.material == MPDG_MATERIAL_GAS_LIQUID_SOLID) ?
                 d.material :
                 0;
    prm.sl_sharp = d.solid_liquid_sharp;
    prm.s_rho = (Number)d.solid.density;
    prm.s_cp = (Number)d.solid.specific_heat_capacity;
    prm.s_k = (Number)d.solid.thermal_conductivity;
    prm.s_rhocp = (Number)(d.solid.density * d.solid.specific_heat_capacity);
    prm.solidus = (Number)d.solidus_temperature;
    prm.liquidus = (Number)d.liquidus_temperature;
    // Material::inv_mushy_interval (material.hpp:648-654); unused by the sharp law
    prm.inv_mushy = (Number)(d.liquidus_temperature > d.solidus_temperature ?
                               1.0 / (d.liquidus_temperature - d.solidus_temperature) :
                               0.0);
    prm.ac_type = d.apparent_capacity_type;
    prm.latent = (Number)(prm.tmat != 0 ? d.latent_heat_of_fusion : 0.0);
    const mpdg_tuning &tu = H->tune[sizeof(Number) == 4 ? MPDG_FP32 : MPDG_FP64];
    prm.zlayers = std::max(1, std::min(tu.z_layers, d.n_cells[2]));
    prm.blocks_per_cu = tu.blocks_per_cu;
    prm.variant = (tu.variant == MPDG_VMULT_CELL && d.degree <= kCellMaxDegree) ? kVariantCell :
                  (tu.variant == MPDG_VMULT_LINE_MFMA && d.degree >= kMfmaMinDegree &&
                   sizeof(Number) == 4) ? kVariantLineMfma :
                  (tu.variant == MPDG_VMULT_CELL_MFMA && em_eligible(d)) ? kVariantCellMfma :
                                                                            kVariantLine;
    prm.tile_log2 = tu.tile_log2;
    const int pi = sizeof(Number) == 4 ? MPDG_FP32 : MPDG_FP64;
    const bool cellish = prm.variant == kVariantCell || prm.variant == kVariantCellMfma;
    const bool ordered = tu.schedule == MPDG_SCHEDULE_ORDERED && !cellish &&
                         H->order[pi].ptr != nullptr;
    prm.order = ordered ? H->order[pi].ptr : nullptr;
    prm.order_flags = ordered ? H->order_flags[pi].ptr : nullptr;
    prm.slabs = tu.schedule == MPDG_SCHEDULE_SLABS && !cellish ? 2 : 1;
    // per-direction tables (see coefficient_tables): [sin2pi dir 0 | dir 1 | dir 2 |
    // u_x x-factor | u_x y-factor | u_y x-factor | u_y y-factor]
    size_t off = 0;
    for (int k = 0; k < 3; ++k)
      {
        prm.sin2pi[k] = tab + off;
        off += (size_t)d.n_cells[k] * n;
      }
    for (int c = 0; c < 2; ++c)
      {
        prm.vel_x[c] = tab + off;
        off += (size_t)d.n_cells[0] * n;
        prm.vel_y[c] = tab + off;
        off += (size_t)d.n_cells[1] * n;
      }
  }

  bool temperature_dependent(const mpdg_heat_desc &d)
  {
    return d.material == MPDG_MATERIAL_LIQUID_SOLID || d.material == MPDG_MATERIAL_GAS_LIQUID_SOLID;
  }

  bool uses_heaviside(const mpdg_heat_desc &d)
  {
    return d.material == MPDG_MATERIAL_GAS_LIQUID || d.material == MPDG_MATERIAL_GAS_LIQUID_SOLID;
  }

  bool nodal_velocity(const mpdg_heat_desc &d) { return d.velocity == MPDG_VELOCITY_NODAL; }

  // The cell loop evaluates its coefficients per quadrature point from nodal
  // fields (material_kernels.hip, FP64): temperature-dependent materials and
  // nodal velocity fields.
  bool generic_cells(const mpdg_heat_desc &d) { return temperature_dependent(d) || nodal_velocity(d); }

  // The element-matrix form needs the same cell matrix in every cell: the
  // constant material without a conductivity field, no or a constant velocity.
  bool em_eligible(const mpdg_heat_desc &d)
  {
    return d.degree <= kEmMaxDegree && d.material == MPDG_MATERIAL_GAS &&
           d.conductivity_field == MPDG_CONDUCTIVITY_MATERIAL &&
           (d.velocity == MPDG_VELOCITY_NONE || d.velocity == MPDG_VELOCITY_CONSTANT);
  }

  // Cell matrix of the tangent of a constant material on an affine box cell,
  // test node m (rows) x trial node n, nodes lexicographic (x fastest):
  //   rho_cp/dt |J| M x M x M + k (|J|/hx^2 K x M x M + ...) + rho_cp (vx |J|/hx C x M x M + ...)
  // with the 1D unit-interval matrices of the operator's QGauss(p+1) rule
  // M_ij = sum_q w phi_i phi_j, K_ij = sum_q w phi_i' phi_j', C_ij = sum_q w phi_i phi_j'
  // (the terms of tangent_local_cell_operation, heat_diffuse_operator.cpp:1029-1042;
  // the rule integrates them exactly). Packed per lane as the 16x16x4 MFMA A
  // operands: entry (rb * T + ks) * 64 + lane holds A[node_r][node_c], column
  // node 4 ks + lane / 16 (0 past the last node). The product of A operand row
  // i = lane % 16 lands (measured on gfx950) in lane group i / 4, register
  // i % 4 for v_mfma_f32_16x16x4f32, and in lane group i % 4, register i / 4
  // for v_mfma_f64_16x16x4f64; row i of block rb is the node of that slot
  // (group g, register r: node g + 4 (4 rb + r)).
  std::vector<double> em_table(const mpdg_heat *H, bool f64)
  {
    const mpdg_heat_desc &d = H->desc;
    const Tables1D &t = H->t1d;
    const int n = t.n, nn = n * n * n;
    double M[kMaxN][kMaxN]{}, K[kMaxN][kMaxN]{}, C[kMaxN][kMaxN]{};
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int q = 0; q < n; ++q)
          {
            M[i][j] += t.wq[q] * t.S[q][i] * t.S[q][j];
            K[i][j] += t.wq[q] * t.D[q][i] * t.D[q][j];
            C[i][j] += t.wq[q] * t.S[q][i] * t.D[q][j];
          }
    const double vol = H->h[0] * H->h[1] * H->h[2];
    const double rhocp = d.gas.density * d.gas.specific_heat_capacity;
    const double mass = (H->dt_set ? 1.0 / H->dt : 0.0) * rhocp * vol;
    const double k = d.gas.thermal_conductivity;
    const bool cst = d.velocity == MPDG_VELOCITY_CONSTANT;
    std::vector<double> A((size_t)nn * nn);
    for (int m = 0; m < nn; ++m)
      {
        const int a = m % n, b = (m / n) % n, c = m / (n * n);
        for (int q = 0; q < nn; ++q)
          {
            const int a2 = q % n, b2 = (q / n) % n, c2 = q / (n * n);
            const double mx = M[a][a2], my = M[b][b2], mz = M[c][c2];
            double v = mass * mx * my * mz;
            v += k * (vol / (H->h[0] * H->h[0]) * K[a][a2] * my * mz +
                      vol / (H->h[1] * H->h[1]) * mx * K[b][b2] * mz +
                      vol / (H->h[2] * H->h[2]) * mx * my * K[c][c2]);
            if (cst)
              v += rhocp * (d.velocity_value[0] * vol / H->h[0] * C[a][a2] * my * mz +
                            d.velocity_value[1] * vol / H->h[1] * mx * C[b][b2] * mz +
                            d.velocity_value[2] * vol / H->h[2] * mx * my * C[c][c2]);
            A[(size_t)m * nn + q] = v;
          }
      }
    const int T = (nn + 3) / 4, RB = (T + 3) / 4;
    std::vector<double> packed((size_t)RB * T * 64, 0.0);
    for (int rb = 0; rb < RB; ++rb)
      for (int ks = 0; ks < T; ++ks)
        for (int lane = 0; lane < 64; ++lane)
          {
            const int i = lane % 16, kk = lane / 16;
            const int g = f64 ? i % 4 : i / 4, r = f64 ? i / 4 : i % 4;
            const int tr = 4 * rb + r;
            const int mr = g + 4 * tr, mc = 4 * ks + kk;
            if (tr < T && mr < nn && mc < nn)
              packed[((size_t)rb * T + ks) * 64 + lane] = A[(size_t)mr * nn + mc];
          }
    return packed;
  }

  // (Re)builds the element-matrix tables of an eligible operator (the mass
  // term follows dt) and points both precisions' parameters at them.
  mpdg_status update_em(mpdg_heat *H)
  {
    H->prm64.em = nullptr;
    H->prm32.em = nullptr;
    if (!em_eligible(H->desc))
      return MPDG_OK;
    H->em_host64 = em_table(H, true);
    const std::vector<double> t32 = em_table(H, false);
    H->em_host32.assign(t32.begin(), t32.end());
    const size_t cnt = H->em_host64.size();
    MPDG_HIP(H->em64.alloc(cnt));
    MPDG_HIP(H->em32.alloc(cnt));
    MPDG_HIP(hipMemcpyAsync(H->em64.ptr, H->em_host64.data(), cnt * sizeof(double),
                            hipMemcpyHostToDevice, H->stream));
    MPDG_HIP(hipMemcpyAsync(H->em32.ptr, H->em_host32.data(), cnt * sizeof(float),
                            hipMemcpyHostToDevice, H->stream));
    H->prm64.em = H->em64.ptr;
    H->prm32.em = H->em32.ptr;
    return MPDG_OK;
  }

  int velocity_mode(const mpdg_heat_desc &d)
  {
    return nodal_velocity(d) ? kVelNodal : (d.velocity != MPDG_VELOCITY_NONE ? kVelTables : kVelNone);
  }

  int coef_mode(const mpdg_heat *H)
  {
    if (temperature_dependent(H->desc))
      return kCoefTemperature;
    if (H->desc.material == MPDG_MATERIAL_GAS_LIQUID)
      return kCoefHeaviside;
    return H->desc.conductivity_field == MPDG_CONDUCTIVITY_SEPARABLE_SINE ? kCoefSine :
                                                                           kCoefConstant;
  }

  mpdg_status check_ready(const mpdg_heat *H)
  {
    if (!H)
      return fail(MPDG_ERR_ARGUMENT, "null handle");
    if (!H->dt_set)
      return fail(MPDG_ERR_ZERO_TIME_INCREMENT,
                  "time increment not set: call mpdg_heat_set_dt (reset_time_increment) first");
    if (uses_heaviside(H->desc) && !H->has_heaviside)
      return fail(MPDG_ERR_ARGUMENT, "gas_liquid(_solid) material needs mpdg_heat_set_heaviside");
    if (nodal_velocity(H->desc) && !H->has_velocity)
      return fail(MPDG_ERR_ARGUMENT, "a nodal velocity field needs mpdg_heat_set_velocity");
    return MPDG_OK;
  }

  mpdg_status poll_order_state(mpdg_heat *H);

  // check_ready plus the sticky ordered-schedule report of earlier calls
  mpdg_status check_call(mpdg_heat *H)
  {
    if (mpdg_status s = check_ready(H))
      return s;
    return poll_order_state(H);
  }

  // The tangent of radiating faces depends on the registered temperature.
  mpdg_status check_tangent(const mpdg_heat *H)
  {
    if (H->robin.rad && !H->temperature)
      return fail(MPDG_ERR_ARGUMENT,
                  "radiation faces need the linearisation temperature: mpdg_heat_set_temperature");
    if (temperature_dependent(H->desc) && (!H->temperature || !H->temperature_old))
      return fail(MPDG_ERR_ARGUMENT,
                  "a temperature-dependent material needs the linearisation temperature and the "
                  "old temperature: mpdg_heat_set_temperature / mpdg_heat_set_temperature_old");
    return MPDG_OK;
  }

  // After a synchronising call: a bounded wait of the ordered vmult schedule
  // that gave up (only a defect can cause it) leaves its error bit set, and the
  // results since then are not to be trusted.
  mpdg_status order_failure()
  {
    return fail(MPDG_ERR_HIP, "ordered vmult schedule: a wait for a neighbouring work item "
                              "timed out; results are unreliable");
  }

  // Non-blocking half of the report: a completed copy of the error bit that
  // shows the bit set makes the handle's failure sticky; every later call on
  // the handle returns it.
  mpdg_status poll_order_state(mpdg_heat *H)
  {
    for (int k = 0; k < 2; ++k)
      if (H->order_pending[k] && hipEventQuery(H->order_ev[k]) == hipSuccess)
        {
          H->order_pending[k] = false;
          if (H->host_order_err[k])
            H->order_failed = true;
        }
    return H->order_failed ? order_failure() : MPDG_OK;
  }

  // After an ordered launch of a non-synchronising call: queue the copy of the
  // error bit that poll_order_state reads later.
  mpdg_status queue_order_check(mpdg_heat *H, int k)
  {
    const bool active = k == MPDG_FP64 ? H->prm64.order != nullptr : H->prm32.order != nullptr;
    if (!active)
      return MPDG_OK;
    MPDG_HIP(hipMemcpyAsync(H->host_order_err + k, &H->order[k].ptr->error, sizeof(unsigned),
                            hipMemcpyDeviceToHost, H->stream));
    MPDG_HIP(hipEventRecord(H->order_ev[k], H->stream));
    H->order_pending[k] = true;
    return MPDG_OK;
  }

  mpdg_status check_order_state(mpdg_heat *H)
  {
    if (H->order_failed)
      return order_failure();
    for (int k = 0; k < 2; ++k)
      {
        const bool active = k == MPDG_FP64 ? H->prm64.order != nullptr : H->prm32.order != nullptr;
        if (!active)
          continue;
        unsigned err = 0;
        MPDG_HIP(hipMemcpy(&err, &H->order[k].ptr->error, sizeof(unsigned), hipMemcpyDeviceToHost));
        if (err)
          {
            H->order_failed = true;
            return order_failure();
          }
      }
    return MPDG_OK;
  }

  mpdg_status ensure_workspace(mpdg_heat *H, bool mixed)
  {
    const ColourGrid cg = colour_grid(H->desc.n_cells[0], H->desc.n_cells[1], H->desc.n_cells[2],
                                      cells_per_block(H->desc.degree, false));
    const ColourGrid cg32 = colour_grid(H->desc.n_cells[0], H->desc.n_cells[1],
                                        H->desc.n_cells[2], cells_per_block(H->desc.degree, true));
    const size_t vmult_slots = 8 * (size_t)kMaxResidentPerCu * device_cu_count();
    // the ordered vmult schedule keeps one p.Ap share per work item (at most one per cell)
    const size_t cells = (size_t)H->desc.n_cells[0] * H->desc.n_cells[1] * H->desc.n_cells[2];
    const size_t np = std::max<size_t>({(size_t)cg.total, (size_t)cg32.total, vmult_slots, cells,
                                        (size_t)kMaxBasis * kVecBlocks,
                                        (size_t)robin_blocks(H->robin),
                                        (size_t)tmat_blocks(H->desc.degree, H->desc.n_cells[0],
                                                            H->desc.n_cells[1], H->desc.n_cells[2])});
    MPDG_HIP(H->partials.alloc(np));
    MPDG_HIP(H->r.alloc(H->N));
    MPDG_HIP(H->p.alloc(H->N));
    MPDG_HIP(H->p2.alloc(H->N));
    MPDG_HIP(H->q.alloc(H->N));
    MPDG_HIP(H->inv_diag.alloc(H->N));
    if (!H->scal.ptr)
      {
        // the reduction ticket counter must start at 0; the kernels reset it after each use
        MPDG_HIP(H->scal.alloc(1));
        const SolverScalars zero{};
        MPDG_HIP(hipMemcpy(H->scal.ptr, &zero, sizeof(zero), hipMemcpyHostToDevice));
      }
    if (!H->host_scal)
      MPDG_HIP(hipHostMalloc((void **)&H->host_scal, sizeof(SolverScalars), hipHostMallocDefault));
    if (mixed)
      {
        MPDG_HIP(H->r32.alloc(H->N));
        MPDG_HIP(H->p32.alloc(H->N));
        MPDG_HIP(H->q32.alloc(H->N));
        MPDG_HIP(H->x32.alloc(H->N));
        MPDG_HIP(H->inv32.alloc(H->N));
        MPDG_HIP(H->inv16.alloc(H->N));
      }
    return MPDG_OK;
  }

  // GMRES coefficient block layout (doubles): h1[32] | h2[32] | nrm2
  constexpr int kGmH1 = 0, kGmH2 = kMaxBasis, kGmNrm = 2 * kMaxBasis, kGmSize = 2 * kMaxBasis + 8;

  // Leading dimension of the GMRES basis: N rounded up to 4 entries, so every
  // basis vector starts 16-byte aligned (FP64 and FP32) for the wide sweeps.
  int64_t basis_ld(int64_t n) { return (n + 3) & ~(int64_t)3; }

  mpdg_status ensure_gmres(mpdg_heat *H, int restart, bool fp32)
  {
    const size_t nb = (size_t)(restart + 1) * basis_ld(H->N);
    if (fp32)
      MPDG_HIP(H->V32.alloc(nb));
    else
      MPDG_HIP(H->V64.alloc(nb));
    MPDG_HIP(H->gm.alloc(kGmSize));
    if (!H->gst.ptr)
      {
        // the vmult's p.Ap ticket in gst->sc.counter must start at 0
        MPDG_HIP(H->gst.alloc(1));
        const GmresState zero{};
        MPDG_HIP(hipMemcpy(H->gst.ptr, &zero, sizeof(zero), hipMemcpyHostToDevice));
      }
    if (!H->host_gst)
      MPDG_HIP(hipHostMalloc((void **)&H->host_gst, sizeof(GmresState), hipHostMallocDefault));
    return MPDG_OK;
  }

  mpdg_status read_scalars(mpdg_heat *H)
  {
    MPDG_HIP(hipMemcpyAsync(H->host_scal, H->scal.ptr, sizeof(SolverScalars),
                            hipMemcpyDeviceToHost, H->stream));
    MPDG_HIP(hipStreamSynchronize(H->stream));
    return MPDG_OK;
  }
  // Scratch of the time-step routines: two N-vectors, reduction partials, a
  // zeroed ticket counter and a pinned result slot.
  mpdg_status ensure_newton(mpdg_heat *H)
  {
    MPDG_HIP(H->nrhs.alloc(H->N));
    MPDG_HIP(H->ndu.alloc(H->N));
    const int nb = l2_norm_blocks(H->desc.degree, H->desc.n_cells[0], H->desc.n_cells[1],
                                  H->desc.n_cells[2]);
    MPDG_HIP(H->npart.alloc((size_t)std::max(nb, kVecBlocks)));
    MPDG_HIP(H->nres.alloc(4));
    if (!H->nticket.ptr)
      {
        MPDG_HIP(H->nticket.alloc(1));
        MPDG_HIP(hipMemset(H->nticket.ptr, 0, sizeof(unsigned)));
      }
    if (!H->host_nres)
      MPDG_HIP(hipHostMalloc((void **)&H->host_nres, 4 * sizeof(double), hipHostMallocDefault));
    return MPDG_OK;
  }

  // Euclidean norm of a device vector (deterministic reduction), synchronous.
  mpdg_status device_norm(mpdg_heat *H, const double *v, double &out)
  {
    MPDG_HIP(launch_dot(v, v, H->N, H->npart.ptr, H->nres.ptr, H->nticket.ptr, H->stream));
    MPDG_HIP(hipMemcpyAsync(H->host_nres, H->nres.ptr, sizeof(double), hipMemcpyDeviceToHost,
                            H->stream));
    MPDG_HIP(hipStreamSynchronize(H->stream));
    out = std::sqrt(H->host_nres[0]);
    return MPDG_OK;
  }


  // The cell loop, then the Robin faces (MatrixFree::loop's boundary-face part)
  // on the same stream. With partials and scal, the cell loop leaves p.Ap of the
  // cells in scal->pap and the face launches add the faces' part and set alpha;
  // scal alone only stops the launches once scal->done is set. (Running
  // the face-cell integrals on a second stream beside the cell loop measured
  // slower: the fork and join cost more than the overlap saves.)
  template <typename Number>
  mpdg_status vmult_any(mpdg_heat *H, Number *dst, const Number *src, double *partials,
                        SolverScalars *scal)
  {
    if (mpdg_status s = check_tangent(H))
      return s;
    if (generic_cells(H->desc))
      {
        // coefficients per quadrature point from nodal fields (FP64 only;
        // creation rejects FP32 handles)
        if constexpr (sizeof(Number) == 8)
          MPDG_HIP(launch_tmat_vmult(H->desc.degree, velocity_mode(H->desc), H->prm64, src, dst,
                                     H->temperature, H->temperature_old,
                                     uses_heaviside(H->desc) ? H->heav64.ptr : nullptr,
                                     H->vel64.ptr, partials, scal, H->stream));
        else
          return fail(MPDG_ERR_UNSUPPORTED, "temperature-dependent materials and nodal velocity "
                                            "fields have an FP64 cell loop only");
      }
    else if constexpr (sizeof(Number) == 8)
      MPDG_HIP(launch_vmult<double>(H->desc.degree, coef_mode(H), H->desc.velocity != 0, H->prm64,
                                    H->vt64, src, dst, H->heav64.ptr, partials, scal, H->stream,
                                    H->slabs.side ? &H->slabs : nullptr));
    else
      MPDG_HIP(launch_vmult<float>(H->desc.degree, coef_mode(H), H->desc.velocity != 0, H->prm32,
                                   H->vt32, src, dst, H->heav32.ptr, partials, scal, H->stream,
                                   H->slabs.side ? &H->slabs : nullptr));
    MPDG_HIP(launch_robin<Number>(H->desc.degree, kRobinTangent, H->robin, dst, src,
                                  H->temperature, H->robin_rows.ptr, partials, scal, H->stream));
    return MPDG_OK;
  }

  mpdg_status vmult64(mpdg_heat *H, double *dst, const double *src, double *partials,
                      SolverScalars *scal)
  {
    return vmult_any<double>(H, dst, src, partials, scal);
  }

  mpdg_status vmult32(mpdg_heat *H, float *dst, const float *src, double *partials,
                      SolverScalars *scal)
  {
    return vmult_any<float>(H, dst, src, partials, scal);
  }

  mpdg_status diagonal_inverse(mpdg_heat *H, double *inv)
  {
    if (mpdg_status s = check_tangent(H))
      return s;
    if (generic_cells(H->desc))
      MPDG_HIP(launch_tmat_diagonal(H->desc.degree, velocity_mode(H->desc), H->prm64, H->dt64, inv,
                                    H->temperature, H->temperature_old,
                                    uses_heaviside(H->desc) ? H->heav64.ptr : nullptr,
                                    H->vel64.ptr, H->stream));
    else
      MPDG_HIP(launch_diagonal<double>(H->desc.degree, coef_mode(H), H->desc.velocity != 0,
                                       H->prm64, H->dt64, inv, H->heav64.ptr, H->stream));
    MPDG_HIP(launch_robin<double>(H->desc.degree, kRobinDiagonal, H->robin, inv, nullptr,
                                  H->temperature, H->robin_rows.ptr, nullptr, nullptr, H->stream));
    MPDG_HIP(launch_invert_diagonal(inv, H->N, H->stream));
    return MPDG_OK;
  }

  // Runs one Krylov chunk (`launch` issues its kernels on H->stream). With
  // graphs enabled, a chunk is launched eagerly the first time it is seen with
  // a set of vectors, chunk length and operator state, captured into a graph
  // the second time, and replayed from then on; any change of those starts
  // over. H->stream is then the handle's capture stream (solve_impl moves the
  // whole solve onto it).
  template <typename Launch>
  mpdg_status run_chunk(mpdg_heat *H, mpdg_heat::ChunkGraph &g, int steps,
                        const void *const *key, int nkey, Launch launch)
  {
    if (!H->use_graphs || H->stream == nullptr)
      return launch();
    const void *full[16]{};
    const void *extra[] = {H->temperature, H->temperature_old, H->heav64.ptr, H->heav32.ptr,
                           H->robin_rows.ptr};
    int nk = 0;
    for (int i = 0; i < nkey && nk < 16; ++i)
      full[nk++] = key[i];
    for (const void *e : extra)
      if (nk < 16)
        full[nk++] = e;
    const bool same = g.gen == H->state_gen && g.steps == steps &&
                      std::memcmp(g.key, full, sizeof(full)) == 0;
    if (!same)
      {
        if (g.exec)
          {
            MPDG_HIP(hipStreamSynchronize(H->stream));
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
          }
        g.gen = H->state_gen;
        g.steps = steps;
        std::memcpy(g.key, full, sizeof(full));
        return launch(); // first use: eager
      }
    if (!g.exec)
      {
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(H->stream, hipStreamCaptureModeThreadLocal);
        mpdg_status st = MPDG_OK;
        if (e == hipSuccess)
          {
            st = launch();
            e = hipStreamEndCapture(H->stream, &graph);
          }
        if (st)
          {
            if (graph)
              (void)hipGraphDestroy(graph);
            return st;
          }
        if (e != hipSuccess)
          return fail(MPDG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess)
          {
            g.exec = nullptr;
            return fail(MPDG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
          }
      }
    MPDG_HIP(hipGraphLaunch(g.exec, H->stream));
    return MPDG_OK;
  }

  // Device PCG on precision Number; x and the workspace vectors in Number,
  // b in double, the inverse diagonal in Number or (InvT = uint16_t) bfloat16.
  // Returns through scal (iteration, status, res).
  //
  // With a second direction vector p2 and an even check interval, the
  // iterations run in pairs: the directions alternate between p and p2 (the
  // direction update writes the other vector), the first iteration of a pair
  // leaves x alone and the second applies both steps, so x is streamed every
  // other iteration (56 -> 48 bytes per DoF and iteration of update traffic
  // on average in FP64). Every iterate, residual and x is bitwise that of
  // single updates (same roundings in the same order). A solve that stops on
  // the first iteration of a pair (an odd iteration count) gets the pending
  // x += alpha p afterwards. The FP64 CG runs in pairs (4 % faster at 16.8 M
  // DoFs); the FP32 inner CG of the mixed solve keeps single updates, which
  // measured 0.3 % faster there (its 68 MB vectors largely stay in the
  // last-level cache, and a second direction vector adds to that set).
  template <typename Number, typename InvT = Number>
  mpdg_status pcg(mpdg_heat *H, Number *x, const double *b, Number *r, Number *p, Number *q,
                  const InvT *inv, double rel_tol, double abs_tol, int max_it, int check,
                  bool x_is_zero, Number *p2 = nullptr)
  {
    const bool paired = p2 != nullptr && check % 2 == 0;
    double *part = H->partials.ptr;
    SolverScalars *sc = H->scal.ptr;
    // the residual history is the FP64 CG's (the inner solves of the mixed
    // precision solve do not record)
    double *hist = sizeof(Number) == 8 && H->history_len > 0 ? H->history : nullptr;
    const int hist_len = hist ? H->history_len : 0;
    if (x_is_zero)
      {
        MPDG_HIP(launch_cg_init<Number>(b, nullptr, r, p, inv, H->N, rel_tol, abs_tol, max_it,
                                        part, sc, H->stream));
      }
    else
      {
        mpdg_status st;
        if constexpr (sizeof(Number) == 8)
          st = vmult64(H, q, x, nullptr, nullptr);
        else
          st = vmult32(H, q, x, nullptr, nullptr);
        if (st)
          return st;
        MPDG_HIP(launch_cg_init<Number>(b, q, r, p, inv, H->N, rel_tol, abs_tol, max_it, part, sc,
                                        H->stream));
      }
    if (hist)
      MPDG_HIP(hipMemcpyAsync(hist, &sc->res0, sizeof(double), hipMemcpyDeviceToDevice,
                              H->stream));
    auto op = [&](const Number *src) -> mpdg_status {
      if constexpr (sizeof(Number) == 8)
        return vmult64(H, q, src, part, sc);
      else
        return vmult32(H, q, src, part, sc);
    };
    auto chunk = [&]() -> mpdg_status {
      if (paired)
        for (int k = 0; k < check; k += 2)
          {
            // first of the pair: direction in p, the next one into p2
            if (mpdg_status st = op(p))
              return st;
            MPDG_HIP(launch_cg_update<Number>(x, r, p, q, inv, H->N, part, sc, hist, hist_len,
                                              H->stream, kXDefer, nullptr));
            MPDG_HIP(launch_cg_direction<Number>(p2, r, inv, H->N, sc, H->stream, p));
            // second: x takes both steps, the next direction goes back into p
            if (mpdg_status st = op(p2))
              return st;
            MPDG_HIP(launch_cg_update<Number>(x, r, p2, q, inv, H->N, part, sc, hist, hist_len,
                                              H->stream, kXPair, p));
            MPDG_HIP(launch_cg_direction<Number>(p, r, inv, H->N, sc, H->stream, p2));
          }
      else
        for (int k = 0; k < check; ++k)
          {
            if (mpdg_status st = op(p))
              return st;
            MPDG_HIP(launch_cg_update<Number>(x, r, p, q, inv, H->N, part, sc, hist, hist_len,
                                              H->stream));
            MPDG_HIP(launch_cg_direction<Number>(p, r, inv, H->N, sc, H->stream));
          }
      return MPDG_OK;
    };
    const void *key[] = {x, r, p, q, inv, part, sc, hist, (const void *)(uintptr_t)hist_len,
                         paired ? p2 : nullptr};
    for (;;)
      {
        if (mpdg_status st = run_chunk(H, H->pcg_graph[sizeof(Number) == 4], check, key,
                                       sizeof(key) / sizeof(key[0]), chunk))
          return st;
        if (mpdg_status st = read_scalars(H))
          return st;
        if (H->host_scal->done)
          break;
      }
    // iterations 0, 2, 4, ... are first of a pair: an odd count ended on one
    // (its direction is in p, its alpha in scal->alpha)
    if (paired && (H->host_scal->iteration & 1))
      MPDG_HIP(launch_cg_flush<Number>(x, p, H->N, sc, H->stream));
    return MPDG_OK;
  }

  // Right-preconditioned GMRES(restart) with Jacobi, classical Gram-Schmidt
  // plus one re-orthogonalisation pass; the inner check uses the least-squares
  // residual estimate, each restart the true residual (deal.II SolverGMRES with
  // right preconditioning, linear_solver.hpp:112-119). Number is the precision
  // of the basis and of the operator; b is FP64.
  //
  // The Hessenberg column, Givens rotations and back substitution run in
  // one-thread kernels on the device (GmresState). A restart cycle is cut into
  // chunks of `chunk` Arnoldi steps; the host reads the cycle's stop flag after
  // each chunk, and every kernel of a step does nothing once the cycle has
  // stopped (the flag is the state's sc.done), so a chunk is a fixed launch
  // sequence. The cycle ends with the update x += D V y over its basis size and
  // the true residual for the next cycle. Chunks and the end replay as captured
  // graphs (run_chunk).
  template <typename Number>
  mpdg_status gmres(mpdg_heat *H, Number *x, const double *b, Number *r, Number *z, Number *w,
                    Number *V, const Number *inv, double rel_tol, double abs_tol, int max_it,
                    int restart, bool x_is_zero, int *iterations, double *res0_out,
                    double *res_out, bool *converged)
  {
    const int64_t n = H->N;
    const int64_t ld = basis_ld(n);
    double *part = H->partials.ptr;
    SolverScalars *sc = H->scal.ptr;
    unsigned *counter = &sc->counter;
    double *gm = H->gm.ptr;
    GmresState *st = H->gst.ptr;
    const int chunk = std::min(restart, H->N < (1 << 20) ? 5 : 10);
    const int nchunks = (restart + chunk - 1) / chunk;
    std::vector<mpdg_heat::ChunkGraph> &graphs = H->gmres_graphs[sizeof(Number) == 4];
    if ((int)graphs.size() < nchunks + 1)
      graphs.resize(nchunks + 1);
    // with `stop` the vmult does nothing once the cycle has stopped (no p.Ap)
    auto op = [&](Number *dst, const Number *src, SolverScalars *stop) -> mpdg_status {
      if constexpr (sizeof(Number) == 8)
        return vmult64(H, dst, src, nullptr, stop);
      else
        return vmult32(H, dst, src, nullptr, stop);
    };
    // r = b - A x and ||r|| into sc->res (z receives D r, unused)
    auto residual = [&](bool zero_x) -> mpdg_status {
      if (!zero_x)
        if (mpdg_status e = op(w, x, nullptr))
          return e;
      MPDG_HIP(launch_cg_init<Number>(b, zero_x ? nullptr : w, r, z, inv, n, 0.0, 0.0, 1, part, sc,
                                      H->stream));
      return MPDG_OK;
    };
    auto steps = [&](int j0, int j1) -> mpdg_status {
      const int *stop = &st->sc.done;
      for (int j = j0; j < j1; ++j)
        {
          if (mpdg_status e = op(w, z, &st->sc))
            return e;
          // classical Gram-Schmidt with one re-orthogonalisation pass, in three
          // sweeps over the basis: h1 = V^T w; w -= V h1 with h2 = V^T w; w -= V h2
          // with ||w||^2
          MPDG_HIP(launch_multi_dot<Number>(V, ld, j + 1, w, n, part, gm + kGmH1, counter,
                                            H->stream, stop));
          MPDG_HIP(launch_multi_axpy_dot<Number>(w, V, ld, j + 1, gm + kGmH1, n, part, gm + kGmH2,
                                                 counter, H->stream, stop));
          MPDG_HIP(launch_multi_axpy<Number>(w, V, ld, j + 1, gm + kGmH2, n, part, gm + kGmNrm,
                                             counter, H->stream, stop));
          MPDG_HIP(launch_gmres_givens(st, gm + kGmH1, gm + kGmH2, gm + kGmNrm, j, restart,
                                       H->stream));
          // v_{j+1} = w / ||w|| and the next step's z = D v_{j+1}
          MPDG_HIP(launch_scale_precondition<Number>(V + (j + 1) * ld, z, w, inv, gm + kGmNrm, n,
                                                     H->stream, stop));
        }
      return MPDG_OK;
    };
    auto finish = [&]() -> mpdg_status {
      // x += D V y, y from the triangular system, over the cycle's basis
      MPDG_HIP(launch_gmres_solve(st, H->stream));
      MPDG_HIP((launch_gmres_update<Number, Number>(x, V, ld, restart, st->y, inv, n, H->stream,
                                                    &st->k)));
      return residual(false);
    };
    auto read_state = [&]() -> mpdg_status {
      MPDG_HIP(hipMemcpyAsync(H->host_gst, st, offsetof(GmresState, cs), hipMemcpyDeviceToHost,
                              H->stream));
      return read_scalars(H);
    };
    if (mpdg_status e = residual(x_is_zero))
      return e;
    if (mpdg_status e = read_scalars(H))
      return e;
    const double res0 = H->host_scal->res;
    const double tol = std::max(abs_tol, rel_tol * res0);
    double res = res0;
    int it = 0;
    bool conv = res <= tol;
    if (!conv && max_it > 0)
      MPDG_HIP(launch_gmres_init(st, tol, max_it, H->stream));
    const void *key[] = {x, b, r, z, w, V, inv, part, sc, st, gm};
    const int nkey = sizeof(key) / sizeof(key[0]);
    while (!conv && it < max_it)
      {
        // v_0 = r / ||r||, z = D v_0, g = (||r||, 0, ...)
        MPDG_HIP(launch_gmres_start(st, sc, restart, H->stream));
        MPDG_HIP(launch_scale_precondition<Number>(V, z, r, inv, &sc->rr, n, H->stream));
        for (int c = 0; c < nchunks; ++c)
          {
            const int j0 = c * chunk, j1 = std::min(restart, j0 + chunk);
            if (mpdg_status e = run_chunk(H, graphs[c], restart * 64 + j0, key, nkey,
                                          [&]() { return steps(j0, j1); }))
              return e;
            if (c + 1 < nchunks)
              {
                if (mpdg_status e = read_state())
                  return e;
                if (H->host_gst->sc.done)
                  break;
              }
          }
        if (mpdg_status e = run_chunk(H, graphs[nchunks], restart * 64 + kMaxBasis, key, nkey,
                                      finish))
          return e;
        if (mpdg_status e = read_state())
          return e;
        const GmresState &hs = *H->host_gst;
        it = hs.it;
        if (hs.inner_conv)
          {
            conv = true;
            res = hs.res;
            break;
          }
        if (it >= max_it || std::isnan(hs.res))
          {
            res = hs.res;
            break;
          }
        res = H->host_scal->res;
        conv = res <= tol;
      }
    MPDG_HIP(hipStreamSynchronize(H->stream));
    *iterations = it;
    *res0_out = res0;
    *res_out = res;
    *converged = conv;
    return MPDG_OK;
  }

  // Longest captured Krylov chunk (iterations between host checks). A
  // 128-iteration PCG chunk of the slab-scheduled FP64 operator (about 2,900
  // graph nodes) aborted the process with a host-heap corruption inside the
  // HIP runtime on this ROCm image; 64 ran repeatedly. Longer intervals are
  // refused rather than captured.
  constexpr int kMaxCheckInterval = 64;

  int auto_check(const mpdg_heat *H)
  {
    // poll less often when an iteration is cheap
    return H->N < (1 << 20) ? 16 : 8;
  }
} // namespace

extern "C" {

const char *mpdg_last_error(void) { return g_last_error.c_str(); }

void mpdg_solver_params_default(mpdg_solver_params *p)
{
  if (!p)
    return;
  std::memset(p, 0, sizeof(*p));
  p->solver_type = MPDG_SOLVER_CG;
  p->preconditioner_type = MPDG_PRECONDITIONER_DIAGONAL;
  p->max_iterations = 10000;
  p->rel_tolerance = 1e-12;
  p->abs_tolerance = 1e-20;
  p->gmres_restart = 30;
  p->mixed_precision = 0;
  p->inner_rel_tolerance = 1e-2;
  p->check_interval = 0;
}

mpdg_status mpdg_heat_create(const mpdg_heat_desc *desc, mpdg_heat **out)
{
  if (!desc || !out)
    return fail(MPDG_ERR_ARGUMENT, "null descriptor or output pointer");
  *out = nullptr;
  const mpdg_heat_desc &d = *desc;
  if (d.degree < 1 || d.degree > kMaxDegree)
    return fail(MPDG_ERR_UNSUPPORTED, "degree must be in 1.." + std::to_string(kMaxDegree));
  int64_t N = 1, Nf = 1;
  for (int k = 0; k < 3; ++k)
    {
      if (d.n_cells[k] < 1)
        return fail(MPDG_ERR_ARGUMENT, "n_cells must be >= 1");
      if (!(d.box_hi[k] > d.box_lo[k]))
        return fail(MPDG_ERR_ARGUMENT, "box_hi must exceed box_lo");
      const int64_t n1 = (int64_t)d.n_cells[k] * d.degree + 1;
      N *= n1;
      const int64_t drop = ((d.dirichlet_faces >> (2 * k)) & 1u) + ((d.dirichlet_faces >> (2 * k + 1)) & 1u);
      Nf *= std::max<int64_t>(0, n1 - drop);
    }
  // The cell kernels address each vector through a raw buffer resource whose
  // byte size and offsets are 32-bit, and the cell form marks a skipped node by
  // bit 31 of its byte offset: an FP64 vector must stay below 2^31 bytes.
  if (N * (int64_t)sizeof(double) >= kMaxVectorBytes)
    return fail(MPDG_ERR_UNSUPPORTED,
                "mesh too large: n_dofs() = " + std::to_string(N) + " needs " +
                  std::to_string(N * (int64_t)sizeof(double)) +
                  " bytes per FP64 vector; the cell kernels address vectors below 2^31 bytes "
                  "(at most " + std::to_string(kMaxVectorBytes / (int64_t)sizeof(double) - 1) +
                  " DoFs)");
  if (d.precision != MPDG_FP64 && d.precision != MPDG_FP32)
    return fail(MPDG_ERR_ARGUMENT, "unknown precision");
  if (d.material != MPDG_MATERIAL_GAS && d.material != MPDG_MATERIAL_GAS_LIQUID &&
      d.material != MPDG_MATERIAL_LIQUID_SOLID && d.material != MPDG_MATERIAL_GAS_LIQUID_SOLID)
    return fail(MPDG_ERR_UNSUPPORTED, "material type not supported");
  if (temperature_dependent(d))
    {
      if (d.precision != MPDG_FP64)
        return fail(MPDG_ERR_UNSUPPORTED,
                    "temperature-dependent materials have an FP64 cell loop only");
      if (d.conductivity_field != MPDG_CONDUCTIVITY_MATERIAL)
        return fail(MPDG_ERR_UNSUPPORTED,
                    "separable conductivity applies to the gas material only");
      if (!std::isfinite(d.solidus_temperature) || !std::isfinite(d.liquidus_temperature) ||
          (d.solid_liquid_sharp ? !(d.liquidus_temperature >= d.solidus_temperature) :
                                  !(d.liquidus_temperature > d.solidus_temperature)))
        return fail(MPDG_ERR_ARGUMENT, "the mushy zone needs liquidus > solidus temperature "
                                       "(>= for the sharp solid-liquid transition)");
      if (d.latent_heat_of_fusion != 0.0)
        {
          // Heat::ApparentCapacity's and Material's checks (apparent_capacity.cpp:165-168,
          // material.templates.hpp:276-281)
          if (!(d.latent_heat_of_fusion > 0.0) || !std::isfinite(d.latent_heat_of_fusion) ||
              d.apparent_capacity_type < MPDG_APPARENT_CAPACITY_POLY4_BELL ||
              d.apparent_capacity_type > MPDG_APPARENT_CAPACITY_QLQ)
            return fail(MPDG_ERR_ARGUMENT, "latent heat of fusion must be finite and >= 0 with an "
                                           "apparent capacity type of mpdg_apparent_capacity");
          if (!(d.liquidus_temperature > d.solidus_temperature))
            return fail(MPDG_ERR_ARGUMENT, "The liquidus temperature must be greater than the "
                                           "solidus temperature for the apparent capacity");
          if (d.volume_specific_heat)
            return fail(MPDG_ERR_UNSUPPORTED, "the apparent capacity is not implemented for the "
                                              "volume specific heat capacity");
        }
    }
  if (nodal_velocity(d) && d.precision != MPDG_FP64)
    return fail(MPDG_ERR_UNSUPPORTED, "nodal velocity fields have an FP64 cell loop only");
  if (nodal_velocity(d) && d.conductivity_field != MPDG_CONDUCTIVITY_MATERIAL)
    return fail(MPDG_ERR_UNSUPPORTED,
                "the separable-sine conductivity is not built with a nodal velocity field");
  if (d.velocity != MPDG_VELOCITY_NONE && d.velocity != MPDG_VELOCITY_VORTEX &&
      d.velocity != MPDG_VELOCITY_CONSTANT && d.velocity != MPDG_VELOCITY_NODAL)
    return fail(MPDG_ERR_UNSUPPORTED, "velocity field not supported");
  if (d.velocity == MPDG_VELOCITY_CONSTANT &&
      !(std::isfinite(d.velocity_value[0]) && std::isfinite(d.velocity_value[1]) &&
        std::isfinite(d.velocity_value[2])))
    return fail(MPDG_ERR_ARGUMENT, "constant velocity must be finite");
  if (d.conductivity_field != MPDG_CONDUCTIVITY_MATERIAL &&
      d.conductivity_field != MPDG_CONDUCTIVITY_SEPARABLE_SINE)
    return fail(MPDG_ERR_UNSUPPORTED, "conductivity field not supported");
  if (uses_heaviside(d) && d.conductivity_field != MPDG_CONDUCTIVITY_MATERIAL)
    return fail(MPDG_ERR_UNSUPPORTED, "separable conductivity applies to the gas material only");
  if ((d.dirichlet_faces | d.convection_faces | d.radiation_faces | d.neumann_faces) & ~0x3Fu)
    return fail(MPDG_ERR_ARGUMENT, "face masks have 6 bits (faces 2d + side)");
  if (d.neumann_faces & (d.convection_faces | d.radiation_faces))
    return fail(MPDG_ERR_ARGUMENT, "It is not allowed to specify both Neumann and radiation "
                                   "and/or convection boundary conditions at the same face.");
  for (int f = 0; f < 6; ++f)
    for (int c = 0; c < 3; ++c)
      if (((d.neumann_faces >> f) & 1u) && !std::isfinite(d.neumann_flux[f][c]))
        return fail(MPDG_ERR_ARGUMENT, "Neumann fluxes must be finite");
  if (!std::isfinite(d.convection_coefficient) || !std::isfinite(d.emissivity) ||
      !std::isfinite(d.convection_temperature_infinity) ||
      !std::isfinite(d.radiation_temperature_infinity))
    return fail(MPDG_ERR_ARGUMENT, "Robin face parameters must be finite");

  mpdg_heat *H = new (std::nothrow) mpdg_heat();
  if (!H)
    return fail(MPDG_ERR_ARGUMENT, "out of host memory");
  H->desc = d;
  const int64_t cells = (int64_t)d.n_cells[0] * d.n_cells[1] * d.n_cells[2];
  H->tune[MPDG_FP64] = default_tuning(d.degree, false, cells);
  H->tune[MPDG_FP32] = default_tuning(d.degree, true, cells);
  H->stream = (hipStream_t)d.stream;
  H->N = N;
  H->N_free = Nf;
  H->t1d = make_tables(d.degree);
  for (int k = 0; k < 3; ++k)
    H->h[k] = (d.box_hi[k] - d.box_lo[k]) / d.n_cells[k];
  (void)hipGetDevice(&H->device);

  const Tables1D &t = H->t1d;
  const int n = t.n;
  H->vt64.S = even_odd<double>(t.S, n, false);
  H->vt64.St = even_odd<double>(t.S, n, true);
  H->vt64.C = even_odd<double>(t.C, n, false);
  H->vt64.Ct = even_odd<double>(t.C, n, true);
  H->vt32.S = even_odd<float>(t.S, n, false);
  H->vt32.St = even_odd<float>(t.S, n, true);
  H->vt32.C = even_odd<float>(t.C, n, false);
  H->vt32.Ct = even_odd<float>(t.C, n, true);
  double S2[kMaxN][kMaxN]{}, D2[kMaxN][kMaxN]{}, SD[kMaxN][kMaxN]{};
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      {
        S2[i][j] = t.S[i][j] * t.S[i][j];
        D2[i][j] = t.D[i][j] * t.D[i][j];
        SD[i][j] = t.S[i][j] * t.D[i][j];
      }
  H->dt64.S = even_odd<double>(t.S, n, false);
  H->dt64.S2t = even_odd<double>(S2, n, true);
  H->dt64.D2t = even_odd<double>(D2, n, true);
  H->dt64.SDt = even_odd<double>(SD, n, true);
  for (int q = 0; q < n; ++q)
    {
      H->vt64.W[q] = t.wq[q];
      H->vt32.W[q] = (float)t.wq[q];
      H->dt64.W[q] = t.wq[q];
    }
  if (symmetry_defect(t.S, n, 1) > 1e-13 || symmetry_defect(t.C, n, -1) > 1e-12 ||
      symmetry_defect(t.D, n, -1) > 1e-12 || ctable_defect(d.degree, t) > 1e-13)
    {
      delete H;
      return fail(MPDG_ERR_ARGUMENT, "1D tables failed their symmetry / compile-time consistency check");
    }

  const std::vector<double> tab = coefficient_tables(d, t, H->h);
  std::vector<float> tabf(tab.begin(), tab.end());
  hipError_t e = H->sin_tab64.alloc(tab.size());
  if (e == hipSuccess)
    e = H->sin_tab32.alloc(tabf.size());
  if (e == hipSuccess)
    e = hipMemcpy(H->sin_tab64.ptr, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(H->sin_tab32.ptr, tabf.data(), tabf.size() * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipEventCreate(&H->ev0);
  if (e == hipSuccess)
    e = hipEventCreate(&H->ev1);
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&H->order_ev[k], hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&H->host_order_err, 2 * sizeof(unsigned), hipHostMallocDefault);
  if (e == hipSuccess)
    H->host_order_err[0] = H->host_order_err[1] = 0u;
  if (e == hipSuccess && (H->tune[MPDG_FP64].schedule == MPDG_SCHEDULE_SLABS ||
                          H->tune[MPDG_FP32].schedule == MPDG_SCHEDULE_SLABS))
    e = ensure_slabs(H);
  // ordered vmult schedule: counters and epoch at 0, one flag per possible work
  // item (an item holds at least one cell)
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    {
      e = H->order[k].alloc(1);
      if (e == hipSuccess)
        e = H->order_flags[k].alloc((size_t)cells);
      if (e == hipSuccess)
        e = hipMemset(H->order[k].ptr, 0, sizeof(OrderState));
      if (e == hipSuccess)
        e = hipMemset(H->order_flags[k].ptr, 0, (size_t)cells * sizeof(unsigned));
    }
  if (e != hipSuccess)
    {
      delete H;
      return fail(MPDG_ERR_HIP, std::string("set-up: ") + hipGetErrorString(e));
    }
  fill_params<double>(H, H->prm64, H->sin_tab64.ptr);
  fill_params<float>(H, H->prm32, H->sin_tab32.ptr);
  if (mpdg_status st = update_em(H))
    {
      delete H;
      return st;
    }
  H->robin = robin_params(d, H->h);
  H->robin_rhs = robin_params(d, H->h, true);
  if (H->robin_rhs.nfaces > 0)
    {
      // rows of the residual's face set, which contains the tangent's
      const hipError_t er = H->robin_rows.alloc((size_t)robin_scratch(H->robin_rhs, d.degree));
      if (er != hipSuccess)
        {
          delete H;
          return fail(MPDG_ERR_HIP, std::string("Robin face scratch: ") + hipGetErrorString(er));
        }
    }
  *out = H;
  return MPDG_OK;
}

mpdg_status mpdg_heat_destroy(mpdg_heat *h)
{
  if (!h)
    return MPDG_OK;
  if (h->stream)
    (void)hipStreamSynchronize(h->stream);
  else
    (void)hipDeviceSynchronize();
  delete h;
  return MPDG_OK;
}

int64_t mpdg_heat_n_dofs(const mpdg_heat *h) { return h ? h->N : -1; }
int64_t mpdg_heat_n_free_dofs(const mpdg_heat *h) { return h ? h->N_free : -1; }

mpdg_status mpdg_heat_set_dt(mpdg_heat *h, double dt)
{
  if (!h)
    return fail(MPDG_ERR_ARGUMENT, "null handle");
  if (dt == 0.0 || !std::isfinite(dt))
    return fail(MPDG_ERR_ZERO_TIME_INCREMENT, "time increment must be finite and nonzero");
  if (h->dt_set && dt == h->dt)
    return MPDG_OK; // nothing changes (captured solver graphs stay valid)
  h->dt = dt;
  h->dt_set = true;
  fill_params<double>(h, h->prm64, h->sin_tab64.ptr);
  fill_params<float>(h, h->prm32, h->sin_tab32.ptr);
  return update_em(h); // the element matrix's mass term follows dt
}

mpdg_status mpdg_heat_set_velocity(mpdg_heat *h, const double *velocity)
{
  if (!h || !velocity)
    return fail(MPDG_ERR_ARGUMENT, "null handle or velocity");
  if (!nodal_velocity(h->desc))
    return fail(MPDG_ERR_ARGUMENT, "the handle was not created with MPDG_VELOCITY_NODAL");
  MPDG_HIP(h->vel64.alloc(3 * (size_t)h->N));
  MPDG_HIP(hipMemcpyAsync(h->vel64.ptr, velocity, 3 * (size_t)h->N * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
  h->has_velocity = true;
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_heaviside(mpdg_heat *h, const double *heaviside)
{
  if (!h || !heaviside)
    return fail(MPDG_ERR_ARGUMENT, "null handle or heaviside");
  MPDG_HIP(h->heav64.alloc(h->N));
  MPDG_HIP(h->heav32.alloc(h->N));
  MPDG_HIP(hipMemcpyAsync(h->heav64.ptr, heaviside, h->N * sizeof(double),
                          hipMemcpyDeviceToDevice, h->stream));
  MPDG_HIP(launch_to_f32(h->heav32.ptr, h->heav64.ptr, h->N, h->stream));
  h->has_heaviside = true;
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_heaviside_iso_q1(mpdg_heat *h, const double *heaviside_iso_q1)
{
  if (!h || !heaviside_iso_q1)
    return fail(MPDG_ERR_ARGUMENT, "null handle or heaviside");
  const int p = h->desc.degree;
  // GLL support point a of a cell in the iso element's sub-interval s = floor(x p)
  // (the last one for x = 1) at fraction t = x p - s
  IsoQ1Map m{};
  for (int a = 0; a <= p; ++a)
    {
      const double xp = h->t1d.gll[a] * p;
      const int sa = std::min(std::max((int)std::floor(xp), 0), p - 1);
      m.s[a] = sa;
      m.t[a] = a == 0 ? 0.0 : a == p ? 1.0 : xp - sa;
    }
  MPDG_HIP(h->heav64.alloc(h->N));
  MPDG_HIP(h->heav32.alloc(h->N));
  MPDG_HIP(launch_iso_q1_to_gll(h->heav64.ptr, heaviside_iso_q1, h->prm64.n1x, h->prm64.n1y,
                                h->prm64.n1z, p, m, h->stream));
  MPDG_HIP(launch_to_f32(h->heav32.ptr, h->heav64.ptr, h->N, h->stream));
  h->has_heaviside = true;
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_temperature(mpdg_heat *h, const double *temperature)
{
  if (!h)
    return fail(MPDG_ERR_ARGUMENT, "null handle");
  h->temperature = temperature;
  return MPDG_OK;
}

mpdg_status mpdg_lattice_index_of_points(const mpdg_heat_desc *desc, const double *points,
                                         int64_t count, uint32_t *lattice_index)
{
  if (!desc || !points || !lattice_index)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  const mpdg_heat_desc &d = *desc;
  if (d.degree < 1 || d.degree > kMaxDegree)
    return fail(MPDG_ERR_UNSUPPORTED, "degree must be in 1.." + std::to_string(kMaxDegree));
  const Tables1D t = make_tables(d.degree);
  // support-point coordinates of each lattice line, ascending
  std::vector<double> coords[3];
  double tol[3];
  int64_t N = 1;
  for (int k = 0; k < 3; ++k)
    {
      if (d.n_cells[k] < 1 || !(d.box_hi[k] > d.box_lo[k]))
        return fail(MPDG_ERR_ARGUMENT, "invalid box");
      const double h = (d.box_hi[k] - d.box_lo[k]) / d.n_cells[k];
      for (int c = 0; c < d.n_cells[k]; ++c)
        for (int a = 0; a < d.degree; ++a)
          coords[k].push_back(d.box_lo[k] + (c + t.gll[a]) * h);
      coords[k].push_back(d.box_hi[k]);
      // a point matches the support point within a quarter of the smallest gap
      double gap = h;
      for (size_t i = 1; i < coords[k].size(); ++i)
        gap = std::min(gap, coords[k][i] - coords[k][i - 1]);
      tol[k] = 0.25 * gap;
      N *= (int64_t)coords[k].size();
    }
  if (count != N)
    return fail(MPDG_ERR_ARGUMENT, "count must be the number of support points n_dofs()");
  if (N > ((int64_t)1 << 32))
    return fail(MPDG_ERR_ARGUMENT, "lattice too large for 32-bit indices");
  std::vector<unsigned char> seen((size_t)N, 0);
  for (int64_t i = 0; i < count; ++i)
    {
      int64_t idx[3];
      for (int k = 0; k < 3; ++k)
        {
          const double x = points[3 * i + k];
          const std::vector<double> &c = coords[k];
          const auto it = std::lower_bound(c.begin(), c.end(), x);
          int64_t j = it - c.begin();
          if (j == (int64_t)c.size() || (j > 0 && x - c[j - 1] < c[j] - x))
            --j;
          if (!(std::fabs(c[j] - x) <= tol[k]))
            return fail(MPDG_ERR_ARGUMENT, "point " + std::to_string(i) +
                                             " is not a support point of the box");
          idx[k] = j;
        }
      const int64_t I = idx[0] + (int64_t)coords[0].size() * (idx[1] + (int64_t)coords[1].size() * idx[2]);
      if (seen[I])
        return fail(MPDG_ERR_ARGUMENT, "support point " + std::to_string(I) + " given twice");
      seen[I] = 1;
      lattice_index[i] = (uint32_t)I;
    }
  return MPDG_OK;
}

mpdg_status mpdg_heat_to_lattice(mpdg_heat *h, double *lattice, const double *dofs,
                                 const uint32_t *lattice_index)
{
  if (!h || !lattice || !dofs || !lattice_index)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if ((const void *)lattice == (const void *)dofs)
    return fail(MPDG_ERR_ARGUMENT, "the permutation cannot run in place");
  MPDG_HIP(launch_permute(lattice, dofs, lattice_index, h->N, true, h->stream));
  return MPDG_OK;
}

mpdg_status mpdg_heat_from_lattice(mpdg_heat *h, double *dofs, const double *lattice,
                                   const uint32_t *lattice_index)
{
  if (!h || !lattice || !dofs || !lattice_index)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if ((const void *)lattice == (const void *)dofs)
    return fail(MPDG_ERR_ARGUMENT, "the permutation cannot run in place");
  MPDG_HIP(launch_permute(dofs, lattice, lattice_index, h->N, false, h->stream));
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_graphs(mpdg_heat *h, int enable)
{
  if (!h)
    return fail(MPDG_ERR_ARGUMENT, "null handle");
  h->use_graphs = enable != 0;
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_residual_history(mpdg_heat *h, double *history, int length)
{
  if (!h || length < 0 || (length > 0 && !history))
    return fail(MPDG_ERR_ARGUMENT, "history needs a device buffer of length >= 0 entries");
  h->history = length > 0 ? history : nullptr;
  h->history_len = length;
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_temperature_old(mpdg_heat *h, const double *temperature_old)
{
  if (!h)
    return fail(MPDG_ERR_ARGUMENT, "null handle");
  h->temperature_old = temperature_old;
  return MPDG_OK;
}

mpdg_status mpdg_heat_vmult(mpdg_heat *h, double *dst, const double *src)
{
  if (mpdg_status s = check_ready(h))
    return s;
  if (!dst || !src)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  if (dst == src)
    return fail(MPDG_ERR_ARGUMENT, "vmult cannot run in place");
  if (mpdg_status s = poll_order_state(h))
    return s;
  if (mpdg_status s = vmult64(h, dst, src, nullptr, nullptr))
    return s;
  return queue_order_check(h, MPDG_FP64);
}

mpdg_status mpdg_heat_vmult_f32(mpdg_heat *h, float *dst, const float *src)
{
  if (mpdg_status s = check_ready(h))
    return s;
  if (!dst || !src)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  if ((const void *)dst == (const void *)src)
    return fail(MPDG_ERR_ARGUMENT, "vmult cannot run in place");
  if (mpdg_status s = poll_order_state(h))
    return s;
  if (mpdg_status s = vmult32(h, dst, src, nullptr, nullptr))
    return s;
  return queue_order_check(h, MPDG_FP32);
}

mpdg_status mpdg_heat_inverse_diagonal(mpdg_heat *h, double *inv_diag)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (!inv_diag)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  return diagonal_inverse(h, inv_diag);
}

mpdg_status mpdg_heat_create_rhs(mpdg_heat *h, double *dst, const double *temperature,
                                 const double *temperature_old, const double *heat_source)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (!dst || !temperature || !temperature_old)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  MPDG_HIP(launch_create_rhs(h->desc.degree, coef_mode(h), velocity_mode(h->desc), h->prm64, dst,
                             temperature, temperature_old, heat_source, h->heav64.ptr,
                             h->vel64.ptr, h->stream));
  // rhs_boundary_loop: the Robin faces with T read plain
  MPDG_HIP(launch_robin<double>(h->desc.degree, kRobinRhs, h->robin_rhs, dst, nullptr, temperature,
                                h->robin_rows.ptr, nullptr, nullptr, h->stream));
  return MPDG_OK;
}

void mpdg_newton_params_default(mpdg_newton_params *p)
{
  if (!p)
    return;
  p->max_nonlinear_iterations = 10;
  p->field_correction_tolerance = 1e-10;
  p->residual_tolerance = 1e-9;
  p->max_nonlinear_iterations_alt = 0;
  p->field_correction_tolerance_alt = 1e-9;
  p->residual_tolerance_alt = 1e-8;
  p->use_relative_tolerances = 0;
}

mpdg_status mpdg_heat_distribute_constraints(mpdg_heat *h, double *v)
{
  if (!h || !v)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  MPDG_HIP(launch_zero_constrained(v, h->prm64.n1x, h->prm64.n1y, h->prm64.n1z,
                                   h->desc.dirichlet_faces, h->stream));
  return MPDG_OK;
}

mpdg_status mpdg_heat_l2_norm(mpdg_heat *h, const double *v, double *norm)
{
  if (!h || !v || !norm)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if (mpdg_status s = ensure_newton(h))
    return s;
  MPDG_HIP(launch_l2_norm_sq(h->desc.degree, h->prm64, v, h->npart.ptr, h->nticket.ptr,
                             h->nres.ptr, h->stream));
  MPDG_HIP(hipMemcpyAsync(h->host_nres, h->nres.ptr, sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
  MPDG_HIP(hipStreamSynchronize(h->stream));
  *norm = std::sqrt(h->host_nres[0]);
  return MPDG_OK;
}

mpdg_status mpdg_heat_linear_predictor(mpdg_heat *h, double *predictor, const double *old_vec,
                                       const double *old_old_vec, double dt, double dt_old)
{
  if (!h || !predictor || !old_vec || !old_old_vec)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if (std::abs(dt_old) < 1e-12)
    {
      if (predictor != old_vec)
        MPDG_HIP(hipMemcpyAsync(predictor, old_vec, h->N * sizeof(double),
                                hipMemcpyDeviceToDevice, h->stream));
      return MPDG_OK;
    }
  const double frac = dt / dt_old;
  MPDG_HIP(launch_axpby(predictor, frac + 1.0, old_vec, -frac, old_old_vec, h->N, h->stream));
  return MPDG_OK;
}

// LinearSolver::solve; with update_preconditioner the Jacobi preconditioner is
// updated first (JacobiPreconditioner::update), otherwise the handle's current
// inverse diagonal is used as it is.
static mpdg_status solve_impl(mpdg_heat *h, double *x, const double *b,
                              const mpdg_solver_params *params, mpdg_solver_report *report,
                              bool update_preconditioner, const double *user_inv = nullptr,
                              bool user_identity = false);
static mpdg_status solve_on_stream(mpdg_heat *h, double *x, const double *b,
                                   const mpdg_solver_params &prm, mpdg_solver_report *report,
                                   bool update_preconditioner, bool mixed, bool use_gmres,
                                   int check, const double *user_inv, bool user_identity);

mpdg_status mpdg_heat_newton_solve(mpdg_heat *h, double *temperature,
                                   const double *temperature_old, const double *heat_source,
                                   const mpdg_newton_params *newton,
                                   const mpdg_solver_params *linear, mpdg_newton_report *report)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (!temperature || !temperature_old)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  mpdg_newton_params np;
  if (newton)
    np = *newton;
  else
    mpdg_newton_params_default(&np);
  mpdg_solver_params lp;
  if (linear)
    lp = *linear;
  else
    {
      mpdg_solver_params_default(&lp);
      lp.solver_type = MPDG_SOLVER_GMRES; // heat defaults (heat_data.cpp:11-12)
      lp.preconditioner_type = MPDG_PRECONDITIONER_DIAGONAL;
    }
  if (np.max_nonlinear_iterations < 0 || np.max_nonlinear_iterations_alt < 0)
    return fail(MPDG_ERR_ARGUMENT, "negative Newton iteration limit");
  if (mpdg_status s = ensure_newton(h))
    return s;
  if (mpdg_status s = ensure_workspace(h, lp.mixed_precision != 0))
    return s;
  const auto t0 = std::chrono::steady_clock::now();
  // the operator reads the current solution (heat_diffuse_operator.hpp:128), and
  // the preconditioner is updated once, at the start value
  // (heat_diffuse_operation.cpp:275-276)
  h->temperature = temperature;
  h->temperature_old = temperature_old;
  if (lp.preconditioner_type == MPDG_PRECONDITIONER_DIAGONAL)
    {
      if (mpdg_status s = diagonal_inverse(h, h->inv_diag.ptr))
        return s;
    }
  else
    MPDG_HIP(launch_fill(h->inv_diag.ptr, 1.0, h->N, h->stream));
  mpdg_newton_report rep{};
  double res_tol = np.residual_tolerance, corr_tol = np.field_correction_tolerance;
  const int max_total = std::max(np.max_nonlinear_iterations + np.max_nonlinear_iterations_alt, 1);
  double r0 = 0.0;
  for (int i = 0; i < max_total; ++i)
    {
      if (i == np.max_nonlinear_iterations)
        {
          res_tol = np.residual_tolerance_alt;
          corr_tol = np.field_correction_tolerance_alt;
        }
      // solve_increment: rhs = 0, dT = 0, residual, Jacobian solve
      MPDG_HIP(launch_fill(h->nrhs.ptr, 0.0, h->N, h->stream));
      MPDG_HIP(launch_fill(h->ndu.ptr, 0.0, h->N, h->stream));
      if (mpdg_status s = mpdg_heat_create_rhs(h, h->nrhs.ptr, temperature, temperature_old,
                                               heat_source))
        return s;
      mpdg_solver_report lr{};
      const mpdg_status ls = solve_impl(h, h->ndu.ptr, h->nrhs.ptr, &lp, &lr, false);
      rep.linear_iterations += lr.iterations;
      if (ls != MPDG_OK)
        return ls;
      double res = 0.0, upd = 0.0;
      if (mpdg_status s = device_norm(h, h->nrhs.ptr, res))
        return s;
      if (mpdg_status s = device_norm(h, h->ndu.ptr, upd))
        return s;
      if (i == 0)
        rep.initial_residual = r0 = res;
      if (np.use_relative_tolerances)
        {
          if (r0 > 0.0)
            res /= r0;
          double tn = 0.0;
          if (mpdg_status s = device_norm(h, temperature, tn))
            return s;
          if (tn > 0.0)
            upd /= tn;
        }
      rep.residual = res;
      rep.update = upd;
      rep.iterations = i;
      if (res < res_tol && upd < corr_tol)
        {
          rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
          if (report)
            *report = rep;
          // a timed-out ordered wait makes the step's temperatures unreliable
          return check_order_state(h);
        }
      // solution += solution_update (homogeneous constraints: dT is 0 there)
      MPDG_HIP(launch_axpby(temperature, 1.0, temperature, 1.0, h->ndu.ptr, h->N, h->stream));
    }
  MPDG_HIP(hipStreamSynchronize(h->stream));
  rep.iterations = max_total;
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (report)
    *report = rep;
  if (mpdg_status s = check_order_state(h))
    return s;
  return fail(MPDG_ERR_NOT_CONVERGED, "Newton-Raphson solver did not converge");
}

mpdg_status mpdg_heat_jacobi_vmult(mpdg_heat *h, double *dst, const double *src,
                                   const double *inv_diag)
{
  if (!h || !dst || !src || !inv_diag)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  MPDG_HIP(launch_jacobi(dst, src, inv_diag, h->N, h->stream));
  return MPDG_OK;
}

mpdg_status mpdg_heat_solve(mpdg_heat *h, double *x, const double *b,
                            const mpdg_solver_params *params, mpdg_solver_report *report)
{
  const mpdg_status s = solve_impl(h, x, b, params, report, true);
  if (s == MPDG_OK || s == MPDG_ERR_NOT_CONVERGED)
    if (mpdg_status e = check_order_state(h))
      return e;
  return s;
}

mpdg_status mpdg_heat_solve_preconditioned(mpdg_heat *h, double *x, const double *b,
                                           const double *inv_diag,
                                           const mpdg_solver_params *params,
                                           mpdg_solver_report *report)
{
  if (!h)
    return fail(MPDG_ERR_ARGUMENT, "null handle");
  if (inv_diag && (inv_diag == x || inv_diag == b))
    return fail(MPDG_ERR_ARGUMENT, "the inverse diagonal must not alias x or b");
  const mpdg_status s = solve_impl(h, x, b, params, report, false, inv_diag, inv_diag == nullptr);
  if (s == MPDG_OK || s == MPDG_ERR_NOT_CONVERGED)
    if (mpdg_status e = check_order_state(h))
      return e;
  return s;
}

static mpdg_status solve_impl(mpdg_heat *h, double *x, const double *b,
                              const mpdg_solver_params *params, mpdg_solver_report *report,
                              bool update_preconditioner, const double *user_inv,
                              bool user_identity)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (!x || !b)
    return fail(MPDG_ERR_ARGUMENT, "null vector");
  mpdg_solver_params prm;
  if (params)
    prm = *params;
  else
    mpdg_solver_params_default(&prm);
  if (prm.solver_type != MPDG_SOLVER_CG && prm.solver_type != MPDG_SOLVER_GMRES)
    return fail(MPDG_ERR_UNSUPPORTED, "solver type must be CG or GMRES");
  const bool use_gmres = prm.solver_type == MPDG_SOLVER_GMRES;
  if (use_gmres && (prm.gmres_restart < 1 || prm.gmres_restart > kMaxBasis - 1))
    return fail(MPDG_ERR_ARGUMENT, "gmres_restart must be in 1..31");
  if (prm.max_iterations < 1 || !(prm.rel_tolerance >= 0) || !(prm.abs_tolerance >= 0))
    return fail(MPDG_ERR_ARGUMENT, "invalid solver parameters");
  if (prm.check_interval < 0 || prm.check_interval > kMaxCheckInterval)
    return fail(MPDG_ERR_ARGUMENT, "check_interval must be in 0..64 (0: automatic)");
  const bool mixed = prm.mixed_precision != 0;
  if (mixed && generic_cells(h->desc))
    return fail(MPDG_ERR_UNSUPPORTED, "temperature-dependent materials and nodal velocity fields "
                                      "have an FP64 cell loop only (no mixed precision)");
  if (mpdg_status s = ensure_workspace(h, mixed))
    return s;
  if (use_gmres)
    if (mpdg_status s = ensure_gmres(h, prm.gmres_restart, mixed))
      return s;
  const int check = prm.check_interval > 0 ? prm.check_interval : auto_check(h);
  // With graphs the solve runs on the handle's capture stream (the caller's
  // stream may be the legacy NULL stream, which cannot be captured): one fork
  // event orders it after the caller's work, one join event orders the
  // caller's later work after it.
  hipStream_t user = h->stream;
  if (h->use_graphs)
    {
      if (!h->gstream)
        {
          MPDG_HIP(hipStreamCreateWithFlags(&h->gstream, hipStreamNonBlocking));
          MPDG_HIP(hipEventCreateWithFlags(&h->gfork, hipEventDisableTiming));
          MPDG_HIP(hipEventCreateWithFlags(&h->gjoin, hipEventDisableTiming));
        }
      MPDG_HIP(hipEventRecord(h->gfork, user));
      MPDG_HIP(hipStreamWaitEvent(h->gstream, h->gfork, 0));
      h->stream = h->gstream;
    }
  const mpdg_status st = solve_on_stream(h, x, b, prm, report, update_preconditioner, mixed,
                                         use_gmres, check, user_inv, user_identity);
  if (h->stream != user)
    {
      h->stream = user;
      MPDG_HIP(hipEventRecord(h->gjoin, h->gstream));
      MPDG_HIP(hipStreamWaitEvent(user, h->gjoin, 0));
    }
  return st;
}

static mpdg_status solve_on_stream(mpdg_heat *h, double *x, const double *b,
                                   const mpdg_solver_params &prm, mpdg_solver_report *report,
                                   bool update_preconditioner, bool mixed, bool use_gmres,
                                   int check, const double *user_inv, bool user_identity)
{
  mpdg_solver_report rep{};
  const auto t0 = std::chrono::steady_clock::now();

  // The inverse diagonal the Krylov loop applies: the caller's (a
  // JacobiPreconditioner updated by the caller) or the handle's own.
  const double *inv = user_inv ? user_inv : h->inv_diag.ptr;
  // JacobiPreconditioner::update -> compute_inverse_diagonal_from_matrixfree
  if (user_inv)
    ;
  else if (user_identity)
    MPDG_HIP(launch_fill(h->inv_diag.ptr, 1.0, h->N, h->stream));
  else if (!update_preconditioner)
    ;
  else if (prm.preconditioner_type == MPDG_PRECONDITIONER_DIAGONAL)
    {
      if (mpdg_status s = diagonal_inverse(h, h->inv_diag.ptr))
        return s;
    }
  else
    MPDG_HIP(launch_fill(h->inv_diag.ptr, 1.0, h->N, h->stream));

  mpdg_status status = MPDG_OK;
  if (!mixed && use_gmres)
    {
      int its = 0;
      double r0 = 0.0, rf = 0.0;
      bool conv = false;
      if (mpdg_status s = gmres<double>(h, x, b, h->r.ptr, h->p.ptr, h->q.ptr, h->V64.ptr,
                                        inv, prm.rel_tolerance, prm.abs_tolerance,
                                        prm.max_iterations, prm.gmres_restart, false, &its, &r0,
                                        &rf, &conv))
        return s;
      rep.iterations = its;
      rep.initial_residual = r0;
      rep.final_residual = rf;
      if (!conv)
        status = fail(MPDG_ERR_NOT_CONVERGED, "iterative solver did not converge");
    }
  else if (!mixed)
    {
      if (mpdg_status s = pcg<double>(h, x, b, h->r.ptr, h->p.ptr, h->q.ptr, inv,
                                      prm.rel_tolerance, prm.abs_tolerance, prm.max_iterations,
                                      check, false, h->p2.ptr))
        return s;
      rep.iterations = h->host_scal->iteration;
      rep.outer_iterations = 0;
      rep.initial_residual = h->host_scal->res0;
      rep.final_residual = h->host_scal->res;
      if (h->host_scal->status != 0)
        status = fail(MPDG_ERR_NOT_CONVERGED, "iterative solver did not converge");
    }
  else
    {
      // Defect correction: r = b - A64 x in FP64, inner FP32 PCG for A d = r,
      // x += d in FP64, until ||r|| <= max(abs, rel ||r0||). The inner PCG
      // applies the Jacobi preconditioner from a bfloat16 copy of the inverse
      // diagonal (relative rounding <= 2^-9): a preconditioner only has to be
      // a fixed SPD approximation, the FP64 outer residual decides accuracy,
      // and the copy takes 2 of the 28 (update) and 16 (direction) bytes per
      // DoF and iteration off the inner loop's vector traffic. Inner GMRES
      // keeps the FP32 copy.
      if (use_gmres)
        MPDG_HIP(launch_to_f32(h->inv32.ptr, inv, h->N, h->stream));
      else
        MPDG_HIP(launch_to_bf16(h->inv16.ptr, inv, h->N, h->stream));
      double tol = 0.0, res0 = 0.0;
      int inner_total = 0, outer = 0;
      for (;;)
        {
          if (mpdg_status s = vmult64(h, h->q.ptr, x, nullptr, nullptr))
            return s;
          MPDG_HIP(launch_residual(h->r.ptr, b, h->q.ptr, h->N, h->partials.ptr, h->scal.ptr,
                                   h->stream));
          if (mpdg_status s = read_scalars(h))
            return s;
          const double res = h->host_scal->res;
          if (outer == 0)
            {
              res0 = res;
              tol = std::max(prm.abs_tolerance, prm.rel_tolerance * res);
            }
          rep.final_residual = res;
          if (res <= tol)
            break;
          if (inner_total >= prm.max_iterations || !std::isfinite(res))
            {
              status = fail(MPDG_ERR_NOT_CONVERGED, "defect correction did not converge");
              break;
            }
          ++outer;
          // inner solve: the relative reduction is capped so the last step
          // does not oversolve past the outer tolerance
          const double inner_rel = std::max(prm.inner_rel_tolerance, 0.5 * tol / res);
          MPDG_HIP(launch_fill_f32(h->x32.ptr, 0.0f, h->N, h->stream));
          if (use_gmres)
            {
              int its = 0;
              double r0 = 0.0, rf = 0.0;
              bool conv = false;
              if (mpdg_status s = gmres<float>(h, h->x32.ptr, h->r.ptr, h->r32.ptr, h->p32.ptr,
                                               h->q32.ptr, h->V32.ptr, h->inv32.ptr, inner_rel, 0.0,
                                               prm.max_iterations - inner_total, prm.gmres_restart,
                                               true, &its, &r0, &rf, &conv))
                return s;
              inner_total += its;
            }
          else
            {
              if (mpdg_status s = pcg<float>(h, h->x32.ptr, h->r.ptr, h->r32.ptr, h->p32.ptr,
                                             h->q32.ptr, h->inv16.ptr, inner_rel, 0.0,
                                             prm.max_iterations - inner_total, check, true))
                return s;
              inner_total += h->host_scal->iteration;
            }
          MPDG_HIP(launch_add_f32(x, h->x32.ptr, h->N, h->stream));
        }
      rep.iterations = inner_total;
      rep.outer_iterations = outer;
      rep.initial_residual = res0;
    }
  MPDG_HIP(hipStreamSynchronize(h->stream));
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (report)
    *report = rep;
  return status;
}

mpdg_status mpdg_heat_get_tuning(const mpdg_heat *h, int precision, mpdg_tuning *t)
{
  if (!h || !t)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if (precision != MPDG_FP64 && precision != MPDG_FP32)
    return fail(MPDG_ERR_ARGUMENT, "precision must be MPDG_FP64 or MPDG_FP32");
  *t = h->tune[precision];
  return MPDG_OK;
}

mpdg_status mpdg_heat_set_tuning(mpdg_heat *h, int precision, const mpdg_tuning *t)
{
  if (!h || !t)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if (precision != MPDG_FP64 && precision != MPDG_FP32)
    return fail(MPDG_ERR_ARGUMENT, "precision must be MPDG_FP64 or MPDG_FP32");
  if (!tuning_valid(*t, h->desc.degree, precision))
    return fail(MPDG_ERR_ARGUMENT,
                "tuning needs z_layers >= 1, blocks_per_cu in 0..16; the cell variant degree <= 2 "
                "and tile_log2 in 3..6; the MFMA variant degree >= 3 and MPDG_FP32; the "
                "element-matrix variant degree <= 2 and tile_log2 in 1..4");
  if (t->variant == MPDG_VMULT_CELL_MFMA && !em_eligible(h->desc))
    return fail(MPDG_ERR_UNSUPPORTED,
                "the element-matrix variant needs the same cell matrix in every cell: gas "
                "material, no conductivity field, no or a constant velocity");
  if (h->stream)
    MPDG_HIP(hipStreamSynchronize(h->stream));
  if (t->schedule == MPDG_SCHEDULE_SLABS)
    MPDG_HIP(ensure_slabs(h));
  h->tune[precision] = *t;
  fill_params<double>(h, h->prm64, h->sin_tab64.ptr);
  fill_params<float>(h, h->prm32, h->sin_tab32.ptr);
  return MPDG_OK;
}

mpdg_status mpdg_heat_autotune(mpdg_heat *h, int precision, int repeats, mpdg_tuning *best)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (precision != MPDG_FP64 && precision != MPDG_FP32)
    return fail(MPDG_ERR_ARGUMENT, "precision must be MPDG_FP64 or MPDG_FP32");
  if (repeats < 1)
    return fail(MPDG_ERR_ARGUMENT, "repeats must be >= 1");
  if (generic_cells(h->desc))
    return fail(MPDG_ERR_UNSUPPORTED, "the tangent of a temperature-dependent material or of a "
                                      "nodal velocity field has one launch form (no tuning)");
  const bool fp32 = precision == MPDG_FP32;
  if (mpdg_status s = ensure_workspace(h, fp32))
    return s;
  MPDG_HIP(ensure_slabs(h));
  // scratch vectors: src = a fill of ones, dst = the PCG's q
  void *src = fp32 ? (void *)h->p32.ptr : (void *)h->p.ptr;
  void *dst = fp32 ? (void *)h->q32.ptr : (void *)h->q.ptr;
  if (fp32)
    MPDG_HIP(launch_fill_f32((float *)src, 1.0f, h->N, h->stream));
  else
    MPDG_HIP(launch_fill((double *)src, 1.0, h->N, h->stream));

  std::vector<mpdg_tuning> cand;
  const int nz = h->desc.n_cells[2];
  for (int zl : {1, 2, 4, 8})
    {
      if (zl > 1 && zl >= 2 * nz)
        continue;
      for (int bpc : {0, 2, 4, 8})
        {
          for (int sch : {MPDG_SCHEDULE_PASSES, MPDG_SCHEDULE_ORDERED, MPDG_SCHEDULE_SLABS})
            {
              cand.push_back(mpdg_tuning{zl, bpc, MPDG_VMULT_LINE, 4, sch});
              if (fp32 && h->desc.degree >= kMfmaMinDegree)
                cand.push_back(mpdg_tuning{zl, bpc, MPDG_VMULT_LINE_MFMA, 4, sch});
            }
          if (h->desc.degree <= kCellMaxDegree)
            for (int tl : {3, 4, 5, 6})
              cand.push_back(mpdg_tuning{zl, bpc, MPDG_VMULT_CELL, tl, MPDG_SCHEDULE_PASSES});
          // element-matrix form: the widest tiles (it measured slower than the
          // line and cell forms on every mesh tried, so only a few shapes)
          if (em_eligible(h->desc) && bpc == 0)
            for (int tl : {3, 4})
              cand.push_back(mpdg_tuning{zl, bpc, MPDG_VMULT_CELL_MFMA, tl, MPDG_SCHEDULE_PASSES});
        }
    }
  const mpdg_tuning saved = h->tune[precision];
  double best_ms = 0.0;
  mpdg_tuning pick = saved;
  for (const mpdg_tuning &t : cand)
    {
      h->tune[precision] = t;
      fill_params<double>(h, h->prm64, h->sin_tab64.ptr);
      fill_params<float>(h, h->prm32, h->sin_tab32.ptr);
      double ms = 0.0;
      mpdg_status st = fp32 ? vmult32(h, (float *)dst, (const float *)src, nullptr, nullptr) :
                              vmult64(h, (double *)dst, (const double *)src, nullptr, nullptr);
      if (st == MPDG_OK)
        st = mpdg_heat_time_vmult(h, (double *)dst, (const double *)src, repeats, precision, &ms);
      if (st != MPDG_OK)
        {
          h->tune[precision] = saved;
          fill_params<double>(h, h->prm64, h->sin_tab64.ptr);
          fill_params<float>(h, h->prm32, h->sin_tab32.ptr);
          return st;
        }
      if (best_ms == 0.0 || ms < best_ms)
        {
          best_ms = ms;
          pick = t;
        }
    }
  h->tune[precision] = pick;
  fill_params<double>(h, h->prm64, h->sin_tab64.ptr);
  fill_params<float>(h, h->prm32, h->sin_tab32.ptr);
  if (best)
    *best = pick;
  return MPDG_OK;
}

mpdg_status mpdg_heat_time_vmult(mpdg_heat *h, double *dst, const double *src, int repeats,
                                 int precision, double *ms_per_vmult)
{
  if (mpdg_status s = check_call(h))
    return s;
  if (!dst || !src || repeats < 1 || !ms_per_vmult)
    return fail(MPDG_ERR_ARGUMENT, "invalid timing arguments");
  MPDG_HIP(hipEventRecord(h->ev0, h->stream));
  for (int i = 0; i < repeats; ++i)
    {
      mpdg_status s = precision == MPDG_FP32 ?
                        vmult32(h, (float *)dst, (const float *)src, nullptr, nullptr) :
                        vmult64(h, dst, src, nullptr, nullptr);
      if (s)
        return s;
    }
  MPDG_HIP(hipEventRecord(h->ev1, h->stream));
  MPDG_HIP(hipEventSynchronize(h->ev1));
  float ms = 0.0f;
  MPDG_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  *ms_per_vmult = ms / repeats;
  return check_order_state(h);
}

} // extern "C"

// ---------------------------------------------------------------------------
// Batches of independent solves
// ---------------------------------------------------------------------------
namespace
{
  template <typename Number>
  struct BatchGroup
  {
    int degree = 0;
    int nops = 0;
    int variant = kVariantLine;        // kernel form of ops (precision Number)
    int variant64 = kVariantLine;      // kernel form of opsx (FP64)
    std::vector<int> members;          // solve index of each operator
    std::vector<int> block_off;        // 8 x (nops + 1), host
    BatchVmultPlan plan{};
    int pbase = 0;                     // first vmult partial slot of the group
    DevBuf<BatchOperator<Number>> ops; // src/dst: p -> q
    DevBuf<BatchOperator<double>> opsx;// FP64 x -> q (defect-correction residual)
    DevBuf<int> dev_block_off;
    DevBuf<int> dev_members;
  };
} // namespace

struct mpdg_heat_batch
{
  int count = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false; // graphs cannot be captured on the legacy NULL stream
  std::vector<mpdg_heat *> members;
  std::vector<int64_t> off;  // member offsets in the concatenated vectors
  int64_t total = 0;
  int nblocks = 0;           // workgroups of the segmented vector kernels
  int coef = 0;
  bool vel = false;
  DevBuf<BatchSeg> seg;
  DevBuf<double> X, B, R, P, Q, D, partials, rel, res;
  DevBuf<float> R32, P32, Q32, X32, D32;
  DevBuf<int> active;
  DevBuf<SolverScalars> scal;
  DevBuf<SolverScalars> ogate; // .done = 1: member out of the FP64 residual vmult
  DevBuf<BatchOuter> outer;    // per-member defect-correction state (mixed solves)
  SolverScalars *host_scal = nullptr;
  BatchOuter *host_outer = nullptr;
  double *host_res = nullptr;
  double clock_hz = 0.0;       // device wall clock rate
  hipGraphExec_t graph_mixed = nullptr;
  int graph_steps_mixed = 0;
  double graph_abs = -1.0, graph_rel = -1.0, graph_inner = -1.0;
  int graph_max_it = -1;
  std::vector<BatchGroup<double>> g64;
  std::vector<BatchGroup<float>> g32;
  int vmult_slots64 = 0, vmult_slots32 = 0;
  hipGraphExec_t graph64 = nullptr, graph32 = nullptr;
  int graph_steps64 = 0, graph_steps32 = 0;

  ~mpdg_heat_batch()
  {
    if (own_stream && stream)
      (void)hipStreamSynchronize(stream);
    if (graph64)
      (void)hipGraphExecDestroy(graph64);
    if (graph32)
      (void)hipGraphExecDestroy(graph32);
    if (graph_mixed)
      (void)hipGraphExecDestroy(graph_mixed);
    for (auto *m : members)
      delete m;
    for (auto *b : {&X, &B, &R, &P, &Q, &D, &partials, &rel, &res})
      b->release();
    for (auto *b : {&R32, &P32, &Q32, &X32, &D32})
      b->release();
    active.release();
    scal.release();
    ogate.release();
    outer.release();
    seg.release();
    for (auto &g : g64)
      {
        g.ops.release();
        g.opsx.release();
        g.dev_block_off.release();
        g.dev_members.release();
      }
    for (auto &g : g32)
      {
        g.ops.release();
        g.opsx.release();
        g.dev_block_off.release();
        g.dev_members.release();
      }
    if (host_scal)
      (void)hipHostFree(host_scal);
    if (host_outer)
      (void)hipHostFree(host_outer);
    if (host_res)
      (void)hipHostFree(host_res);
    if (own_stream && stream)
      (void)hipStreamDestroy(stream);
  }
};

namespace
{
  template <typename Number>
  mpdg_status build_groups(mpdg_heat_batch *B, std::vector<BatchGroup<Number>> &groups,
                           int &slots, Number *src, Number *dst)
  {
    const bool fp32 = sizeof(Number) == 4;
    groups.clear();
    std::vector<int> degrees;
    for (auto *m : B->members)
      if (std::find(degrees.begin(), degrees.end(), m->desc.degree) == degrees.end())
        degrees.push_back(m->desc.degree);
    std::sort(degrees.begin(), degrees.end());
    slots = 0;
    for (int deg : degrees)
      {
        BatchGroup<Number> g;
        g.degree = deg;
        std::vector<int> items; // per operator and colour, in workgroup units of the plan
        std::vector<BatchOperator<Number>> ops;
        std::vector<BatchOperator<double>> opsx;
        int variant = -1, tile = 4, variant64 = -1, tile64 = 4;
        for (int s = 0; s < B->count; ++s)
          {
            mpdg_heat *m = B->members[s];
            if (m->desc.degree != deg)
              continue;
            CellParams<Number> prm = *reinterpret_cast<const CellParams<Number> *>(
              fp32 ? (const void *)&m->prm32 : (const void *)&m->prm64);
            CellParams<double> prm64 = m->prm64;
            if (variant < 0)
              {
                variant = prm.variant;
                tile = prm.tile_log2;
                variant64 = prm64.variant;
                tile64 = prm64.tile_log2;
                // batched launches have no element-matrix form: its members take the line form
                if (variant == kVariantCellMfma)
                  variant = kVariantLine;
                if (variant64 == kVariantCellMfma)
                  variant64 = kVariantLine;
              }
            // one kernel form per group and precision
            prm.variant = variant;
            prm.tile_log2 = tile;
            prm64.variant = variant64;
            prm64.tile_log2 = tile64;
            const VmultGrid vg = vmult_items<Number>(deg, prm);
            const int per_block = variant == kVariantCell ? cell_items_per_block() : 1;
            BatchOperator<Number> op{};
            op.prm = prm;
            op.src = src + B->off[s];
            op.dst = dst + B->off[s];
            op.heav = fp32 ? (const Number *)(const void *)m->heav32.ptr :
                             (const Number *)(const void *)m->heav64.ptr;
            op.solve = s;
            for (int c = 0; c < 8; ++c)
              {
                op.items[c] = vg.items[c];
                items.push_back((vg.items[c] + per_block - 1) / per_block);
              }
            ops.push_back(op);
            BatchOperator<double> ox{};
            ox.prm = prm64;
            ox.src = B->X.ptr + B->off[s];
            ox.dst = B->Q.ptr + B->off[s];
            ox.heav = m->heav64.ptr;
            ox.solve = s;
            const VmultGrid vg64 = vmult_items<double>(deg, prm64);
            for (int c = 0; c < 8; ++c)
              ox.items[c] = vg64.items[c];
            opsx.push_back(ox);
            g.members.push_back(s);
          }
        g.variant = variant;
        g.variant64 = variant64;
        g.nops = (int)ops.size();
        g.block_off.assign(8 * (g.nops + 1), 0);
        const int max_blocks =
          vmult_resident_blocks(deg, fp32, B->coef, B->vel, variant) * device_cu_count();
        g.plan = batch_vmult_plan(items.data(), g.nops, max_blocks, g.block_off.data());
        g.pbase = slots;
        slots += g.plan.total_slots;
        MPDG_HIP(g.ops.alloc(g.nops));
        MPDG_HIP(g.opsx.alloc(g.nops));
        MPDG_HIP(g.dev_block_off.alloc(g.block_off.size()));
        MPDG_HIP(g.dev_members.alloc(g.nops));
        MPDG_HIP(hipMemcpy(g.ops.ptr, ops.data(), ops.size() * sizeof(ops[0]), hipMemcpyHostToDevice));
        MPDG_HIP(hipMemcpy(g.opsx.ptr, opsx.data(), opsx.size() * sizeof(opsx[0]), hipMemcpyHostToDevice));
        MPDG_HIP(hipMemcpy(g.dev_block_off.ptr, g.block_off.data(), g.block_off.size() * sizeof(int),
                           hipMemcpyHostToDevice));
        MPDG_HIP(hipMemcpy(g.dev_members.ptr, g.members.data(), g.members.size() * sizeof(int),
                           hipMemcpyHostToDevice));
        groups.push_back(std::move(g));
      }
    return MPDG_OK;
  }

  // q = A p for every member whose PCG is still running, then alpha per member.
  template <typename Number>
  mpdg_status batch_vmult_pq(mpdg_heat_batch *B, std::vector<BatchGroup<Number>> &groups)
  {
    for (auto &g : groups)
      MPDG_HIP(launch_vmult_batch<Number>(g.degree, B->coef, B->vel, g.variant, g.ops.ptr,
                                          g.dev_block_off.ptr,
                                          g.nops, g.plan, B->partials.ptr + g.pbase, B->scal.ptr,
                                          B->stream));
    for (auto &g : groups)
      MPDG_HIP(launch_batch_pap_finalize(g.dev_block_off.ptr, g.dev_members.ptr, g.nops, g.plan,
                                         B->partials.ptr + g.pbase, B->scal.ptr, B->stream));
    return MPDG_OK;
  }

  template <typename Number>
  mpdg_status batch_pcg_steps(mpdg_heat_batch *B, std::vector<BatchGroup<Number>> &groups,
                              int steps, int vslots, Number *x, Number *r, Number *p, Number *q,
                              const Number *d)
  {
    double *vpart = B->partials.ptr + vslots;
    for (int k = 0; k < steps; ++k)
      {
        if (mpdg_status st = batch_vmult_pq<Number>(B, groups))
          return st;
        MPDG_HIP(launch_batch_cg_update<Number>(x, r, p, q, d, B->seg.ptr, B->count, B->nblocks,
                                                vpart, B->scal.ptr, B->stream));
        MPDG_HIP(launch_batch_cg_direction<Number>(p, r, d, B->seg.ptr, B->count, B->nblocks,
                                                   B->scal.ptr, B->stream));
      }
    return MPDG_OK;
  }

  // Captures `steps` PCG iterations of the whole batch into one executable graph.
  template <typename Number>
  mpdg_status batch_graph(mpdg_heat_batch *B, std::vector<BatchGroup<Number>> &groups, int steps,
                          int vslots, Number *x, Number *r, Number *p, Number *q, const Number *d,
                          hipGraphExec_t *exec)
  {
    hipGraph_t graph = nullptr;
    MPDG_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    mpdg_status st = batch_pcg_steps<Number>(B, groups, steps, vslots, x, r, p, q, d);
    const hipError_t e = hipStreamEndCapture(B->stream, &graph);
    if (st)
      {
        if (graph)
          (void)hipGraphDestroy(graph);
        return st;
      }
    if (e != hipSuccess)
      return fail(MPDG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess)
      return fail(MPDG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    return MPDG_OK;
  }

  // One chunk of the per-member defect correction of a mixed batch solve:
  // the FP64 residual vmult of the Outer members (the others are gated off),
  // their outer step, the FP32 PCG start of the InnerStart members, `steps`
  // FP32 PCG iterations of the Inner members and the finish (x += d) of those
  // whose inner solve stopped. Every member moves through its own outer steps,
  // so a member's convergence does not wait for the others'.
  mpdg_status batch_mixed_chunk(mpdg_heat_batch *B, int steps, const mpdg_solver_params &prm)
  {
    double *vpart = B->partials.ptr + B->vmult_slots32;
    for (auto &g : B->g64)
      MPDG_HIP(launch_vmult_batch<double>(g.degree, B->coef, B->vel, g.variant64, g.opsx.ptr,
                                          g.dev_block_off.ptr, g.nops, g.plan, nullptr,
                                          B->ogate.ptr, B->stream));
    MPDG_HIP(launch_batch_outer_step(B->R.ptr, B->B.ptr, B->Q.ptr, B->seg.ptr, B->count,
                                     B->nblocks, vpart, B->outer.ptr, B->ogate.ptr,
                                     prm.abs_tolerance, prm.rel_tolerance,
                                     prm.inner_rel_tolerance, prm.max_iterations, B->stream));
    MPDG_HIP(launch_batch_inner_start(B->R.ptr, B->R32.ptr, B->P32.ptr, B->X32.ptr, B->D32.ptr,
                                      B->seg.ptr, B->count, B->nblocks, vpart, B->outer.ptr,
                                      B->scal.ptr, prm.max_iterations, B->stream));
    if (mpdg_status st = batch_pcg_steps<float>(B, B->g32, steps, B->vmult_slots32, B->X32.ptr,
                                                B->R32.ptr, B->P32.ptr, B->Q32.ptr, B->D32.ptr))
      return st;
    MPDG_HIP(launch_batch_inner_finish(B->X.ptr, B->X32.ptr, B->seg.ptr, B->count, B->nblocks,
                                       B->outer.ptr, B->scal.ptr, B->ogate.ptr, B->stream));
    return MPDG_OK;
  }

  mpdg_status batch_mixed_graph(mpdg_heat_batch *B, int steps, const mpdg_solver_params &prm,
                                hipGraphExec_t *exec)
  {
    hipGraph_t graph = nullptr;
    MPDG_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
    mpdg_status st = batch_mixed_chunk(B, steps, prm);
    const hipError_t e = hipStreamEndCapture(B->stream, &graph);
    if (st)
      {
        if (graph)
          (void)hipGraphDestroy(graph);
        return st;
      }
    if (e != hipSuccess)
      return fail(MPDG_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    const hipError_t e2 = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess)
      return fail(MPDG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e2));
    return MPDG_OK;
  }

  mpdg_status read_batch_scalars(mpdg_heat_batch *B)
  {
    MPDG_HIP(hipMemcpyAsync(B->host_scal, B->scal.ptr, B->count * sizeof(SolverScalars),
                            hipMemcpyDeviceToHost, B->stream));
    MPDG_HIP(hipStreamSynchronize(B->stream));
    return MPDG_OK;
  }
} // namespace

extern "C" {

mpdg_status mpdg_heat_batch_create(const mpdg_heat_desc *descs, int count, mpdg_heat_batch **out)
{
  if (!descs || !out || count < 1)
    return fail(MPDG_ERR_ARGUMENT, "batch needs a descriptor array and count >= 1");
  *out = nullptr;
  for (int s = 1; s < count; ++s)
    if (descs[s].material != descs[0].material ||
        descs[s].conductivity_field != descs[0].conductivity_field ||
        descs[s].velocity != descs[0].velocity || descs[s].stream != descs[0].stream)
      return fail(MPDG_ERR_UNSUPPORTED,
                  "batch members must share material kind, conductivity and velocity fields and stream");
  if (descs[0].material != MPDG_MATERIAL_GAS)
    return fail(MPDG_ERR_UNSUPPORTED, "batches are built for the gas material only");
  if (nodal_velocity(descs[0]))
    return fail(MPDG_ERR_UNSUPPORTED, "batches are not built for nodal velocity fields");
  for (int s = 0; s < count; ++s)
    if ((descs[s].convection_faces | descs[s].radiation_faces) & 0x3Fu)
      return fail(MPDG_ERR_UNSUPPORTED, "batches of operators with Robin faces are not built");
  // each member's own limits are checked by mpdg_heat_create; the vector size
  // limit is checked here as well, before the batch creates its stream
  for (int s = 0; s < count; ++s)
    {
      int64_t n = 1;
      for (int k = 0; k < 3; ++k)
        n *= std::max<int64_t>(1, (int64_t)descs[s].n_cells[k] * descs[s].degree + 1);
      if (n * (int64_t)sizeof(double) >= kMaxVectorBytes)
        return fail(MPDG_ERR_UNSUPPORTED,
                    "batch member " + std::to_string(s) + ": mesh too large (n_dofs() = " +
                      std::to_string(n) + "); the cell kernels address vectors below 2^31 bytes");
    }
  mpdg_heat_batch *B = new (std::nothrow) mpdg_heat_batch();
  if (!B)
    return fail(MPDG_ERR_ARGUMENT, "out of host memory");
  B->count = count;
  B->stream = (hipStream_t)descs[0].stream;
  if (!B->stream)
    {
      if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
        {
          delete B;
          return fail(MPDG_ERR_HIP, "could not create the batch stream");
        }
      B->own_stream = true;
    }
  for (int s = 0; s < count; ++s)
    {
      mpdg_heat *h = nullptr;
      mpdg_heat_desc d = descs[s];
      d.stream = B->stream;
      if (mpdg_status st = mpdg_heat_create(&d, &h))
        {
          delete B;
          return st;
        }
      B->members.push_back(h);
    }
  B->coef = coef_mode(B->members[0]);
  B->vel = descs[0].velocity != 0;
  std::vector<BatchSeg> seg(count);
  int64_t off = 0;
  int blk = 0;
  for (int s = 0; s < count; ++s)
    {
      const int64_t n = B->members[s]->N;
      B->off.push_back(off);
      seg[s].off = off;
      seg[s].n = n;
      seg[s].blk0 = blk;
      seg[s].nblk = (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + 4 * kBatchThreads - 1) /
                                                                        (4 * kBatchThreads)));
      blk += seg[s].nblk;
      off += n;
    }
  B->total = off;
  B->nblocks = blk;
  hipError_t e = B->seg.alloc(count);
  for (auto *buf : {&B->X, &B->B, &B->R, &B->P, &B->Q, &B->D})
    if (e == hipSuccess)
      e = buf->alloc(B->total);
  for (auto *buf : {&B->R32, &B->P32, &B->Q32, &B->X32, &B->D32})
    if (e == hipSuccess)
      e = buf->alloc(B->total);
  if (e == hipSuccess)
    e = B->rel.alloc(count);
  if (e == hipSuccess)
    e = B->res.alloc(2 * count);
  if (e == hipSuccess)
    e = B->active.alloc(count);
  if (e == hipSuccess)
    e = B->scal.alloc(count);
  if (e == hipSuccess)
    e = B->ogate.alloc(count);
  if (e == hipSuccess)
    e = B->outer.alloc(count);
  if (e == hipSuccess)
    e = hipMemcpy(B->seg.ptr, seg.data(), count * sizeof(BatchSeg), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    {
      std::vector<SolverScalars> zero(count);
      std::memset((void *)zero.data(), 0, count * sizeof(SolverScalars));
      e = hipMemcpy(B->scal.ptr, zero.data(), count * sizeof(SolverScalars), hipMemcpyHostToDevice);
    }
  if (e == hipSuccess)
    {
      std::vector<double> zero(2 * count, 0.0);
      e = hipMemcpy(B->res.ptr, zero.data(), 2 * count * sizeof(double), hipMemcpyHostToDevice);
    }
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->host_scal, count * sizeof(SolverScalars), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->host_res, 2 * count * sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipHostMalloc((void **)&B->host_outer, count * sizeof(BatchOuter), hipHostMallocDefault);
  if (e == hipSuccess)
    {
      int dev = 0, khz = 0;
      e = hipGetDevice(&dev);
      if (e == hipSuccess)
        e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
      B->clock_hz = 1e3 * (khz > 0 ? khz : 100000);
    }
  if (e != hipSuccess)
    {
      delete B;
      return fail(MPDG_ERR_HIP, std::string("batch set-up: ") + hipGetErrorString(e));
    }
  *out = B;
  return MPDG_OK;
}

mpdg_status mpdg_heat_batch_destroy(mpdg_heat_batch *b)
{
  if (!b)
    return MPDG_OK;
  if (b->stream)
    (void)hipStreamSynchronize(b->stream);
  else
    (void)hipDeviceSynchronize();
  delete b;
  return MPDG_OK;
}

int64_t mpdg_heat_batch_n_dofs(const mpdg_heat_batch *b, int member)
{
  if (!b || member < 0 || member >= b->count)
    return -1;
  return b->members[member]->N;
}

mpdg_status mpdg_heat_batch_set_dt(mpdg_heat_batch *b, double dt)
{
  if (!b)
    return fail(MPDG_ERR_ARGUMENT, "null batch");
  for (auto *m : b->members)
    if (mpdg_status st = mpdg_heat_set_dt(m, dt))
      return st;
  // operator descriptors carry dt-dependent constants: rebuild them and the graphs
  if (b->graph64)
    (void)hipGraphExecDestroy(b->graph64);
  if (b->graph32)
    (void)hipGraphExecDestroy(b->graph32);
  if (b->graph_mixed)
    (void)hipGraphExecDestroy(b->graph_mixed);
  b->graph64 = b->graph32 = b->graph_mixed = nullptr;
  if (mpdg_status st = build_groups<double>(b, b->g64, b->vmult_slots64, b->P.ptr, b->Q.ptr))
    return st;
  if (mpdg_status st = build_groups<float>(b, b->g32, b->vmult_slots32, b->P32.ptr, b->Q32.ptr))
    return st;
  const size_t need = (size_t)std::max(b->vmult_slots64, b->vmult_slots32) + 2 * (size_t)b->nblocks;
  MPDG_HIP(b->partials.alloc(need));
  return MPDG_OK;
}

mpdg_status mpdg_heat_batch_solve(mpdg_heat_batch *b, double *const *x, const double *const *rhs,
                                  const mpdg_solver_params *params, mpdg_batch_report *reports)
{
  if (!b || !x || !rhs)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  if (b->g64.empty())
    return fail(MPDG_ERR_ZERO_TIME_INCREMENT, "call mpdg_heat_batch_set_dt first");
  mpdg_solver_params prm;
  if (params)
    prm = *params;
  else
    mpdg_solver_params_default(&prm);
  if (prm.solver_type != MPDG_SOLVER_CG)
    return fail(MPDG_ERR_UNSUPPORTED, "batched solves use CG");
  if (prm.check_interval < 0 || prm.check_interval > kMaxCheckInterval)
    return fail(MPDG_ERR_ARGUMENT, "check_interval must be in 0..64 (0: automatic)");
  const int S = b->count;
  const bool mixed = prm.mixed_precision != 0;
  const int steps = prm.check_interval > 0 ? prm.check_interval : 8;
  if (b->own_stream)
    MPDG_HIP(hipDeviceSynchronize()); // inputs were produced on other streams
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  std::vector<mpdg_batch_report> rep(S);
  for (auto &r : rep)
    {
      std::memset(&r, 0, sizeof(r));
      r.status = 1;
      r.seconds = -1.0;
    }

  // Jacobi diagonals (JacobiPreconditioner::update per member) and inputs
  for (int s = 0; s < S; ++s)
    {
      mpdg_heat *m = b->members[s];
      if (mpdg_status st = diagonal_inverse(m, b->D.ptr + b->off[s]))
        return st;
      MPDG_HIP(hipMemcpyAsync(b->B.ptr + b->off[s], rhs[s], m->N * sizeof(double),
                              hipMemcpyDeviceToDevice, b->stream));
      MPDG_HIP(hipMemcpyAsync(b->X.ptr + b->off[s], x[s], m->N * sizeof(double),
                              hipMemcpyDeviceToDevice, b->stream));
    }
  std::vector<double> relv(S, prm.rel_tolerance);
  MPDG_HIP(hipMemcpyAsync(b->rel.ptr, relv.data(), S * sizeof(double), hipMemcpyHostToDevice, b->stream));

  // opsx was built for the FP64 group layout; its block offsets equal g64's
  // (same degree and precision), so the same dev_block_off serves both.

  if (!mixed)
    {
      if (!b->graph64 || b->graph_steps64 != steps)
        {
          if (b->graph64)
            (void)hipGraphExecDestroy(b->graph64);
          b->graph64 = nullptr;
          if (mpdg_status st = batch_graph<double>(b, b->g64, steps, b->vmult_slots64, b->X.ptr,
                                                   b->R.ptr, b->P.ptr, b->Q.ptr, b->D.ptr, &b->graph64))
            return st;
          b->graph_steps64 = steps;
        }
      // r = b - A x0 for all members, then the PCG start
      MPDG_HIP(launch_batch_begin(b->outer.ptr, b->scal.ptr, b->ogate.ptr, S, b->stream));
      for (auto &g : b->g64)
        MPDG_HIP(launch_vmult_batch<double>(g.degree, b->coef, b->vel, g.variant64, g.opsx.ptr,
                                          g.dev_block_off.ptr,
                                            g.nops, g.plan, nullptr, nullptr, b->stream));
      MPDG_HIP(launch_batch_cg_init<double>(b->B.ptr, b->Q.ptr, b->R.ptr, b->P.ptr, b->D.ptr,
                                            b->seg.ptr, S, b->nblocks,
                                            b->partials.ptr + b->vmult_slots64, b->scal.ptr,
                                            b->rel.ptr, prm.abs_tolerance, prm.max_iterations,
                                            nullptr, b->stream));
      for (;;)
        {
          if (mpdg_status st = read_batch_scalars(b))
            return st;
          bool all = true;
          for (int s = 0; s < S; ++s)
            {
              const SolverScalars &sc = b->host_scal[s];
              if (sc.done && rep[s].seconds < 0)
                {
                  rep[s].seconds = 0.0; // device time stamps below
                  rep[s].status = sc.status == 0 ? 0 : 1;
                  rep[s].iterations = sc.iteration;
                  rep[s].initial_residual = sc.res0;
                  rep[s].final_residual = sc.res;
                }
              all = all && sc.done;
            }
          if (all)
            break;
          MPDG_HIP(hipGraphLaunch(b->graph64, b->stream));
        }
      // per-member completion: device wall clock at the stop, from the start stamp
      MPDG_HIP(hipMemcpyAsync(b->host_outer, b->outer.ptr, S * sizeof(BatchOuter),
                              hipMemcpyDeviceToHost, b->stream));
      MPDG_HIP(hipStreamSynchronize(b->stream));
      for (int s = 0; s < S; ++s)
        rep[s].seconds =
          (double)(b->host_scal[s].t_done - b->host_outer[s].t_start) / b->clock_hz;
    }
  else
    {
      if (!b->graph_mixed || b->graph_steps_mixed != steps || b->graph_abs != prm.abs_tolerance ||
          b->graph_rel != prm.rel_tolerance || b->graph_inner != prm.inner_rel_tolerance ||
          b->graph_max_it != prm.max_iterations)
        {
          if (b->graph_mixed)
            (void)hipGraphExecDestroy(b->graph_mixed);
          b->graph_mixed = nullptr;
          if (mpdg_status st = batch_mixed_graph(b, steps, prm, &b->graph_mixed))
            return st;
          b->graph_steps_mixed = steps;
          b->graph_abs = prm.abs_tolerance;
          b->graph_rel = prm.rel_tolerance;
          b->graph_inner = prm.inner_rel_tolerance;
          b->graph_max_it = prm.max_iterations;
        }
      MPDG_HIP(launch_to_f32(b->D32.ptr, b->D.ptr, b->total, b->stream));
      MPDG_HIP(launch_batch_begin(b->outer.ptr, b->scal.ptr, b->ogate.ptr, S, b->stream));
      // every chunk moves each unfinished member by at least one outer step
      // or `steps` inner iterations; the cap only guards against a stop test
      // that never fires
      const long max_chunks = 2L * (prm.max_iterations / steps + 2) + 64;
      bool all = false;
      for (long chunk = 0; chunk < max_chunks && !all; ++chunk)
        {
          MPDG_HIP(hipGraphLaunch(b->graph_mixed, b->stream));
          MPDG_HIP(hipMemcpyAsync(b->host_outer, b->outer.ptr, S * sizeof(BatchOuter),
                                  hipMemcpyDeviceToHost, b->stream));
          MPDG_HIP(hipStreamSynchronize(b->stream));
          all = true;
          for (int s = 0; s < S; ++s)
            all = all && b->host_outer[s].phase == kPhaseDone;
        }
      for (int s = 0; s < S; ++s)
        {
          const BatchOuter &o = b->host_outer[s];
          rep[s].iterations = o.inner_total;
          rep[s].outer_iterations = o.outer;
          rep[s].initial_residual = o.res0;
          rep[s].final_residual = o.res;
          rep[s].status = o.phase == kPhaseDone && o.status == 0 ? 0 : 1;
          if (o.phase == kPhaseDone)
            rep[s].seconds = (double)(o.t_done - o.t_start) / b->clock_hz;
        }
    }
  for (int s = 0; s < S; ++s)
    MPDG_HIP(hipMemcpyAsync(x[s], b->X.ptr + b->off[s], b->members[s]->N * sizeof(double),
                            hipMemcpyDeviceToDevice, b->stream));
  MPDG_HIP(hipStreamSynchronize(b->stream));
  bool ok = true;
  for (int s = 0; s < S; ++s)
    ok = ok && rep[s].status == 0;
  const double t_end = now();
  for (int s = 0; s < S; ++s)
    if (rep[s].seconds < 0)
      rep[s].seconds = t_end;
  if (reports)
    std::copy(rep.begin(), rep.end(), reports);
  return ok ? MPDG_OK : fail(MPDG_ERR_NOT_CONVERGED, "at least one batch member did not converge");
}

} // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free FE L2 norm (mpdg_heat_l2_norm_noalloc)
// ---------------------------------------------------------------------------
// The reduction scratch of the L2 norm kernel (one partial per workgroup, the
// last-workgroup ticket) lives in one file-scope device array per GPU instead
// of the handle's lazily allocated workspace. Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the symbol):
// each device's address comes from hipGetSymbolAddress, once per device with
// that device current, and only that address is used.
namespace
{
  constexpr int kNoAllocPartials = 1 << 19; // workgroups of one norm launch (4 MiB of partials)
  constexpr int kNoAllocTicket = kNoAllocPartials;      // slot of the unsigned ticket
  constexpr int kNoAllocDoubles = kNoAllocPartials + 32; // + ticket, padded to 256 bytes
  constexpr int kMaxDevices = 64;
} // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace
{
  std::mutex g_scratch_mutex;
  double *g_scratch_ptr[kMaxDevices] = {};

  // Device address of this device's scratch array (set-up on first use).
  mpdg_status scratch_address(int device, double **out)
  {
    if (device < 0 || device >= kMaxDevices)
      return fail(MPDG_ERR_UNSUPPORTED, "allocation-free entry points: device index above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device])
      {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
          return fail(MPDG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                      hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double *>(ptr);
      }
    *out = g_scratch_ptr[device];
    return MPDG_OK;
  }

  // The ticket must start at 0 (the kernel resets it on exit; this clears a
  // ticket a failed launch may have left). A kernel rather than a memset, so
  // that the call can be captured into a graph.
  __global__ void clear_ticket_kernel(unsigned *ticket) { *ticket = 0u; }

  __global__ void sqrt_in_place_kernel(double *v) { *v = sqrt(*v); }
} // namespace

extern "C" mpdg_status mpdg_heat_l2_norm_noalloc(mpdg_heat *h, const double *v, double *norm)
{
  if (!h || !v || !norm)
    return fail(MPDG_ERR_ARGUMENT, "null argument");
  const int blocks = l2_norm_blocks(h->desc.degree, h->desc.n_cells[0], h->desc.n_cells[1],
                                    h->desc.n_cells[2]);
  if (blocks > kNoAllocPartials)
    return fail(MPDG_ERR_UNSUPPORTED,
                "mpdg_heat_l2_norm_noalloc: the mesh needs " + std::to_string(blocks) +
                  " reduction workgroups; the static scratch holds " +
                  std::to_string(kNoAllocPartials));
  int device = 0;
  MPDG_HIP(hipGetDevice(&device));
  if (device != h->device)
    return fail(MPDG_ERR_ARGUMENT, "mpdg_heat_l2_norm_noalloc: the handle's device is not current");
  double *base = nullptr;
  if (mpdg_status s = scratch_address(device, &base))
    return s;
  unsigned *ticket = reinterpret_cast<unsigned *>(base + kNoAllocTicket);
  clear_ticket_kernel<<<1, 1, 0, h->stream>>>(ticket);
  MPDG_HIP(hipGetLastError());
  MPDG_HIP(launch_l2_norm_sq(h->desc.degree, h->prm64, v, base, ticket, norm, h->stream));
  sqrt_in_place_kernel<<<1, 1, 0, h->stream>>>(norm);
  MPDG_HIP(hipGetLastError());
  return MPDG_OK;
}
