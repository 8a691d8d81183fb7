// CPU oracle for the MeltPoolDG diffuse heat operator hot path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU baseline
// timed beside it. The product library never links, loads or calls this file.
// See heat_oracle.h for the list of reference functions restated here.
//
// The cell kernel is deliberately the textbook form: values and gradients are
// interpolated straight from the nodal coefficients with the 1D value matrix S
// and derivative matrix D (no collocation transform, no even-odd split), so it
// is an independent code path from the GPU kernels.
#include "heat_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace
{
  constexpr long double PI_L = 3.141592653589793238462643383279502884L;

  // Legendre P_n and P_n' on [-1,1] by the three-term recurrence.
  void legendre(int n, long double x, long double &p, long double &dp)
  {
    long double p0 = 1.0L, p1 = x;
    if (n == 0)
      {
        p = 1.0L;
        dp = 0.0L;
        return;
      }
    for (int k = 2; k <= n; ++k)
      {
        long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1.0L);
  }

  // Gauss-Legendre points/weights on [0,1], ascending (deal.II QGauss(n)).
  void gauss_rule(int n, long double *x, long double *w)
  {
    for (int i = 0; i < n; ++i)
      {
        long double z = std::cos(PI_L * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; ++it)
          {
            long double p, dp;
            legendre(n, z, p, dp);
            long double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-30L)
              break;
          }
        long double p, dp;
        legendre(n, z, p, dp);
        // z runs from +1 down to -1 as i grows; store ascending on [0,1]
        x[n - 1 - i] = 0.5L * (1.0L + z);
        w[n - 1 - i] = 1.0L / ((1.0L - z * z) * dp * dp); // (2/((1-z^2)P'^2)) / 2
      }
  }

  // Gauss-Lobatto points of degree p on [0,1]: 0, roots of P_p', 1.
  void gll_nodes(int p, long double *x)
  {
    x[0] = 0.0L;
    x[p] = 1.0L;
    for (int i = 1; i < p; ++i)
      {
        long double z = -std::cos(PI_L * i / p);
        for (int it = 0; it < 100; ++it)
          {
            long double P, dP;
            legendre(p, z, P, dP);
            // (1-z^2) P'' = 2 z P' - p(p+1) P
            long double d2P = (2.0L * z * dP - p * (p + 1.0L) * P) / (1.0L - z * z);
            long double dz = dP / d2P;
            z -= dz;
            if (std::fabs(dz) < 1e-30L)
              break;
          }
        x[i] = 0.5L * (1.0L + z);
      }
  }

  long double lagrange(int p, const long double *nodes, int i, long double x)
  {
    long double v = 1.0L;
    for (int j = 0; j <= p; ++j)
      if (j != i)
        v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    return v;
  }

  long double lagrange_deriv(int p, const long double *nodes, int i, long double x)
  {
    long double s = 0.0L;
    for (int k = 0; k <= p; ++k)
      {
        if (k == i)
          continue;
        long double t = 1.0L / (nodes[i] - nodes[k]);
        for (int j = 0; j <= p; ++j)
          if (j != i && j != k)
            t *= (x - nodes[j]) / (nodes[i] - nodes[j]);
        s += t;
      }
    return s;
  }

  constexpr int MAXN = ORACLE_MAX_DEGREE + 1;

  struct Tables
  {
    int p, n;
    double S[MAXN][MAXN]; // S[q][i] = phi_i(x_q)
    double D[MAXN][MAXN]; // D[q][i] = phi_i'(x_q)
    double W[MAXN];       // Gauss weights on [0,1]
    double X[MAXN];       // Gauss points on [0,1]
  };

  Tables make_tables(int p)
  {
    Tables t;
    t.p = p;
    t.n = p + 1;
    long double nodes[MAXN], gx[MAXN], gw[MAXN];
    gll_nodes(p, nodes);
    gauss_rule(p + 1, gx, gw);
    for (int q = 0; q <= p; ++q)
      {
        t.W[q] = (double)gw[q];
        t.X[q] = (double)gx[q];
        for (int i = 0; i <= p; ++i)
          {
            t.S[q][i] = (double)lagrange(p, nodes, i, gx[q]);
            t.D[q][i] = (double)lagrange_deriv(p, nodes, i, gx[q]);
          }
      }
    return t;
  }

  struct Mesh
  {
    int n[3];
    int p;
    int64_t N1[3];
    int64_t N;
    double h[3];
    double lo[3];
    // sin(2 pi x) and sin(pi x) at the Gauss coordinates of each direction,
    // index c (p + 1) + q: the same std::sin of the same argument the
    // quadrature-point physics evaluates, computed once per coordinate
    std::vector<double> sin2pi[3], sinpi[3];
  };

  Mesh make_mesh(const oracle_problem *pb)
  {
    Mesh m;
    m.p = pb->degree;
    m.N = 1;
    for (int d = 0; d < 3; ++d)
      {
        m.n[d] = pb->n[d];
        m.N1[d] = (int64_t)pb->n[d] * pb->degree + 1;
        m.N *= m.N1[d];
        m.h[d] = (pb->hi[d] - pb->lo[d]) / pb->n[d];
        m.lo[d] = pb->lo[d];
      }
    if (pb->degree >= 1 && pb->degree <= ORACLE_MAX_DEGREE)
      {
        const Tables t = make_tables(pb->degree);
        const double pi = (double)PI_L, two_pi = 2.0 * (double)PI_L;
        for (int d = 0; d < 3; ++d)
          {
            m.sin2pi[d].resize((size_t)std::max(m.n[d], 0) * t.n);
            m.sinpi[d].resize(m.sin2pi[d].size());
            for (int c = 0; c < m.n[d]; ++c)
              for (int q = 0; q < t.n; ++q)
                {
                  const double xq = m.lo[d] + (c + t.X[q]) * m.h[d];
                  m.sin2pi[d][(size_t)c * t.n + q] = std::sin(two_pi * xq);
                  m.sinpi[d][(size_t)c * t.n + q] = std::sin(pi * xq);
                }
          }
      }
    return m;
  }

  bool valid(const oracle_problem *pb)
  {
    if (!pb || pb->degree < 1 || pb->degree > ORACLE_MAX_DEGREE)
      return false;
    for (int d = 0; d < 3; ++d)
      if (pb->n[d] < 1 || !(pb->hi[d] > pb->lo[d]))
        return false;
    if (!(pb->dt != 0.0))
      return false;
    if (pb->material < 0 || pb->material > 3)
      return false;
    if ((pb->material == 1 || pb->material == 3) && !pb->heaviside)
      return false;
    if (pb->velocity < 0 || pb->velocity > 3 || (pb->velocity == 3 && !pb->velocity_field))
      return false;
    // no face may be both Neumann and Robin (heat_diffuse_operator.cpp:744-747)
    if (pb->neumann_faces & (pb->convection_faces | pb->radiation_faces) & 0x3Fu)
      return false;
    if (pb->material >= 2)
      {
        if (pb->consistent_with_evaporation)
          return false;
        if (pb->solid_liquid_sharp ? !(pb->liquidus >= pb->solidus) :
                                     !(pb->liquidus > pb->solidus))
          return false;
        // the reference rejects the apparent capacity with volume-specific
        // interpolation and needs liquidus > solidus for it
        if (pb->latent_heat > 0.0 &&
            (pb->volume_specific || !(pb->liquidus > pb->solidus) || pb->apparent_capacity < 0 ||
             pb->apparent_capacity > 2))
          return false;
      }
    return true;
  }

  inline bool uses_heaviside(const oracle_problem *pb)
  {
    return pb->material == 1 || pb->material == 3;
  }

  inline bool temperature_dependent(const oracle_problem *pb)
  {
    return pb->material >= 2;
  }

  // the tangent of a temperature-dependent material reads T and T_old
  bool tangent_ready(const oracle_problem *pb)
  {
    return !temperature_dependent(pb) || (pb->temperature && pb->temperature_old);
  }

  inline bool is_constrained(const oracle_problem *pb, const Mesh &m, int64_t ix, int64_t iy,
                             int64_t iz)
  {
    const unsigned f = pb->dirichlet_faces;
    const int64_t c[3] = {ix, iy, iz};
    for (int d = 0; d < 3; ++d)
      {
        if ((f >> (2 * d)) & 1u)
          if (c[d] == 0)
            return true;
        if ((f >> (2 * d + 1)) & 1u)
          if (c[d] == m.N1[d] - 1)
            return true;
      }
    return false;
  }

  // Material::compute_parameters for gas / gas_liquid
  // (include/meltpooldg/core/material.templates.hpp:223-295, 510-539;
  //  LevelSet::Tools::interpolate include/meltpooldg/level_set/level_set_tools.hpp:40-43;
  //  rho*cp combination: source/heat/heat_diffuse_operator.cpp:1167-1210).
  inline void material_at(const oracle_problem *pb, double H, double &rho_cp, double &k)
  {
    if (pb->material == 0)
      {
        rho_cp = pb->gas_rho * pb->gas_cp;
        k = pb->gas_k;
        return;
      }
    const double w = pb->sharp ? (H > 0.5 ? 1.0 : 0.0) : H;
    k = (1. - w) * pb->gas_k + w * pb->liq_k;
    if (pb->volume_specific)
      {
        rho_cp = (1. - w) * (pb->gas_rho * pb->gas_cp) + w * (pb->liq_rho * pb->liq_cp);
      }
    else
      {
        const double cp = (1. - w) * pb->gas_cp + w * pb->liq_cp;
        const double rho = pb->consistent_with_evaporation ?
                             1. / ((1. - H) / pb->gas_rho + H / pb->liq_rho) :
                             (1. - w) * pb->gas_rho + w * pb->liq_rho;
        rho_cp = rho * cp;
      }
  }

  // Material state at one quadrature point, with the temperature derivatives
  // the tangent needs.
  struct MatState
  {
    double rho_cp, k, d_rho_cp, d_k, rho, cp, d_rho, d_cp;
  };

  // compute_temperature_dependent_solid_fraction (material.templates.hpp:740-756):
  // mushy zone: (liquidus - T) / (liquidus - solidus) limited to [0, 1]; sharp:
  // 1 below the solidus temperature, else 0.
  inline double solid_fraction(const oracle_problem *pb, double T, double inv_mushy)
  {
    if (pb->solid_liquid_sharp)
      return T < pb->solidus ? 1.0 : 0.0;
    const double f = (pb->liquidus - T) * inv_mushy;
    return f > 1.0 ? 1.0 : (f < 0.0 ? 0.0 : f);
  }

  // C1 ramp from (x1, y1) to (x2, y2): quadratic with zero slope at either end
  // over a width delta, linear in between with the slope s that makes the
  // pieces meet, s = (y2 - y1) / (x2 - x1 - delta) (the quadratic-linear-
  // quadratic profile of source/heat/apparent_capacity.cpp:42-155).
  inline void qlq_ramp(double x, double x1, double x2, double y1, double y2, double delta,
                       double &v, double &dv)
  {
    const double s = (y2 - y1) / (x2 - x1 - delta);
    if (x < x1 + delta)
      {
        const double d = x - x1;
        v = y1 + 0.5 * s * d * d / delta;
        dv = s * d / delta;
      }
    else if (x > x2 - delta)
      {
        const double d = x2 - x;
        v = y2 - 0.5 * s * d * d / delta;
        dv = s * d / delta;
      }
    else
      {
        v = y1 + s * (x - x1 - 0.5 * delta);
        dv = s;
      }
  }

  // Heat::ApparentCapacity (source/heat/apparent_capacity.cpp:157-311): the
  // latent heat L spread over the mushy zone as an extra c_p, zero outside
  // 0 < fs < 1, and its temperature derivative:
  //   poly4_bell  L (2/dT) b(xi), b(x) = 15/16 (1 - x^2)^2, xi = 1 - 2 fs;
  //   constant    L / dT;
  //   qlq         a C1 ramp from 0 at the solidus to 2 L / dT at the mid
  //               temperature and back to 0 at the liquidus (width dT / 6).
  inline void apparent_capacity(const oracle_problem *pb, double fs, double &c, double &dc)
  {
    c = dc = 0.0;
    if (!(pb->latent_heat > 0.0) || fs <= 0.0 || fs >= 1.0)
      return;
    const double Ts = pb->solidus, Tl = pb->liquidus, dT = Tl - Ts, L = pb->latent_heat;
    if (pb->apparent_capacity == 1)
      c = L / dT;
    else if (pb->apparent_capacity == 0)
      {
        const double r = 2.0 / dT, xi = 1.0 - 2.0 * fs, q = 1.0 - xi * xi;
        c = L * r * (15.0 / 16.0) * q * q;
        dc = L * r * r * (15.0 / 4.0) * xi * (xi * xi - 1.0);
      }
    else
      {
        const double T = Tl - fs * dT, Tm = 0.5 * (Ts + Tl), peak = 2.0 * L / dT;
        if (T <= Tm)
          qlq_ramp(T, Ts, Tm, 0.0, peak, dT / 6.0, c, dc);
        else
          qlq_ramp(T, Tm, Tl, peak, 0.0, dT / 6.0, c, dc);
      }
  }

  // Liquid/solid blend of one property by the cubic smoothstep of the solid
  // fraction fs, and its temperature derivative (compute_solid_liquid_phases_property
  // and compute_temperature_derivative_of_solid_liquid_phases_property,
  // material.templates.hpp:540-600; LevelSet::Tools::interpolate_cubic(_derivative),
  // level_set_tools.hpp:66-90): pure phases take the phase value and derivative 0.
  inline void solid_liquid(double fs, double inv_mushy, double liquid, double solid, double &v,
                           double &dv)
  {
    if (fs == 0.0 || fs == 1.0)
      {
        v = fs == 0.0 ? liquid : solid;
        dv = 0.0;
        return;
      }
    v = liquid + (solid - liquid) * (fs * fs * (3.0 - 2.0 * fs));
    dv = -inv_mushy * (solid - liquid) * (6.0 * fs * (1.0 - fs));
  }

  // Material::compute_parameters_internal for all four diffuse material types
  // (material.templates.hpp:223-495) and the rho_cp combination of
  // get_material_parameters_and_derivatives (heat_diffuse_operator.cpp:1167-1210):
  // rho_cp = rho cp with d(rho cp) = d_cp rho + d_rho cp, or the blended
  // volume-specific capacity and its derivative.
  MatState material_state(const oracle_problem *pb, double H, double T)
  {
    MatState s;
    if (!temperature_dependent(pb))
      {
        material_at(pb, H, s.rho_cp, s.k);
        const double w = pb->sharp ? (H > 0.5 ? 1.0 : 0.0) : H;
        s.rho = pb->material == 0 ? pb->gas_rho :
                (pb->consistent_with_evaporation ? 1. / ((1. - H) / pb->gas_rho + H / pb->liq_rho) :
                                                   (1. - w) * pb->gas_rho + w * pb->liq_rho);
        s.cp = pb->material == 0 ? pb->gas_cp : (1. - w) * pb->gas_cp + w * pb->liq_cp;
        s.d_rho_cp = s.d_k = s.d_rho = s.d_cp = 0.0;
        return s;
      }
    const double inv_mushy = 1.0 / (pb->liquidus - pb->solidus);
    const double fs = solid_fraction(pb, T, inv_mushy);
    double vc, dvc;
    solid_liquid(fs, inv_mushy, pb->liq_k, pb->sol_k, s.k, s.d_k);
    solid_liquid(fs, inv_mushy, pb->liq_cp, pb->sol_cp, s.cp, s.d_cp);
    // compute_solid_liquid_phases_specific_heat_capacity (material.templates.hpp:
    // 563-577, 600-617): c_p and its derivative carry the apparent capacity
    double ac, dac;
    apparent_capacity(pb, fs, ac, dac);
    s.cp += ac;
    s.d_cp += dac;
    solid_liquid(fs, inv_mushy, pb->liq_rho, pb->sol_rho, s.rho, s.d_rho);
    solid_liquid(fs, inv_mushy, pb->liq_rho * pb->liq_cp, pb->sol_rho * pb->sol_cp, vc, dvc);
    if (pb->material == 3)
      {
        // gas / (liquid-solid) blend by the heaviside; the derivatives carry the
        // blend weight (material.templates.hpp:378-495, 629-709)
        const double w = pb->sharp ? (H > 0.5 ? 1.0 : 0.0) : H;
        s.k = (1. - w) * pb->gas_k + w * s.k;
        s.d_k *= w;
        s.cp = (1. - w) * pb->gas_cp + w * s.cp;
        s.d_cp *= w;
        s.rho = (1. - w) * pb->gas_rho + w * s.rho;
        s.d_rho *= w;
        vc = (1. - w) * (pb->gas_rho * pb->gas_cp) + w * vc;
        dvc *= w;
      }
    if (pb->volume_specific)
      {
        s.rho_cp = vc;
        s.d_rho_cp = dvc;
      }
    else
      {
        s.rho_cp = s.rho * s.cp;
        s.d_rho_cp = s.d_cp * s.rho + s.d_rho * s.cp;
      }
    return s;
  }

  // Prescribed velocity at a quadrature point: 1 = the vortex field (analytic at
  // q), 2 = a constant field (what VectorTools::interpolate of a constant
  // prescribed_velocity function gives at any point).
  // (the sines at the Gauss coordinate c (p + 1) + q of each direction come
  // from the mesh's tables: ix, iy index them)
  inline void velocity_at(const oracle_problem *pb, const Mesh &m, size_t ix, size_t iy,
                          double v[3])
  {
    if (pb->velocity == 2)
      {
        v[0] = pb->velocity_value[0];
        v[1] = pb->velocity_value[1];
        v[2] = pb->velocity_value[2];
        return;
      }
    const double sx = m.sinpi[0][ix], sy = m.sinpi[1][iy];
    v[0] = -sx * sx * m.sin2pi[1][iy];
    v[1] = m.sin2pi[0][ix] * sy * sy;
    v[2] = pb->w0;
  }

  // Nodal velocity (velocity == 3) at the Gauss points of one cell: each
  // component's nodal values read plain and interpolated with S in x, y, z
  // (vel_eval.read_dof_values_plain + evaluate(values)); defined after the
  // cell helpers it uses.
  struct CellGeom;
  void nodal_velocity_q(const oracle_problem *pb, const struct Tables &t, const struct Mesh &m,
                        const CellGeom &g, double (*Vq)[MAXN * MAXN * MAXN]);

  // 1D contraction along one tensor direction of a local n^3 array
  // (index (c*n+b)*n+a, a = x fastest). out = M in along `dir`, or M^T.
  template <bool transpose, int n, int dir>
  void apply_dir_n(const double (*M)[MAXN], const double *in, double *out)
  {
    constexpr int stride = dir == 0 ? 1 : (dir == 1 ? n : n * n);
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const int idx[3] = {a, b, c};
            const int o = idx[dir];
            const int base = (c * n + b) * n + a - o * stride;
            double s = 0.0;
            for (int r = 0; r < n; ++r)
              s += (transpose ? M[r][o] : M[o][r]) * in[base + r * stride];
            out[(c * n + b) * n + a] = s;
          }
  }

  // the same loops with n and the direction compile-time constants (the
  // compiler unrolls them; each sum runs over r in the same order)
  template <bool transpose>
  void apply_dir(const double (*M)[MAXN], int n, int dir, const double *in, double *out)
  {
    switch (n)
      {
        case 2: return dir == 0 ? apply_dir_n<transpose, 2, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 2, 1>(M, in, out) :
                                  apply_dir_n<transpose, 2, 2>(M, in, out);
        case 3: return dir == 0 ? apply_dir_n<transpose, 3, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 3, 1>(M, in, out) :
                                  apply_dir_n<transpose, 3, 2>(M, in, out);
        case 4: return dir == 0 ? apply_dir_n<transpose, 4, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 4, 1>(M, in, out) :
                                  apply_dir_n<transpose, 4, 2>(M, in, out);
        case 5: return dir == 0 ? apply_dir_n<transpose, 5, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 5, 1>(M, in, out) :
                                  apply_dir_n<transpose, 5, 2>(M, in, out);
        case 6: return dir == 0 ? apply_dir_n<transpose, 6, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 6, 1>(M, in, out) :
                                  apply_dir_n<transpose, 6, 2>(M, in, out);
        case 7: return dir == 0 ? apply_dir_n<transpose, 7, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 7, 1>(M, in, out) :
                                  apply_dir_n<transpose, 7, 2>(M, in, out);
        case 8: return dir == 0 ? apply_dir_n<transpose, 8, 0>(M, in, out) :
                       dir == 1 ? apply_dir_n<transpose, 8, 1>(M, in, out) :
                                  apply_dir_n<transpose, 8, 2>(M, in, out);
        default: break;
      }
    const int stride = dir == 0 ? 1 : (dir == 1 ? n : n * n);
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const int idx[3] = {a, b, c};
            const int o = idx[dir];
            const int base = (c * n + b) * n + a - o * stride;
            double s = 0.0;
            for (int r = 0; r < n; ++r)
              s += (transpose ? M[r][o] : M[o][r]) * in[base + r * stride];
            out[(c * n + b) * n + a] = s;
          }
  }

  struct CellGeom
  {
    int c[3];
  };

  // Coefficient fields of one cell, read plain (read_dof_values_plain): the
  // heaviside (materials 1, 3) and, for a temperature-dependent material, the
  // linearisation temperature and the old temperature.
  struct CellFields
  {
    double H[MAXN * MAXN * MAXN], T[MAXN * MAXN * MAXN], To[MAXN * MAXN * MAXN];
  };

  // values (and gradients, when G is given) of a nodal cell field at the Gauss points
  void evaluate_field(const Tables &t, const double *u, double *U, double (*G)[MAXN * MAXN * MAXN])
  {
    const int n = t.n;
    double tS[MAXN * MAXN * MAXN], tD[MAXN * MAXN * MAXN];
    double tSS[MAXN * MAXN * MAXN], tSD[MAXN * MAXN * MAXN], tDS[MAXN * MAXN * MAXN];
    apply_dir<false>(t.S, n, 0, u, tS);
    apply_dir<false>(t.S, n, 1, tS, tSS);
    apply_dir<false>(t.S, n, 2, tSS, U);
    if (G)
      {
        apply_dir<false>(t.D, n, 0, u, tD);
        apply_dir<false>(t.D, n, 1, tS, tSD);
        apply_dir<false>(t.S, n, 1, tD, tDS);
        apply_dir<false>(t.D, n, 2, tSS, G[2]);
        apply_dir<false>(t.S, n, 2, tSD, G[1]);
        apply_dir<false>(t.S, n, 2, tDS, G[0]);
      }
  }

  // tangent_local_cell_operation for one cell (heat_diffuse_operator.cpp:955-1086):
  // evaluate values|gradients, quadrature-point physics, integrate values|gradients.
  // With a temperature-dependent material the linearisation adds, at q,
  //   val      += dt^-1 d(rho_cp)/dT (T - T_old) u  (+ d(rho_cp)/dT (grad T . v) u)
  //   val_grad += dk/dT grad T u                                     (:1044-1056)
  void local_apply(const oracle_problem *pb, const Tables &t, const Mesh &m, const CellGeom &g,
                   const double *u, const CellFields &fl, double *y)
  {
    const int n = t.n;
    const int n3 = n * n * n;
    double tS[MAXN * MAXN * MAXN], tD[MAXN * MAXN * MAXN];
    double tSS[MAXN * MAXN * MAXN], tSD[MAXN * MAXN * MAXN], tDS[MAXN * MAXN * MAXN];
    double U[MAXN * MAXN * MAXN], Gx[MAXN * MAXN * MAXN], Gy[MAXN * MAXN * MAXN],
      Gz[MAXN * MAXN * MAXN], Hq[MAXN * MAXN * MAXN];

    // evaluate(values | gradients)
    apply_dir<false>(t.S, n, 0, u, tS);
    apply_dir<false>(t.D, n, 0, u, tD);
    apply_dir<false>(t.S, n, 1, tS, tSS);
    apply_dir<false>(t.D, n, 1, tS, tSD);
    apply_dir<false>(t.S, n, 1, tD, tDS);
    apply_dir<false>(t.S, n, 2, tSS, U);
    apply_dir<false>(t.D, n, 2, tSS, Gz);
    apply_dir<false>(t.S, n, 2, tSD, Gy);
    apply_dir<false>(t.S, n, 2, tDS, Gx);

    const bool td = temperature_dependent(pb);
    if (uses_heaviside(pb))
      evaluate_field(t, fl.H, Hq, nullptr); // heaviside_eval.evaluate(values)
    double TU[MAXN * MAXN * MAXN], TG[3][MAXN * MAXN * MAXN], ToU[MAXN * MAXN * MAXN];
    if (td)
      {
        evaluate_field(t, fl.T, TU, TG);      // T_new_eval.evaluate(values | gradients)
        evaluate_field(t, fl.To, ToU, nullptr); // T_old_eval.evaluate(values)
      }

    const double hx = m.h[0], hy = m.h[1], hz = m.h[2];
    const double jac_det = hx * hy * hz;
    const double dt_inv = 1. / pb->dt;

    double Vq[3][MAXN * MAXN * MAXN];
    if (pb->velocity == 3)
      nodal_velocity_q(pb, t, m, g, Vq);

    double V[MAXN * MAXN * MAXN], Fx[MAXN * MAXN * MAXN], Fy[MAXN * MAXN * MAXN],
      Fz[MAXN * MAXN * MAXN];
    for (int qz = 0; qz < n; ++qz)
      for (int qy = 0; qy < n; ++qy)
        for (int qx = 0; qx < n; ++qx)
          {
            const int q = (qz * n + qy) * n + qx;
            const double JxW = t.W[qx] * t.W[qy] * t.W[qz] * jac_det;

            const MatState ms = material_state(pb, uses_heaviside(pb) ? Hq[q] : 0.0,
                                               td ? TU[q] : 0.0);
            const double rho_cp = ms.rho_cp;
            double k = ms.k;
            if (pb->conductivity_field == 1)
              k += pb->k_amplitude * m.sin2pi[0][(size_t)g.c[0] * n + qx] *
                   m.sin2pi[1][(size_t)g.c[1] * n + qy] * m.sin2pi[2][(size_t)g.c[2] * n + qz];

            const double gx = Gx[q] / hx, gy = Gy[q] / hy, gz = Gz[q] / hz;
            double val = dt_inv * rho_cp * U[q];
            double fx = k * gx, fy = k * gy, fz = k * gz;
            double v[3] = {0.0, 0.0, 0.0};
            if (pb->velocity == 3)
              {
                v[0] = Vq[0][q];
                v[1] = Vq[1][q];
                v[2] = Vq[2][q];
                val += rho_cp * (gx * v[0] + gy * v[1] + gz * v[2]);
              }
            else if (pb->velocity != 0)
              {
                velocity_at(pb, m, (size_t)g.c[0] * n + qx, (size_t)g.c[1] * n + qy, v);
                val += rho_cp * (gx * v[0] + gy * v[1] + gz * v[2]);
              }
            if (td)
              {
                const double tgx = TG[0][q] / hx, tgy = TG[1][q] / hy, tgz = TG[2][q] / hz;
                val += dt_inv * ms.d_rho_cp * (TU[q] - ToU[q]) * U[q];
                fx += ms.d_k * tgx * U[q];
                fy += ms.d_k * tgy * U[q];
                fz += ms.d_k * tgz * U[q];
                if (pb->velocity != 0)
                  val += ms.d_rho_cp * (tgx * v[0] + tgy * v[1] + tgz * v[2]) * U[q];
              }
            V[q] = val * JxW;
            Fx[q] = fx * JxW / hx;
            Fy[q] = fy * JxW / hy;
            Fz[q] = fz * JxW / hz;
          }

    // integrate(values | gradients)
    double acc[MAXN * MAXN * MAXN];
    double a1[MAXN * MAXN * MAXN], a2[MAXN * MAXN * MAXN];
    // value part: S^T S^T S^T V
    apply_dir<true>(t.S, n, 2, V, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, acc);
    // x gradient: S^T(z) S^T(y) D^T(x)
    apply_dir<true>(t.S, n, 2, Fx, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.D, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      acc[i] += a1[i];
    apply_dir<true>(t.S, n, 2, Fy, a1);
    apply_dir<true>(t.D, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      acc[i] += a1[i];
    apply_dir<true>(t.D, n, 2, Fz, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      y[i] = acc[i] + a1[i];
  }

  inline int64_t node_index(const Mesh &m, const CellGeom &g, int a, int b, int c)
  {
    const int64_t ix = (int64_t)g.c[0] * m.p + a;
    const int64_t iy = (int64_t)g.c[1] * m.p + b;
    const int64_t iz = (int64_t)g.c[2] * m.p + c;
    return ix + m.N1[0] * (iy + m.N1[1] * iz);
  }

  void nodal_velocity_q(const oracle_problem *pb, const Tables &t, const Mesh &m,
                        const CellGeom &g, double (*Vq)[MAXN * MAXN * MAXN])
  {
    const int n = t.n;
    double loc[MAXN * MAXN * MAXN], t1[MAXN * MAXN * MAXN], t2[MAXN * MAXN * MAXN];
    for (int comp = 0; comp < 3; ++comp)
      {
        const double *v = pb->velocity_field + comp * m.N;
        for (int c = 0; c < n; ++c)
          for (int b = 0; b < n; ++b)
            for (int a = 0; a < n; ++a)
              loc[(c * n + b) * n + a] = v[node_index(m, g, a, b, c)];
        apply_dir<false>(t.S, n, 0, loc, t1);
        apply_dir<false>(t.S, n, 1, t1, t2);
        apply_dir<false>(t.S, n, 2, t2, Vq[comp]);
      }
  }

  // read_dof_values_plain of the heaviside and, for a temperature-dependent
  // material, of T and T_old
  void gather_fields(const oracle_problem *pb, const Mesh &m, const CellGeom &g, const double *T,
                     const double *To, CellFields &fl)
  {
    const int n = m.p + 1;
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const int l = (c * n + b) * n + a;
            const int64_t I = node_index(m, g, a, b, c);
            fl.H[l] = uses_heaviside(pb) ? pb->heaviside[I] : 0.0;
            fl.T[l] = temperature_dependent(pb) ? T[I] : 0.0;
            fl.To[l] = temperature_dependent(pb) ? To[I] : 0.0;
          }
  }

  template <typename Body>
  void for_cells_coloured(const Mesh &m, Body body)
  {
    for (int colour = 0; colour < 8; ++colour)
      {
        const int ox = colour & 1, oy = (colour >> 1) & 1, oz = (colour >> 2) & 1;
        const int mx = (m.n[0] - ox + 1) / 2, my = (m.n[1] - oy + 1) / 2,
                  mz = (m.n[2] - oz + 1) / 2;
        const int64_t total = (int64_t)mx * my * mz;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < total; ++i)
          {
            CellGeom g;
            g.c[0] = ox + 2 * (int)(i % mx);
            g.c[1] = oy + 2 * (int)((i / mx) % my);
            g.c[2] = oz + 2 * (int)(i / ((int64_t)mx * my));
            body(g);
          }
      }
  }

  void robin_vmult(const oracle_problem *pb, const Tables &t, const Mesh &m, const double *src,
                   double *dst);

  // HeatDiffuseMultiPhaseOperator::vmult (heat_diffuse_operator.cpp:179-197):
  // the cell loop, then the boundary-face loop of the Robin faces.
  void vmult_impl(const oracle_problem *pb, const Tables &t, const Mesh &m, const double *src,
                  double *dst)
  {
    const int n = t.n;
    std::fill(dst, dst + m.N, 0.0);
    for_cells_coloured(m, [&](const CellGeom &g) {
      double u[MAXN * MAXN * MAXN], y[MAXN * MAXN * MAXN];
      CellFields fl;
      gather_fields(pb, m, g, pb->temperature, pb->temperature_old, fl);
      for (int c = 0; c < n; ++c)
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a)
            {
              const int l = (c * n + b) * n + a;
              const int64_t I = node_index(m, g, a, b, c);
              const bool con = is_constrained(pb, m, (int64_t)g.c[0] * m.p + a,
                                              (int64_t)g.c[1] * m.p + b, (int64_t)g.c[2] * m.p + c);
              // read_dof_values: homogeneous constraints resolve to 0
              u[l] = con ? 0.0 : src[I];
            }
      local_apply(pb, t, m, g, u, fl, y);
      for (int c = 0; c < n; ++c)
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a)
            {
              const bool con = is_constrained(pb, m, (int64_t)g.c[0] * m.p + a,
                                              (int64_t)g.c[1] * m.p + b, (int64_t)g.c[2] * m.p + c);
              // distribute_local_to_global drops constrained rows
              if (!con)
                dst[node_index(m, g, a, b, c)] += y[(c * n + b) * n + a];
            }
    });
    robin_vmult(pb, t, m, src, dst);
  }

  // rhs_cell_loop restated (source/heat/heat_diffuse_operator.cpp:393-529)
  // without evaporation, the material evaluated at T_new(q): T_new, T_old and
  // the heat source read plain (constrained values included), at q
  //   val      = dt^-1 rho_cp (T_new - T_old) - f  (+ rho_cp grad T_new . u)
  //   val_grad = k grad T_new
  // submitted with the factor -1 and integrated (values | gradients).
  void local_rhs(const oracle_problem *pb, const Tables &t, const Mesh &m, const CellGeom &g,
                 const double *tn, const double *to, const double *fl, const double *Hloc,
                 double *y)
  {
    const int n = t.n;
    const int n3 = n * n * n;
    double tS[MAXN * MAXN * MAXN], tD[MAXN * MAXN * MAXN];
    double tSS[MAXN * MAXN * MAXN], tSD[MAXN * MAXN * MAXN], tDS[MAXN * MAXN * MAXN];
    double U[MAXN * MAXN * MAXN], Gx[MAXN * MAXN * MAXN], Gy[MAXN * MAXN * MAXN],
      Gz[MAXN * MAXN * MAXN], Hq[MAXN * MAXN * MAXN], Uo[MAXN * MAXN * MAXN],
      Fq[MAXN * MAXN * MAXN];

    // T_new_eval.evaluate(values | gradients)
    apply_dir<false>(t.S, n, 0, tn, tS);
    apply_dir<false>(t.D, n, 0, tn, tD);
    apply_dir<false>(t.S, n, 1, tS, tSS);
    apply_dir<false>(t.D, n, 1, tS, tSD);
    apply_dir<false>(t.S, n, 1, tD, tDS);
    apply_dir<false>(t.S, n, 2, tSS, U);
    apply_dir<false>(t.D, n, 2, tSS, Gz);
    apply_dir<false>(t.S, n, 2, tSD, Gy);
    apply_dir<false>(t.S, n, 2, tDS, Gx);
    // T_old_eval / heat_source_eval.evaluate(values)
    apply_dir<false>(t.S, n, 0, to, tS);
    apply_dir<false>(t.S, n, 1, tS, tSS);
    apply_dir<false>(t.S, n, 2, tSS, Uo);
    apply_dir<false>(t.S, n, 0, fl, tS);
    apply_dir<false>(t.S, n, 1, tS, tSS);
    apply_dir<false>(t.S, n, 2, tSS, Fq);
    if (uses_heaviside(pb))
      {
        apply_dir<false>(t.S, n, 0, Hloc, tS);
        apply_dir<false>(t.S, n, 1, tS, tSS);
        apply_dir<false>(t.S, n, 2, tSS, Hq);
      }

    const double hx = m.h[0], hy = m.h[1], hz = m.h[2];
    const double jac_det = hx * hy * hz;
    const double dt_inv = 1. / pb->dt;

    double Vq[3][MAXN * MAXN * MAXN];
    if (pb->velocity == 3)
      nodal_velocity_q(pb, t, m, g, Vq);

    double V[MAXN * MAXN * MAXN], Fx[MAXN * MAXN * MAXN], Fy[MAXN * MAXN * MAXN],
      Fz[MAXN * MAXN * MAXN];
    for (int qz = 0; qz < n; ++qz)
      for (int qy = 0; qy < n; ++qy)
        for (int qx = 0; qx < n; ++qx)
          {
            const int q = (qz * n + qy) * n + qx;
            const double JxW = t.W[qx] * t.W[qy] * t.W[qz] * jac_det;

            // get_material_parameters (heat_diffuse_operator.cpp:1133-1165) at T_new(q)
            const MatState ms = material_state(pb, uses_heaviside(pb) ? Hq[q] : 0.0, U[q]);
            const double rho_cp = ms.rho_cp;
            double k = ms.k;
            if (pb->conductivity_field == 1)
              k += pb->k_amplitude * m.sin2pi[0][(size_t)g.c[0] * n + qx] *
                   m.sin2pi[1][(size_t)g.c[1] * n + qy] * m.sin2pi[2][(size_t)g.c[2] * n + qz];

            const double gx = Gx[q] / hx, gy = Gy[q] / hy, gz = Gz[q] / hz;
            double val = dt_inv * rho_cp * (U[q] - Uo[q]) - Fq[q];
            if (pb->velocity != 0)
              {
                double v[3];
                if (pb->velocity == 3)
                  {
                    v[0] = Vq[0][q];
                    v[1] = Vq[1][q];
                    v[2] = Vq[2][q];
                  }
                else
                  velocity_at(pb, m, (size_t)g.c[0] * n + qx, (size_t)g.c[1] * n + qy, v);
                val += rho_cp * (gx * v[0] + gy * v[1] + gz * v[2]);
              }
            // submit_value(-1 val), submit_gradient(-1 val_grad)
            V[q] = -val * JxW;
            Fx[q] = -k * gx * JxW / hx;
            Fy[q] = -k * gy * JxW / hy;
            Fz[q] = -k * gz * JxW / hz;
          }

    double acc[MAXN * MAXN * MAXN];
    double a1[MAXN * MAXN * MAXN], a2[MAXN * MAXN * MAXN];
    apply_dir<true>(t.S, n, 2, V, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, acc);
    apply_dir<true>(t.S, n, 2, Fx, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.D, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      acc[i] += a1[i];
    apply_dir<true>(t.S, n, 2, Fy, a1);
    apply_dir<true>(t.D, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      acc[i] += a1[i];
    apply_dir<true>(t.D, n, 2, Fz, a1);
    apply_dir<true>(t.S, n, 1, a1, a2);
    apply_dir<true>(t.S, n, 0, a2, a1);
    for (int i = 0; i < n3; ++i)
      y[i] = acc[i] + a1[i];
  }

  // ---- Robin boundary faces ------------------------------------------------
  // Box face f = 2d + side (deal.II's colorized boundary id of hyper_rectangle):
  // normal direction d, tangential directions e1 < e2.
  struct BoxFace
  {
    int f, d, side, e1, e2;
  };

  BoxFace box_face(int f)
  {
    BoxFace b;
    b.f = f;
    b.d = f >> 1;
    b.side = f & 1;
    b.e1 = b.d == 0 ? 1 : 0;
    b.e2 = b.d == 2 ? 1 : 2;
    return b;
  }

  unsigned robin_faces(const oracle_problem *pb)
  {
    return (pb->convection_faces | pb->radiation_faces) & 0x3Fu;
  }

  // Lattice coordinates of local node (a1, a2) of face cell (c1, c2) on face b.
  inline void face_node(const Mesh &m, const BoxFace &b, int c1, int c2, int a1, int a2,
                        int64_t I[3])
  {
    I[b.d] = b.side ? m.N1[b.d] - 1 : 0;
    I[b.e1] = (int64_t)c1 * m.p + a1;
    I[b.e2] = (int64_t)c2 * m.p + a2;
  }

  enum FaceMode
  {
    kFaceTangent = 0,
    kFaceRhs = 1
  };

  // One face cell: face values of T (and of the increment u for the tangent) at
  // the (p+1)^2 Gauss points, the quadrature-point physics, and the integration
  // against the face shape functions (FEFaceEvaluation evaluate/integrate of
  // values). Node order a1 + n a2.
  //   tangent (heat_diffuse_operator.cpp:1101-1125):
  //     val = alpha u + 4 eps sigma T^3 u
  //   rhs (heat_diffuse_operator.cpp:748-776):
  //     val = -eps sigma (T^4 - T_inf^4) - alpha (T - T_inf)
  void face_local(const oracle_problem *pb, const Tables &t, const Mesh &m, const BoxFace &b,
                  int mode, const double *Tl, const double *ul, double *y)
  {
    const int n = t.n;
    const bool conv = (pb->convection_faces >> b.f) & 1u;
    const bool rad = (pb->radiation_faces >> b.f) & 1u;
    const double area = m.h[b.e1] * m.h[b.e2];
    double V[MAXN][MAXN];
    for (int q2 = 0; q2 < n; ++q2)
      for (int q1 = 0; q1 < n; ++q1)
        {
          double T = 0.0, u = 0.0;
          for (int a2 = 0; a2 < n; ++a2)
            for (int a1 = 0; a1 < n; ++a1)
              {
                const double phi = t.S[q1][a1] * t.S[q2][a2];
                T += phi * Tl[a2 * n + a1];
                if (mode == kFaceTangent)
                  u += phi * ul[a2 * n + a1];
              }
          double val = 0.0;
          if (mode == kFaceTangent)
            {
              if (conv)
                val += pb->convection_coefficient * u;
              if (rad)
                val += 4. * pb->emissivity * ORACLE_STEFAN_BOLTZMANN * (T * T * T) * u;
            }
          else
            {
              // Neumann: -q . n, n the outward normal of face f (-e_d or +e_d)
              if ((pb->neumann_faces >> b.f) & 1u)
                val -= (b.side ? 1.0 : -1.0) * pb->neumann_flux[b.f][b.d];
              const double Ti = pb->radiation_t_inf;
              if (rad)
                val -= pb->emissivity * ORACLE_STEFAN_BOLTZMANN *
                       ((T * T) * (T * T) - (Ti * Ti) * (Ti * Ti));
              if (conv)
                val -= pb->convection_coefficient * (T - pb->convection_t_inf);
            }
          V[q1][q2] = val * t.W[q1] * t.W[q2] * area;
        }
    for (int a2 = 0; a2 < n; ++a2)
      for (int a1 = 0; a1 < n; ++a1)
        {
          double s = 0.0;
          for (int q2 = 0; q2 < n; ++q2)
            for (int q1 = 0; q1 < n; ++q1)
              s += t.S[q1][a1] * t.S[q2][a2] * V[q1][q2];
          y[a2 * n + a1] = s;
        }
  }

  // The boundary-face part of MatrixFree::loop: every face cell of every Robin
  // face, faces ascending, face cells lexicographic; body(b, c1, c2).
  template <typename Body>
  void for_robin_face_cells(const oracle_problem *pb, const Mesh &m, Body body,
                            bool with_neumann = false)
  {
    const unsigned mask = robin_faces(pb) | (with_neumann ? pb->neumann_faces & 0x3Fu : 0u);
    for (int f = 0; f < 6; ++f)
      {
        if (!((mask >> f) & 1u))
          continue;
        const BoxFace b = box_face(f);
        for (int c2 = 0; c2 < m.n[b.e2]; ++c2)
          for (int c1 = 0; c1 < m.n[b.e1]; ++c1)
            body(b, c1, c2);
      }
  }

  // Gathers T (constraint-resolved: Dirichlet entries read 0, or plain) and u
  // (constraint-resolved) on one face cell.
  void face_gather(const oracle_problem *pb, const Mesh &m, const BoxFace &b, int c1, int c2,
                   const double *T, bool T_plain, const double *u, double *Tl, double *ul)
  {
    const int n = m.p + 1;
    for (int a2 = 0; a2 < n; ++a2)
      for (int a1 = 0; a1 < n; ++a1)
        {
          int64_t I[3];
          face_node(m, b, c1, c2, a1, a2, I);
          const bool con = is_constrained(pb, m, I[0], I[1], I[2]);
          const int64_t g = I[0] + m.N1[0] * (I[1] + m.N1[1] * I[2]);
          Tl[a2 * n + a1] = T ? ((con && !T_plain) ? 0.0 : T[g]) : 0.0;
          if (ul)
            ul[a2 * n + a1] = con ? 0.0 : u[g];
        }
  }

  // distribute_local_to_global of a face cell: constrained rows dropped.
  void face_scatter(const oracle_problem *pb, const Mesh &m, const BoxFace &b, int c1, int c2,
                    const double *y, double *dst)
  {
    const int n = m.p + 1;
    for (int a2 = 0; a2 < n; ++a2)
      for (int a1 = 0; a1 < n; ++a1)
        {
          int64_t I[3];
          face_node(m, b, c1, c2, a1, a2, I);
          if (!is_constrained(pb, m, I[0], I[1], I[2]))
            dst[I[0] + m.N1[0] * (I[1] + m.N1[1] * I[2])] += y[a2 * n + a1];
        }
  }

  // tangent_boundary_loop (heat_diffuse_operator.cpp:243-269): dst += A_face src.
  void robin_vmult(const oracle_problem *pb, const Tables &t, const Mesh &m, const double *src,
                   double *dst)
  {
    for_robin_face_cells(pb, m, [&](const BoxFace &b, int c1, int c2) {
      double Tl[MAXN * MAXN], ul[MAXN * MAXN], y[MAXN * MAXN];
      face_gather(pb, m, b, c1, c2, pb->temperature, false, src, Tl, ul);
      face_local(pb, t, m, b, kFaceTangent, Tl, ul, y);
      face_scatter(pb, m, b, c1, c2, y, dst);
    });
  }

  bool robin_ready(const oracle_problem *pb)
  {
    return !(pb->radiation_faces & 0x3Fu) || pb->temperature;
  }

  double dot(const std::vector<double> &a, const std::vector<double> &b)
  {
    double s = 0.0;
    const int64_t N = (int64_t)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < N; ++i)
      s += a[i] * b[i];
    return s;
  }

  inline uint64_t splitmix64(uint64_t z)
  {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
} // namespace

extern "C" {

void oracle_set_threads(int n)
{
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int oracle_tables(int p, double *gll, double *gauss_pts, double *gauss_w, double *S, double *D)
{
  if (p < 1 || p > ORACLE_MAX_DEGREE)
    return -1;
  long double nodes[MAXN], gx[MAXN], gw[MAXN];
  gll_nodes(p, nodes);
  gauss_rule(p + 1, gx, gw);
  Tables t = make_tables(p);
  for (int i = 0; i <= p; ++i)
    {
      if (gll)
        gll[i] = (double)nodes[i];
      if (gauss_pts)
        gauss_pts[i] = (double)gx[i];
      if (gauss_w)
        gauss_w[i] = (double)gw[i];
      for (int j = 0; j <= p; ++j)
        {
          if (S)
            S[i * (p + 1) + j] = t.S[i][j];
          if (D)
            D[i * (p + 1) + j] = t.D[i][j];
        }
    }
  return 0;
}

int64_t oracle_n_dofs(const oracle_problem *pb)
{
  if (!valid(pb))
    return -1;
  return make_mesh(pb).N;
}

int oracle_constrained_mask(const oracle_problem *pb, unsigned char *mask)
{
  if (!valid(pb))
    return -1;
  const Mesh m = make_mesh(pb);
  for (int64_t iz = 0; iz < m.N1[2]; ++iz)
    for (int64_t iy = 0; iy < m.N1[1]; ++iy)
      for (int64_t ix = 0; ix < m.N1[0]; ++ix)
        mask[ix + m.N1[0] * (iy + m.N1[1] * iz)] = is_constrained(pb, m, ix, iy, iz) ? 1 : 0;
  return 0;
}

// Seeded uniform[-1,1] on free DoFs, 0 on constrained ones. Keyed by
// (seed, lattice index) so it does not depend on thread order.
void oracle_random_vector(const oracle_problem *pb, uint64_t seed, double *v)
{
  const Mesh m = make_mesh(pb);
  const uint64_t key = splitmix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t iz = 0; iz < m.N1[2]; ++iz)
    for (int64_t iy = 0; iy < m.N1[1]; ++iy)
      for (int64_t ix = 0; ix < m.N1[0]; ++ix)
        {
          const int64_t I = ix + m.N1[0] * (iy + m.N1[1] * iz);
          const uint64_t r = splitmix64(key ^ (uint64_t)I);
          const double u = (double)(r >> 11) * 0x1.0p-53;
          v[I] = is_constrained(pb, m, ix, iy, iz) ? 0.0 : 2.0 * u - 1.0;
        }
}

int oracle_material(int material, int sharp, int volume_specific, int consistent_with_evaporation,
                    double gas_rho, double gas_cp, double gas_k, double liq_rho, double liq_cp,
                    double liq_k, double heaviside, double *rho_cp, double *k, double *rho,
                    double *cp)
{
  oracle_problem pb;
  std::memset(&pb, 0, sizeof(pb));
  pb.material = material;
  pb.sharp = sharp;
  pb.volume_specific = volume_specific;
  pb.consistent_with_evaporation = consistent_with_evaporation;
  pb.gas_rho = gas_rho;
  pb.gas_cp = gas_cp;
  pb.gas_k = gas_k;
  pb.liq_rho = liq_rho;
  pb.liq_cp = liq_cp;
  pb.liq_k = liq_k;
  material_at(&pb, heaviside, *rho_cp, *k);
  if (material == 0)
    {
      *rho = gas_rho;
      *cp = gas_cp;
    }
  else
    {
      const double w = sharp ? (heaviside > 0.5 ? 1.0 : 0.0) : heaviside;
      *cp = (1. - w) * gas_cp + w * liq_cp;
      *rho = consistent_with_evaporation ? 1. / ((1. - heaviside) / gas_rho + heaviside / liq_rho) :
                                           (1. - w) * gas_rho + w * liq_rho;
    }
  return 0;
}

int oracle_material_state(const oracle_problem *pb, double H, double T, double *out)
{
  if (!pb || !out || pb->material < 0 || pb->material > 3)
    return -1;
  const MatState s = material_state(pb, H, T);
  const double v[8] = {s.rho_cp, s.k, s.d_rho_cp, s.d_k, s.rho, s.cp, s.d_rho, s.d_cp};
  std::memcpy(out, v, sizeof(v));
  return 0;
}

int oracle_vmult(const oracle_problem *pb, const double *src, double *dst)
{
  if (!valid(pb) || !robin_ready(pb) || !tangent_ready(pb))
    return -1;
  const Tables t = make_tables(pb->degree);
  const Mesh m = make_mesh(pb);
  vmult_impl(pb, t, m, src, dst);
  return 0;
}

// create_rhs restated (heat_diffuse_operator.cpp:783-806): MatrixFree::loop over
// the cells with zero_dst = false, so dst accumulates; integrate_scatter through
// the Dirichlet-constrained DoF handler drops constrained rows; then the Robin
// faces. No Neumann faces and no cut-cell evaporation term. heat_source may be
// NULL (f = 0).
int oracle_create_rhs(const oracle_problem *pb, const double *temperature,
                      const double *temperature_old, const double *heat_source, double *dst)
{
  if (!valid(pb) || !temperature || !temperature_old || !dst)
    return -1;
  const Tables t = make_tables(pb->degree);
  const Mesh m = make_mesh(pb);
  const int n = t.n;
  for_cells_coloured(m, [&](const CellGeom &g) {
    double tn[MAXN * MAXN * MAXN], to[MAXN * MAXN * MAXN], fl[MAXN * MAXN * MAXN],
      Hl[MAXN * MAXN * MAXN], y[MAXN * MAXN * MAXN];
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const int l = (c * n + b) * n + a;
            const int64_t I = node_index(m, g, a, b, c);
            // read_dof_values_plain: constrained entries as stored
            tn[l] = temperature[I];
            to[l] = temperature_old[I];
            fl[l] = heat_source ? heat_source[I] : 0.0;
            Hl[l] = uses_heaviside(pb) ? pb->heaviside[I] : 0.0;
          }
    local_rhs(pb, t, m, g, tn, to, fl, Hl, y);
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const bool con = is_constrained(pb, m, (int64_t)g.c[0] * m.p + a,
                                            (int64_t)g.c[1] * m.p + b, (int64_t)g.c[2] * m.p + c);
            if (!con)
              dst[node_index(m, g, a, b, c)] += y[(c * n + b) * n + a];
          }
  });
  // rhs_boundary_loop (heat_diffuse_operator.cpp:719-781): T read plain; the
  // Robin and the Neumann faces
  for_robin_face_cells(
    pb, m,
    [&](const BoxFace &b, int c1, int c2) {
      double Tl[MAXN * MAXN], y[MAXN * MAXN];
      face_gather(pb, m, b, c1, c2, temperature, true, nullptr, Tl, nullptr);
      face_local(pb, t, m, b, kFaceRhs, Tl, nullptr, y);
      face_scatter(pb, m, b, c1, c2, y, dst);
    },
    true);
  return 0;
}

// VectorTools::compute_norm (L2) of the FE field restated
// (source/utilities/vector_tools.cpp:170-200 -> compute_global_error_norm against
// the zero function): sqrt(sum over cells and Gauss points of JxW u(x_q)^2),
// values read plain.
int oracle_l2_norm(const oracle_problem *pb, const double *u, double *norm)
{
  if (!valid(pb) || !u || !norm)
    return -1;
  const Tables t = make_tables(pb->degree);
  const Mesh m = make_mesh(pb);
  const int n = t.n;
  const double jac_det = m.h[0] * m.h[1] * m.h[2];
  const int64_t cells = (int64_t)m.n[0] * m.n[1] * m.n[2];
  std::vector<double> per_cell(cells, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < cells; ++i)
    {
      CellGeom g;
      g.c[0] = (int)(i % m.n[0]);
      g.c[1] = (int)((i / m.n[0]) % m.n[1]);
      g.c[2] = (int)(i / ((int64_t)m.n[0] * m.n[1]));
      double ul[MAXN * MAXN * MAXN], a1[MAXN * MAXN * MAXN], a2[MAXN * MAXN * MAXN],
        uq[MAXN * MAXN * MAXN];
      for (int c = 0; c < n; ++c)
        for (int b = 0; b < n; ++b)
          for (int a = 0; a < n; ++a)
            ul[(c * n + b) * n + a] = u[node_index(m, g, a, b, c)];
      apply_dir<false>(t.S, n, 0, ul, a1);
      apply_dir<false>(t.S, n, 1, a1, a2);
      apply_dir<false>(t.S, n, 2, a2, uq);
      double s = 0.0;
      for (int qz = 0; qz < n; ++qz)
        for (int qy = 0; qy < n; ++qy)
          for (int qx = 0; qx < n; ++qx)
            {
              const double v = uq[(qz * n + qy) * n + qx];
              s += t.W[qx] * t.W[qy] * t.W[qz] * jac_det * v * v;
            }
      per_cell[i] = s;
    }
  double total = 0.0;
  for (double v : per_cell)
    total += v;
  *norm = std::sqrt(total);
  return 0;
}

// MatrixFreeTools::compute_diagonal restated: apply the cell operator (and the
// Robin face operator) to each local unit vector and keep the matching entry
// (heat_diffuse_operator.cpp:332-358).
int oracle_diagonal(const oracle_problem *pb, double *diag)
{
  if (!valid(pb) || !robin_ready(pb) || !tangent_ready(pb))
    return -1;
  const Tables t = make_tables(pb->degree);
  const Mesh m = make_mesh(pb);
  const int n = t.n;
  const int n3 = n * n * n;
  std::fill(diag, diag + m.N, 0.0);
  for_cells_coloured(m, [&](const CellGeom &g) {
    double u[MAXN * MAXN * MAXN], y[MAXN * MAXN * MAXN];
    double d[MAXN * MAXN * MAXN];
    CellFields fl;
    gather_fields(pb, m, g, pb->temperature, pb->temperature_old, fl);
    for (int j = 0; j < n3; ++j)
      {
        std::fill(u, u + n3, 0.0);
        u[j] = 1.0;
        local_apply(pb, t, m, g, u, fl, y);
        d[j] = y[j];
      }
    for (int c = 0; c < n; ++c)
      for (int b = 0; b < n; ++b)
        for (int a = 0; a < n; ++a)
          {
            const bool con = is_constrained(pb, m, (int64_t)g.c[0] * m.p + a,
                                            (int64_t)g.c[1] * m.p + b, (int64_t)g.c[2] * m.p + c);
            if (!con)
              diag[node_index(m, g, a, b, c)] += d[(c * n + b) * n + a];
          }
  });
  // boundary operation of compute_diagonal (heat_diffuse_operator.cpp:352-356):
  // each face-local unit vector through tangent_local_boundary_operation
  const int n2 = n * n;
  for_robin_face_cells(pb, m, [&](const BoxFace &b, int c1, int c2) {
    double Tl[MAXN * MAXN], e[MAXN * MAXN], y[MAXN * MAXN], dl[MAXN * MAXN];
    face_gather(pb, m, b, c1, c2, pb->temperature, false, nullptr, Tl, nullptr);
    for (int j = 0; j < n2; ++j)
      {
        std::fill(e, e + n2, 0.0);
        e[j] = 1.0;
        face_local(pb, t, m, b, kFaceTangent, Tl, e, y);
        dl[j] = y[j];
      }
    face_scatter(pb, m, b, c1, c2, dl, diag);
  });
  return 0;
}

int oracle_inverse_diagonal(const oracle_problem *pb, double *inv_diag)
{
  const int rc = oracle_diagonal(pb, inv_diag);
  if (rc)
    return rc;
  const int64_t N = make_mesh(pb).N;
  // heat_diffuse_operator.cpp:282-283
  for (int64_t i = 0; i < N; ++i)
    inv_diag[i] = std::abs(inv_diag[i]) > 1.0e-10 ? 1.0 / inv_diag[i] : 1.0;
  return 0;
}

// LinearSolver::solve restated (linear_solver.hpp:26-134): deal.II SolverCG /
// right-preconditioned SolverGMRES under ReductionControl(max_it, abs, rel),
// Jacobi preconditioner dst = inv_diag .* src (preconditioner_jacobi.hpp:51-56).
// Return 0 on convergence, 1 on NoConvergence, -1 on bad input.
int oracle_solve(const oracle_problem *pb, int solver, int restart, const double *b, double *x,
                 const double *inv_diag, double rel_tol, double abs_tol, int max_it,
                 int *iterations, double *initial_res, double *final_res, double *history,
                 int history_len)
{
  if (!valid(pb) || !robin_ready(pb) || !tangent_ready(pb) || (solver == 1 && restart < 1))
    return -1;
  const Tables t = make_tables(pb->degree);
  const Mesh m = make_mesh(pb);
  const int64_t N = m.N;
  std::vector<double> X(x, x + N), B(b, b + N), D(N, 1.0);
  if (inv_diag)
    D.assign(inv_diag, inv_diag + N);
  std::vector<double> r(N), tmp(N);
  auto record = [&](int it, double res) {
    if (history && it < history_len)
      history[it] = res;
  };

  vmult_impl(pb, t, m, X.data(), tmp.data());
  for (int64_t i = 0; i < N; ++i)
    r[i] = B[i] - tmp[i];
  double res = std::sqrt(dot(r, r));
  *initial_res = res;
  const double tol = std::max(abs_tol, rel_tol * res);
  record(0, res);
  int it = 0;
  int status = 1;
  if (res <= tol)
    status = 0;

  if (status != 0 && solver == 0)
    {
      std::vector<double> z(N), p(N), q(N);
      double rz = 0.0;
      while (true)
        {
          // SolverCG IterationWorker::do_iteration (deal.II, restated)
          ++it;
          for (int64_t i = 0; i < N; ++i)
            z[i] = D[i] * r[i];
          const double rz_old = rz;
          rz = dot(r, z);
          if (it > 1)
            {
              const double beta = rz / rz_old;
              for (int64_t i = 0; i < N; ++i)
                p[i] = z[i] + beta * p[i];
            }
          else
            p = z;
          vmult_impl(pb, t, m, p.data(), q.data());
          const double alpha = rz / dot(p, q);
          for (int64_t i = 0; i < N; ++i)
            {
              X[i] += alpha * p[i];
              r[i] -= alpha * q[i];
            }
          res = std::sqrt(std::abs(dot(r, r)));
          record(it, res);
          if (res <= tol)
            {
              status = 0;
              break;
            }
          if (it >= max_it || std::isnan(res))
            break;
        }
    }
  else if (status != 0 && solver == 1)
    {
      // GMRES(restart), right preconditioning, classical Gram-Schmidt with one
      // re-orthogonalisation pass; the inner check uses the Givens residual estimate.
      const int mr = restart;
      std::vector<std::vector<double>> V(mr + 1, std::vector<double>(N));
      std::vector<double> H((size_t)(mr + 1) * mr), cs(mr), sn(mr), gv(mr + 1), w(N), zz(N),
        hcol(mr + 1);
      bool done = false;
      while (!done)
        {
          const double beta = std::sqrt(dot(r, r));
          for (int64_t i = 0; i < N; ++i)
            V[0][i] = r[i] / beta;
          std::fill(gv.begin(), gv.end(), 0.0);
          gv[0] = beta;
          int j = 0;
          for (; j < mr; ++j)
            {
              ++it;
              for (int64_t i = 0; i < N; ++i)
                zz[i] = D[i] * V[j][i];
              vmult_impl(pb, t, m, zz.data(), w.data());
              std::fill(hcol.begin(), hcol.end(), 0.0);
              for (int pass = 0; pass < 2; ++pass)
                {
                  std::vector<double> hh(j + 1);
                  for (int l = 0; l <= j; ++l)
                    hh[l] = dot(V[l], w);
                  for (int l = 0; l <= j; ++l)
                    {
                      for (int64_t i = 0; i < N; ++i)
                        w[i] -= hh[l] * V[l][i];
                      hcol[l] += hh[l];
                    }
                }
              const double hn = std::sqrt(dot(w, w));
              hcol[j + 1] = hn;
              if (hn > 0)
                for (int64_t i = 0; i < N; ++i)
                  V[j + 1][i] = w[i] / hn;
              for (int l = 0; l < j; ++l)
                {
                  const double a = cs[l] * hcol[l] + sn[l] * hcol[l + 1];
                  const double bb = -sn[l] * hcol[l] + cs[l] * hcol[l + 1];
                  hcol[l] = a;
                  hcol[l + 1] = bb;
                }
              const double den = std::hypot(hcol[j], hcol[j + 1]);
              cs[j] = hcol[j] / den;
              sn[j] = hcol[j + 1] / den;
              hcol[j] = den;
              hcol[j + 1] = 0.0;
              gv[j + 1] = -sn[j] * gv[j];
              gv[j] = cs[j] * gv[j];
              for (int l = 0; l <= j; ++l)
                H[(size_t)l * mr + j] = hcol[l];
              res = std::abs(gv[j + 1]);
              record(it, res);
              if (res <= tol || it >= max_it || std::isnan(res))
                {
                  ++j;
                  break;
                }
            }
          // solve the upper triangular system and update x += P V y
          std::vector<double> yv(j);
          for (int l = j - 1; l >= 0; --l)
            {
              double s = gv[l];
              for (int c = l + 1; c < j; ++c)
                s -= H[(size_t)l * mr + c] * yv[c];
              yv[l] = s / H[(size_t)l * mr + l];
            }
          std::fill(zz.begin(), zz.end(), 0.0);
          for (int l = 0; l < j; ++l)
            for (int64_t i = 0; i < N; ++i)
              zz[i] += yv[l] * V[l][i];
          for (int64_t i = 0; i < N; ++i)
            X[i] += D[i] * zz[i];
          if (res <= tol)
            {
              status = 0;
              done = true;
            }
          else if (it >= max_it || std::isnan(res))
            done = true;
          else
            {
              vmult_impl(pb, t, m, X.data(), tmp.data());
              for (int64_t i = 0; i < N; ++i)
                r[i] = B[i] - tmp[i];
              res = std::sqrt(dot(r, r));
              if (res <= tol)
                {
                  status = 0;
                  done = true;
                }
            }
        }
    }
  std::copy(X.begin(), X.end(), x);
  *iterations = it;
  *final_res = res;
  return status;
}

} // extern "C"
