// Compile-time 1D tables of FE_Q(p) on Gauss-Lobatto points with QGauss(p+1)
// (the same quantities as tables.cpp, evaluated by constexpr code in double).
//
// The cell kernels take their 1D matrices from here, so every table entry is a
// constant the compiler can place directly in the instruction stream instead
// of a kernel argument held in (and spilled from) scalar registers.
// heat_api.cpp checks these against the long-double host tables at set-up.
#pragma once

namespace mpdg
{
  namespace ct
  {
    constexpr double kPi = 3.14159265358979323846264338327950288;

    constexpr double cabs(double x) { return x < 0 ? -x : x; }

    // cos by range reduction to [-pi, pi] and its Taylor series
    constexpr double ccos(double x)
    {
      while (x > kPi)
        x -= 2 * kPi;
      while (x < -kPi)
        x += 2 * kPi;
      double term = 1.0, sum = 1.0;
      for (int k = 1; k < 40; ++k)
        {
          term *= -x * x / ((2 * k - 1) * (2 * k));
          sum += term;
        }
      return sum;
    }

    struct PdP
    {
      double p, dp;
    };

    // P_n(x), P_n'(x) by the Bonnet recurrence
    constexpr PdP legendre(int n, double x)
    {
      if (n == 0)
        return {1.0, 0.0};
      double pm1 = 1.0, p = x;
      for (int k = 1; k < n; ++k)
        {
          const double pk1 = ((2 * k + 1) * x * p - k * pm1) / (k + 1);
          pm1 = p;
          p = pk1;
        }
      return {p, n * (pm1 - x * p) / (1.0 - x * x)};
    }

    template <int N>
    struct Tab
    {
      double gll[N]{};
      double x[N]{};
      double w[N]{};
      double S[N][N]{}; // S[q][i] = phi_i(x_q)
      double C[N][N]{}; // C[q][r] = l_r'(x_q), Lagrange on the Gauss points
      double D[N][N]{}; // D[q][i] = phi_i'(x_q)
    };

    template <int N>
    constexpr double lagrange(const double (&nodes)[N], int i, double x)
    {
      double v = 1.0;
      for (int j = 0; j < N; ++j)
        if (j != i)
          v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
      return v;
    }

    template <int N>
    constexpr double lagrange_derivative(const double (&nodes)[N], int i, double x)
    {
      double sum = 0.0;
      for (int k = 0; k < N; ++k)
        {
          if (k == i)
            continue;
          double prod = 1.0 / (nodes[i] - nodes[k]);
          for (int j = 0; j < N; ++j)
            if (j != i && j != k)
              prod *= (x - nodes[j]) / (nodes[i] - nodes[j]);
          sum += prod;
        }
      return sum;
    }

    template <int P>
    constexpr Tab<P + 1> make_tab()
    {
      constexpr int N = P + 1;
      Tab<N> t{};
      // Gauss-Legendre: roots of P_N, ascending, Newton from -cos(pi (i+3/4)/(N+1/2))
      for (int i = 0; i < N; ++i)
        {
          double z = -ccos(kPi * (i + 0.75) / (N + 0.5));
          for (int it = 0; it < 100; ++it)
            {
              const PdP v = legendre(N, z);
              const double dz = v.p / v.dp;
              z -= dz;
              if (cabs(dz) < 1e-17)
                break;
            }
          const PdP v = legendre(N, z);
          t.x[i] = 0.5 * (z + 1.0);
          t.w[i] = 1.0 / ((1.0 - z * z) * v.dp * v.dp);
        }
      // Gauss-Lobatto: 0, roots of P_P', 1
      t.gll[0] = 0.0;
      t.gll[P] = 1.0;
      for (int i = 1; i < P; ++i)
        {
          double z = -ccos(kPi * i / P);
          for (int it = 0; it < 100; ++it)
            {
              const PdP v = legendre(P, z);
              const double d2 = (2.0 * z * v.dp - P * (P + 1.0) * v.p) / (1.0 - z * z);
              const double dz = v.dp / d2;
              z -= dz;
              if (cabs(dz) < 1e-17)
                break;
            }
          t.gll[i] = 0.5 * (z + 1.0);
        }
      // symmetrise the point sets (exact mirror images about 1/2)
      for (int i = 0; i < N / 2; ++i)
        {
          const double a = 0.5 * (t.x[i] + (1.0 - t.x[N - 1 - i]));
          t.x[i] = a;
          t.x[N - 1 - i] = 1.0 - a;
          const double ww = 0.5 * (t.w[i] + t.w[N - 1 - i]);
          t.w[i] = ww;
          t.w[N - 1 - i] = ww;
          const double g = 0.5 * (t.gll[i] + (1.0 - t.gll[P - i]));
          t.gll[i] = g;
          t.gll[P - i] = 1.0 - g;
        }
      if (N % 2)
        {
          t.x[N / 2] = 0.5;
          t.gll[P / 2] = P % 2 ? t.gll[P / 2] : 0.5;
        }
      for (int q = 0; q < N; ++q)
        for (int i = 0; i < N; ++i)
          {
            t.S[q][i] = lagrange<N>(t.gll, i, t.x[q]);
            t.D[q][i] = lagrange_derivative<N>(t.gll, i, t.x[q]);
            t.C[q][i] = lagrange_derivative<N>(t.x, i, t.x[q]);
          }
      return t;
    }

    // Even-odd split of an N x N matrix (same layout as mpdg::EvenOdd).
    template <int N>
    struct EO
    {
      double e[N / 2 > 0 ? N / 2 : 1][N / 2 > 0 ? N / 2 : 1]{};
      double o[N / 2 > 0 ? N / 2 : 1][N / 2 > 0 ? N / 2 : 1]{};
      double mc[N / 2 > 0 ? N / 2 : 1]{};
      double mr[N / 2 > 0 ? N / 2 : 1]{};
      double mm = 0.0;
    };

    template <int N>
    constexpr EO<N> split(const double (&M)[N][N], bool transpose)
    {
      EO<N> r{};
      const int h = N / 2;
      for (int i = 0; i < h; ++i)
        for (int j = 0; j < h; ++j)
          {
            const double mij = transpose ? M[j][i] : M[i][j];
            const double mik = transpose ? M[N - 1 - j][i] : M[i][N - 1 - j];
            r.e[i][j] = 0.5 * (mij + mik);
            r.o[i][j] = 0.5 * (mij - mik);
          }
      if (N % 2)
        {
          for (int i = 0; i < h; ++i)
            {
              r.mc[i] = transpose ? M[h][i] : M[i][h];
              r.mr[i] = transpose ? M[i][h] : M[h][i];
            }
          r.mm = M[h][h];
        }
      return r;
    }

    // Dense N x N matrix applied as y_i = sum_j a[i][j] x_j.
    template <int N>
    struct Mat
    {
      double a[N][N]{};
    };

    // Dense 1D matrices of the cell-form vmult, with the quadrature weights
    // folded in so that the q-point work needs no weights:
    //   kind 0: S[q][i]                (interpolate nodal -> Gauss)
    //   kind 1: C[q][r]                (collocation derivative on Gauss points)
    //   kind 2: w_r C[r][q] / w_q      (W^-1 C^T W: the flux back to values)
    //   kind 3: S[q][i] w_q, as [i][q] (S^T W: integrate)
    //   kind 4: S^T, kind 5: C^T       (unweighted transposes)
    template <int P>
    constexpr Mat<P + 1> cell_matrix(const Tab<P + 1> &t, int kind)
    {
      constexpr int N = P + 1;
      Mat<N> m{};
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          m.a[i][j] = kind == 0 ? t.S[i][j]
                    : kind == 1 ? t.C[i][j]
                    : kind == 2 ? t.w[j] * t.C[j][i] / t.w[i]
                    : kind == 3 ? t.S[j][i] * t.w[j]
                    : kind == 4 ? t.S[j][i]
                                : t.C[j][i];
      return m;
    }

    template <int P>
    struct CTab
    {
      static constexpr Tab<P + 1> t = make_tab<P>();
      static constexpr Mat<P + 1> S = cell_matrix<P>(t, 0);
      static constexpr Mat<P + 1> C = cell_matrix<P>(t, 1);
      static constexpr Mat<P + 1> CW = cell_matrix<P>(t, 2);
      static constexpr Mat<P + 1> SW = cell_matrix<P>(t, 3);
      static constexpr Mat<P + 1> St = cell_matrix<P>(t, 4);
      static constexpr Mat<P + 1> Ct = cell_matrix<P>(t, 5);
    };

    // All even-odd tables of the vmult kernel for degree P.
    // M with its point symmetry made exact: M[N-1-i][N-1-j] = sign M[i][j]
    // (sign +1 for S on symmetric points, -1 for the derivative C, whose
    // centre entry is then exactly 0).
    template <int N>
    struct Sq
    {
      double a[N][N]{};
    };

    template <int N>
    constexpr Sq<N> symmetrize(const double (&M)[N][N], int sign)
    {
      Sq<N> r{};
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          r.a[i][j] = 0.5 * (M[i][j] + sign * M[N - 1 - i][N - 1 - j]);
      return r;
    }

    // Even-odd form of M^T from that of M (M point-symmetric with `sign`):
    // the blocks are transposes of M's (swapped for sign -1) and the centre
    // column and row trade places, so M and M^T share their constants
    // bitwise (a kernel then holds one set of scalar registers for both).
    template <int N>
    constexpr EO<N> transpose_eo(const EO<N> &m, int sign)
    {
      EO<N> r{};
      const int h = N / 2;
      for (int i = 0; i < h; ++i)
        for (int j = 0; j < h; ++j)
          {
            r.e[i][j] = sign > 0 ? m.e[j][i] : m.o[j][i];
            r.o[i][j] = sign > 0 ? m.o[j][i] : m.e[j][i];
          }
      for (int i = 0; i < h; ++i)
        {
          r.mc[i] = m.mr[i];
          r.mr[i] = m.mc[i];
        }
      r.mm = m.mm;
      return r;
    }

    // Reference-cell 1D mass and stiffness matrices on the Gauss rule of the
    // operator: Mhat[i][j] = sum_q w_q phi_i(x_q) phi_j(x_q) (kind 0) and
    // Khat[i][j] = sum_q w_q phi_i'(x_q) phi_j'(x_q) (kind 1). The rule is exact
    // for these products, so on an affine box cell with constant coefficients
    // the cell matrix is the Kronecker sum of these factors.
    template <int N>
    constexpr Sq<N> gauss_product(const Tab<N> &t, int kind)
    {
      Sq<N> r{};
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
          {
            double s = 0.0;
            for (int q = 0; q < N; ++q)
              s += t.w[q] * (kind == 0 ? t.S[q][i] * t.S[q][j] : t.D[q][i] * t.D[q][j]);
            r.a[i][j] = s;
          }
      return r;
    }

    template <int P>
    struct VTab
    {
      static constexpr Tab<P + 1> t = make_tab<P>();
      static constexpr Sq<P + 1> Mhat = gauss_product<P + 1>(t, 0);
      static constexpr Sq<P + 1> Khat = gauss_product<P + 1>(t, 1);
      static constexpr EO<P + 1> Mk = split<P + 1>(symmetrize<P + 1>(Mhat.a, 1).a, false);
      static constexpr EO<P + 1> Kk = split<P + 1>(symmetrize<P + 1>(Khat.a, 1).a, false);
      static constexpr Sq<P + 1> Ssym = symmetrize<P + 1>(t.S, 1);
      static constexpr Sq<P + 1> Csym = symmetrize<P + 1>(t.C, -1);
      static constexpr EO<P + 1> S = split<P + 1>(Ssym.a, false);
      static constexpr EO<P + 1> St = transpose_eo<P + 1>(S, 1);
      static constexpr EO<P + 1> C = split<P + 1>(Csym.a, false);
      static constexpr EO<P + 1> Ct = transpose_eo<P + 1>(C, -1);
    };
  } // namespace ct
} // namespace mpdg
