// 1D tables: Gauss-Lobatto support points, Gauss-Legendre quadrature, nodal
// basis values/derivatives at the quadrature points and the collocation
// derivative on the Gauss points. Computed in long double by Newton iteration
// on Legendre polynomials, then rounded once to double.
#include "tables.hpp"

#include <cmath>
#include <stdexcept>

namespace mpdg
{
  namespace
  {
    constexpr long double kPi = 3.141592653589793238462643383279502884L;

    // (P_n(x), P_n'(x)) by the Bonnet recurrence.
    void legendre_pair(int n, long double x, long double &pn, long double &dpn)
    {
      long double pm1 = 1.0L, p = x;
      if (n == 0)
        {
          pn = 1.0L;
          dpn = 0.0L;
          return;
        }
      for (int k = 1; k < n; ++k)
        {
          const long double pk1 = ((2 * k + 1) * x * p - k * pm1) / (k + 1);
          pm1 = p;
          p = pk1;
        }
      pn = p;
      dpn = n * (pm1 - x * p) / (1.0L - x * x);
    }

    long double basis(const long double *nodes, int n, int i, long double x)
    {
      long double v = 1.0L;
      for (int j = 0; j < n; ++j)
        if (j != i)
          v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
      return v;
    }

    long double basis_derivative(const long double *nodes, int n, int i, long double x)
    {
      // d/dx prod_{j!=i} (x - x_j)/(x_i - x_j) = sum_k 1/(x_i-x_k) prod_{j!=i,k} ...
      long double sum = 0.0L;
      for (int k = 0; k < n; ++k)
        {
          if (k == i)
            continue;
          long double prod = 1.0L / (nodes[i] - nodes[k]);
          for (int j = 0; j < n; ++j)
            if (j != i && j != k)
              prod *= (x - nodes[j]) / (nodes[i] - nodes[j]);
          sum += prod;
        }
      return sum;
    }
  } // namespace

  Tables1D make_tables(int p)
  {
    if (p < 1 || p > kMaxDegree)
      throw std::invalid_argument("degree out of range");
    Tables1D t;
    t.p = p;
    t.n = p + 1;
    const int n = p + 1;
    long double gll[kMaxN], xg[kMaxN], wg[kMaxN];

    // Gauss-Legendre: roots of P_n, ascending, mapped to [0,1].
    for (int i = 0; i < n; ++i)
      {
        long double x = -std::cos(kPi * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 200; ++it)
          {
            long double pn, dpn;
            legendre_pair(n, x, pn, dpn);
            const long double dx = pn / dpn;
            x -= dx;
            if (std::fabs(dx) <= 1e-32L)
              break;
          }
        long double pn, dpn;
        legendre_pair(n, x, pn, dpn);
        xg[i] = 0.5L * (x + 1.0L);
        wg[i] = 1.0L / ((1.0L - x * x) * dpn * dpn);
      }

    // Gauss-Lobatto: endpoints and the roots of P_p'.
    gll[0] = 0.0L;
    gll[p] = 1.0L;
    for (int i = 1; i < p; ++i)
      {
        long double x = -std::cos(kPi * i / p);
        for (int it = 0; it < 200; ++it)
          {
            long double pn, dpn;
            legendre_pair(p, x, pn, dpn);
            const long double d2 = (2.0L * x * dpn - p * (p + 1.0L) * pn) / (1.0L - x * x);
            const long double dx = dpn / d2;
            x -= dx;
            if (std::fabs(dx) <= 1e-32L)
              break;
          }
        gll[i] = 0.5L * (x + 1.0L);
      }

    for (int q = 0; q < n; ++q)
      {
        t.gll[q] = (double)gll[q];
        t.xq[q] = (double)xg[q];
        t.wq[q] = (double)wg[q];
        for (int i = 0; i < n; ++i)
          {
            t.S[q][i] = (double)basis(gll, n, i, xg[q]);
            t.D[q][i] = (double)basis_derivative(gll, n, i, xg[q]);
            t.C[q][i] = (double)basis_derivative(xg, n, i, xg[q]);
          }
      }
    return t;
  }

  template <typename Number>
  EvenOdd<Number> even_odd(const double (&M)[kMaxN][kMaxN], int n, bool transpose)
  {
    auto at = [&](int i, int j) { return transpose ? M[j][i] : M[i][j]; };
    EvenOdd<Number> r{};
    const int h = n / 2;
    for (int i = 0; i < h; ++i)
      for (int j = 0; j < h; ++j)
        {
          r.e[i][j] = (Number)(0.5 * (at(i, j) + at(i, n - 1 - j)));
          r.o[i][j] = (Number)(0.5 * (at(i, j) - at(i, n - 1 - j)));
        }
    if (n % 2)
      {
        for (int i = 0; i < h; ++i)
          {
            r.mc[i] = (Number)at(i, h);
            r.mr[i] = (Number)at(h, i);
          }
        r.mm = (Number)at(h, h);
      }
    return r;
  }

  double symmetry_defect(const double (&M)[kMaxN][kMaxN], int n, int s)
  {
    double d = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        d = std::fmax(d, std::fabs(M[n - 1 - i][n - 1 - j] - s * M[i][j]));
    return d;
  }

  template EvenOdd<double> even_odd<double>(const double (&)[kMaxN][kMaxN], int, bool);
  template EvenOdd<float> even_odd<float>(const double (&)[kMaxN][kMaxN], int, bool);
} // namespace mpdg
