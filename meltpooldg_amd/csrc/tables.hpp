// 1D finite-element tables for FE_Q(p) on Gauss-Lobatto support points with
// QGauss(p+1) (source/utilities/fe_util.cpp:16-58, source/core/finite_element_data.cpp:61-65),
// and their even-odd (symmetric/antisymmetric) splits used by the cell kernels.
#pragma once

#include <array>

namespace mpdg
{
  constexpr int kMaxDegree = 6;
  constexpr int kMaxN = kMaxDegree + 1;

  struct Tables1D
  {
    int p = 0;
    int n = 0;                                   // points per direction = p + 1
    std::array<double, kMaxN> gll{};             // support points on [0,1]
    std::array<double, kMaxN> xq{};              // Gauss points on [0,1]
    std::array<double, kMaxN> wq{};              // Gauss weights on [0,1]
    double S[kMaxN][kMaxN]{};                    // S[q][i]  = phi_i(x_q)
    double D[kMaxN][kMaxN]{};                    // D[q][i]  = phi_i'(x_q)
    double C[kMaxN][kMaxN]{};                    // C[q][r]  = l_r'(x_q), l = Lagrange on Gauss points
  };

  Tables1D make_tables(int p);

  // Even-odd form of an n x n matrix M with M[n-1-i][n-1-j] = s M[i][j]:
  //   e[i][j] = (M[i][j] + M[i][n-1-j]) / 2, o[i][j] = (M[i][j] - M[i][n-1-j]) / 2
  //   mc[i] = M[i][m], mr[j] = M[m][j], mm = M[m][m]   (m = n/2 for odd n)
  // so that, with E_j = x_j + x_{n-1-j}, O_j = x_j - x_{n-1-j}:
  //   y_i = A_i + B_i, y_{n-1-i} = s (A_i - B_i), A_i = e E + mc x_m, B_i = o O,
  //   y_m = mr (x_j + s x_{n-1-j}) + mm x_m.
  template <typename Number>
  struct EvenOdd
  {
    Number e[kMaxN / 2][kMaxN / 2];
    Number o[kMaxN / 2][kMaxN / 2];
    Number mc[kMaxN / 2];
    Number mr[kMaxN / 2];
    Number mm;
  };

  template <typename Number>
  EvenOdd<Number> even_odd(const double (&M)[kMaxN][kMaxN], int n, bool transpose);

  // Largest violation of M[n-1-i][n-1-j] = s M[i][j]; used to assert the split is exact.
  double symmetry_defect(const double (&M)[kMaxN][kMaxN], int n, int s);
} // namespace mpdg
