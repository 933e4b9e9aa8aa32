// lsmf_fe.hpp -- 1D finite element tables for the tensor-product Q_p element.
//
// FE_Q(p) in deal.II places its support points at the Gauss-Lobatto points
// (equispaced for p <= 2); the operator integrates with QGauss(n_q_1d)
// (reference: include/mf_elasticity.h:1752, src/main.cc:34-36).  All tables
// live on the unit interval [0,1], ascending.
#pragma once

#include <cmath>
#include <vector>

namespace lsmf {

inline void legendre_eval(int n, double x, double &P, double &dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    P = 1.0;
    dP = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  P = p1;
  dP = n * (x * p1 - p0) / (x * x - 1.0);
}

// Gauss-Legendre rule with n points on [0,1].
inline void gauss_rule(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre_eval(n, z, P, dP);
      const double dz = P / dP;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double P, dP;
    legendre_eval(n, z, P, dP);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dP * dP);
  }
}

// Gauss-Lobatto support points of degree p on [0,1].
inline std::vector<double> lobatto_points(int p) {
  std::vector<double> x(p + 1, 0.0);
  x[p] = 1.0;
  for (int i = 1; i < p; ++i) {
    double z = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre_eval(p, z, P, dP);
      const double d2P = (2.0 * z * dP - p * (p + 1.0) * P) / (1.0 - z * z);
      const double dz = dP / d2P;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
  return x;
}

inline double lagrange_value(const std::vector<double> &nodes, int i, double x) {
  double v = 1.0;
  for (size_t j = 0; j < nodes.size(); ++j)
    if ((int)j != i) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
  return v;
}

inline double lagrange_derivative(const std::vector<double> &nodes, int i, double x) {
  double s = 0.0;
  for (size_t k = 0; k < nodes.size(); ++k) {
    if ((int)k == i) continue;
    double v = 1.0 / (nodes[i] - nodes[k]);
    for (size_t j = 0; j < nodes.size(); ++j)
      if ((int)j != i && j != k) v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    s += v;
  }
  return s;
}

struct Tables1D {
  int p = 0, nq = 0;
  std::vector<double> nodes, qpts, qw;
  std::vector<double> S, D;  // [nq][p+1]: shape value / derivative at Gauss points
  void init(int p_, int nq_) {
    p = p_;
    nq = nq_;
    nodes = lobatto_points(p);
    gauss_rule(nq, qpts, qw);
    S.assign(nq * (p + 1), 0.0);
    D.assign(nq * (p + 1), 0.0);
    for (int q = 0; q < nq; ++q)
      for (int a = 0; a <= p; ++a) {
        S[q * (p + 1) + a] = lagrange_value(nodes, a, qpts[q]);
        D[q * (p + 1) + a] = lagrange_derivative(nodes, a, qpts[q]);
      }
  }
};

}  // namespace lsmf
