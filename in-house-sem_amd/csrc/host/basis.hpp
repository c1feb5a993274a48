// 1D quadrature rules and Lagrange bases used by the host setup.
//
// Follows the reference's basis layer (cpp/src/basis/quadrature.cpp:46-132,
// cpp/src/basis/lagrange.cpp:10-86): Jacobi polynomials by their three-term
// recurrence, roots by Newton iteration with deflation, Gauss-Lobatto-Jacobi
// weights from a moment solve, Lagrange values by the barycentric formula.
#pragma once

#include <vector>

namespace insem_amd {

struct Rule {
  std::vector<double> points;  // ascending on [-1, 1]
  std::vector<double> weights;
  int size() const { return static_cast<int>(points.size()); }
};

double jacobi_p(int n, double a, double b, double x);
double jacobi_dp(int n, double a, double b, double x);
std::vector<double> jacobi_zeros(int n, double a, double b);
Rule gauss_legendre_rule(int n);
// Lobatto rule for the weight (1 - x)^alpha (1 + x)^beta with both end points.
Rule lobatto_jacobi_rule(int n, int alpha, int beta);
inline Rule gll_rule(int n) { return lobatto_jacobi_rule(n, 0, 0); }
inline Rule glj01_rule(int n) { return lobatto_jacobi_rule(n, 0, 1); }

class Lagrange1D {
 public:
  explicit Lagrange1D(std::vector<double> nodes);
  int size() const { return static_cast<int>(x_.size()); }
  const std::vector<double>& nodes() const { return x_; }
  void values(double t, double* out) const;
  void derivatives(double t, double* out) const;
  std::vector<double> values(double t) const;
  std::vector<double> derivatives(double t) const;
  // D[i * n + j] = l_j'(x_i)
  std::vector<double> derivative_matrix() const;

 private:
  int coincident(double t) const;
  std::vector<double> x_;
  std::vector<double> lambda_;  // barycentric weights
};

}  // namespace insem_amd
