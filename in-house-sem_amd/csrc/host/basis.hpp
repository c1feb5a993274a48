// One-dimensional quadrature and interpolation for the host setup.
//
// Lobatto rules for the weights (1 - x)^alpha (1 + x)^beta: GLL (0, 0) for
// every element direction, GLJ (0, 1) along r on elements that touch the
// axis (the rules of cpp/src/basis/quadrature.cpp:46-132).  The nodes come
// from the symmetric Jacobi matrix of the interior weight (Golub & Welsch
// 1969) and the weights from its Christoffel numbers; Lagrange cardinal
// functions are evaluated in barycentric form (Berrut & Trefethen 2004).
#pragma once

#include <vector>

namespace insem_amd {

struct Rule {
  std::vector<double> points;  // ascending on [-1, 1], both end points included
  std::vector<double> weights;
  int size() const { return static_cast<int>(points.size()); }
};

// n-point Lobatto rule for the weight (1 - x)^alpha (1 + x)^beta
// (integer exponents >= 0, n >= 2): exact for polynomials of degree 2n - 3.
Rule lobatto_rule(int n, int alpha, int beta);
inline Rule gll_rule(int n) { return lobatto_rule(n, 0, 0); }
inline Rule glj01_rule(int n) { return lobatto_rule(n, 0, 1); }

// The Lagrange cardinal functions l_j of a node set (l_j(x_k) = delta_jk).
class Lagrange1D {
 public:
  explicit Lagrange1D(std::vector<double> nodes);
  int size() const { return static_cast<int>(nodes_.size()); }
  const std::vector<double>& nodes() const { return nodes_; }
  // l_j(t) and l_j'(t) for every j, into out[0 .. size)
  void values(double t, double* out) const;
  void derivatives(double t, double* out) const;
  std::vector<double> values(double t) const;
  std::vector<double> derivatives(double t) const;
  // D[i * n + j] = l_j'(x_i)
  std::vector<double> derivative_matrix() const;

 private:
  int node_at(double t) const;  // the node t coincides with, or -1
  void derivatives_at_node(int i, double* out) const;
  std::vector<double> nodes_;
  std::vector<double> bary_;  // 1 / prod_{k != j} (x_j - x_k)
};

}  // namespace insem_amd
