#include "basis.hpp"

#include <cmath>
#include <utility>

#include "error.hpp"

namespace insem_amd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Dense solve with partial pivoting for the small moment systems.
std::vector<double> dense_solve(std::vector<double> m, std::vector<double> rhs, int n) {
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(m[r * n + c]) > std::fabs(m[p * n + c])) p = r;
    INSEM_REQUIRE(m[p * n + c] != 0.0, "singular quadrature moment system");
    if (p != c) {
      for (int k = 0; k < n; ++k) std::swap(m[c * n + k], m[p * n + k]);
      std::swap(rhs[c], rhs[p]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double l = m[r * n + c] / m[c * n + c];
      for (int k = c; k < n; ++k) m[r * n + k] -= l * m[c * n + k];
      rhs[r] -= l * rhs[c];
    }
  }
  std::vector<double> x(n);
  for (int r = n - 1; r >= 0; --r) {
    double s = rhs[r];
    for (int k = r + 1; k < n; ++k) s -= m[r * n + k] * x[k];
    x[r] = s / m[r * n + r];
  }
  return x;
}

}  // namespace

// P_n^{(a,b)} from P_0 = 1, P_1 = ((a+b+2)x + a-b)/2 and the standard
// recurrence 2(k+1)(k+a+b+1)(2k+a+b) P_{k+1} =
//   (2k+a+b+1)[(2k+a+b+2)(2k+a+b) x + a^2-b^2] P_k - 2(k+a)(k+b)(2k+a+b+2) P_{k-1}.
double jacobi_p(int n, double a, double b, double x) {
  INSEM_REQUIRE(n >= 0, "Jacobi degree must be non-negative");
  if (n == 0) return 1.0;
  double prev = 1.0;
  double cur = 0.5 * ((a + b + 2.0) * x + (a - b));
  for (int k = 1; k < n; ++k) {
    const double s = 2.0 * k + a + b;
    const double lead = 2.0 * (k + 1) * (k + a + b + 1) * s;
    const double shift = (s + 1) * (a * a - b * b);
    const double slope = (s + 1) * (s + 2) * s;
    const double back = 2.0 * (k + a) * (k + b) * (s + 2);
    const double next = ((shift + slope * x) * cur - back * prev) / lead;
    prev = cur;
    cur = next;
  }
  return cur;
}

double jacobi_dp(int n, double a, double b, double x) {
  if (n == 0) return 0.0;
  return 0.5 * (n + a + b + 1) * jacobi_p(n - 1, a + 1, b + 1, x);
}

// Newton iteration with deflation of the roots already found; start values
// are Chebyshev points, averaged with the previous root.
std::vector<double> jacobi_zeros(int n, double a, double b) {
  INSEM_REQUIRE(n >= 0, "number of roots must be non-negative");
  std::vector<double> z(n);
  for (int k = 0; k < n; ++k) {
    double t = -std::cos((2.0 * k + 1.0) * kPi / (2.0 * n));
    if (k > 0) t = 0.5 * (t + z[k - 1]);
    for (int it = 0; it < 200; ++it) {
      double defl = 0.0;
      for (int i = 0; i < k; ++i) defl += 1.0 / (t - z[i]);
      const double p = jacobi_p(n, a, b, t);
      const double step = -p / (jacobi_dp(n, a, b, t) - defl * p);
      t += step;
      if (std::fabs(step) < 1e-16) break;
    }
    z[k] = t;
  }
  return z;
}

Rule gauss_legendre_rule(int n) {
  INSEM_REQUIRE(n >= 1, "Gauss-Legendre rule needs at least one point");
  Rule r;
  r.points = jacobi_zeros(n, 0.0, 0.0);
  r.weights.resize(n);
  for (int i = 0; i < n; ++i) {
    const double x = r.points[i];
    const double d = jacobi_dp(n, 0.0, 0.0, x);
    r.weights[i] = 2.0 / ((1.0 - x * x) * d * d);
  }
  return r;
}

// End points plus the zeros of P_{n-2}^{(alpha+1, beta+1)}; weights make the
// rule exact for the Legendre polynomials P_0..P_{n-1} against the weight
// function, with the moments taken by a Gauss-Legendre rule that is exact
// for them.
Rule lobatto_jacobi_rule(int n, int alpha, int beta) {
  INSEM_REQUIRE(n >= 2, "Gauss-Lobatto rule needs at least two points");
  INSEM_REQUIRE(alpha >= 0 && beta >= 0, "Gauss-Lobatto-Jacobi exponents must be non-negative integers");
  Rule r;
  r.points.push_back(-1.0);
  for (double z : jacobi_zeros(n - 2, alpha + 1.0, beta + 1.0)) r.points.push_back(z);
  r.points.push_back(1.0);
  const Rule gl = gauss_legendre_rule(n + alpha + beta + 2);
  std::vector<double> mom(n, 0.0);
  for (int q = 0; q < gl.size(); ++q) {
    const double x = gl.points[q];
    const double w = gl.weights[q] * std::pow(1.0 - x, alpha) * std::pow(1.0 + x, beta);
    for (int k = 0; k < n; ++k) mom[k] += w * jacobi_p(k, 0.0, 0.0, x);
  }
  std::vector<double> v(static_cast<std::size_t>(n) * n);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) v[k * n + j] = jacobi_p(k, 0.0, 0.0, r.points[j]);
  r.weights = dense_solve(std::move(v), std::move(mom), n);
  return r;
}

Lagrange1D::Lagrange1D(std::vector<double> nodes) : x_(std::move(nodes)) {
  const int n = size();
  INSEM_REQUIRE(n >= 1, "Lagrange basis needs at least one node");
  lambda_.assign(n, 1.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      if (k == j) continue;
      const double d = x_[j] - x_[k];
      INSEM_REQUIRE(d != 0.0, "Lagrange nodes must be distinct");
      lambda_[j] /= d;
    }
}

int Lagrange1D::coincident(double t) const {
  for (int j = 0; j < size(); ++j)
    if (std::fabs(t - x_[j]) <= 1e-14 * (1.0 + std::fabs(x_[j]))) return j;
  return -1;
}

void Lagrange1D::values(double t, double* out) const {
  const int n = size();
  const int at = coincident(t);
  if (at >= 0) {
    for (int j = 0; j < n; ++j) out[j] = j == at ? 1.0 : 0.0;
    return;
  }
  double den = 0.0;
  for (int j = 0; j < n; ++j) {
    out[j] = lambda_[j] / (t - x_[j]);
    den += out[j];
  }
  for (int j = 0; j < n; ++j) out[j] /= den;
}

void Lagrange1D::derivatives(double t, double* out) const {
  const int n = size();
  const int at = coincident(t);
  if (at >= 0) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == at) continue;
      out[j] = (lambda_[j] / lambda_[at]) / (x_[at] - x_[j]);
      diag -= out[j];
    }
    out[at] = diag;
    return;
  }
  values(t, out);
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int k = 0; k < n; ++k)
      if (k != j) s += 1.0 / (t - x_[k]);
    out[j] *= s;
  }
}

std::vector<double> Lagrange1D::values(double t) const {
  std::vector<double> v(size());
  values(t, v.data());
  return v;
}

std::vector<double> Lagrange1D::derivatives(double t) const {
  std::vector<double> v(size());
  derivatives(t, v.data());
  return v;
}

std::vector<double> Lagrange1D::derivative_matrix() const {
  const int n = size();
  std::vector<double> d(static_cast<std::size_t>(n) * n);
  for (int i = 0; i < n; ++i) derivatives(x_[i], &d[static_cast<std::size_t>(i) * n]);
  return d;
}

}  // namespace insem_amd
