#include "basis.hpp"

#include <cmath>
#include <utility>

#include "error.hpp"

namespace insem_amd {
namespace {

// The weight w(x) = (1 - x)^a (1 + x)^b on [-1, 1] through its three-term
// recurrence: the monic orthogonal polynomials satisfy
//   p_{k+1} = (x - diag[k]) p_k - off2[k] p_{k-1},
// so the symmetric tridiagonal matrix with diagonal diag and off-diagonal
// sqrt(off2[1..m-1]) has the zeros of p_m as its eigenvalues (Golub-Welsch).
struct JacobiWeight {
  int a, b;

  // integral of w(x) x^0: 2^(a+b+1) a! b! / (a+b+1)!
  template <class T = double>
  T mass() const {
    T m = std::ldexp(T(1), a + b + 1);
    for (int i = 1; i <= a; ++i) m *= i;
    for (int i = 1; i <= b; ++i) m *= i;
    for (int i = 2; i <= a + b + 1; ++i) m /= i;
    return m;
  }
  // integral of w(x) x^1
  template <class T = double>
  T first_moment() const {
    return mass<T>() * T(b - a) / T(a + b + 2);
  }

  template <class T = double>
  T diag(int k) const {
    const T s = T(2 * k + a + b);
    if (k == 0) return T(b - a) / T(a + b + 2);
    return T(b * b - a * a) / (s * (s + 2));
  }
  template <class T = double>
  T off2(int k) const {  // k >= 1
    const T s = T(2 * k + a + b);
    const T num = T(4) * k * (k + a) * (k + b) * (k + a + b);
    return num / (s * s * (s + 1) * (s - 1));
  }

  // Orthonormal polynomials q_0 .. q_m at x (q_k = p_k / ||p_k||), and q_m'.
  template <class T>
  void orthonormal(int m, T x, T* q, T* dqm) const {
    T prev = 0, dprev = 0;
    T cur = T(1) / std::sqrt(mass<T>()), dcur = 0;
    q[0] = cur;
    for (int k = 0; k < m; ++k) {
      const T bk = k > 0 ? std::sqrt(off2<T>(k)) : T(0);
      const T bn = std::sqrt(off2<T>(k + 1));
      const T next = ((x - diag<T>(k)) * cur - bk * prev) / bn;
      const T dnext = (cur + (x - diag<T>(k)) * dcur - bk * dprev) / bn;
      prev = cur;
      dprev = dcur;
      cur = next;
      dcur = dnext;
      q[k + 1] = cur;
    }
    if (dqm) *dqm = dcur;
  }

  // Number of eigenvalues of the m x m Jacobi matrix below x (Sturm count
  // of the LDL^T pivots of T - x I).
  int count_below(int m, double x) const {
    int n = 0;
    double piv = 1.0;
    for (int k = 0; k < m; ++k) {
      piv = (diag(k) - x) - (k > 0 ? off2(k) / piv : 0.0);
      if (piv == 0.0) piv = -1e-300;
      if (piv < 0.0) ++n;
    }
    return n;
  }

  // The m zeros of p_m, ascending: bisection on the Sturm count to the
  // resolution of doubles, then Newton steps on q_m in extended precision
  // (coefficients included), so that each zero rounds to the double nearest
  // the true root.
  std::vector<long double> zeros(int m) const {
    std::vector<long double> z(m);
    std::vector<long double> q(static_cast<std::size_t>(m) + 1);
    double lo_prev = -1.0;
    for (int j = 0; j < m; ++j) {
      double lo = lo_prev, hi = 1.0;
      for (int it = 0; it < 1100; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (count_below(m, mid) > j)
          hi = mid;
        else
          lo = mid;
      }
      lo_prev = lo;
      long double x = 0.5L * (static_cast<long double>(lo) + hi);
      for (int it = 0; it < 8; ++it) {
        long double dq = 0;
        orthonormal(m, x, q.data(), &dq);
        if (dq == 0) break;
        const long double step = q[m] / dq;
        if (!(std::fabs(step) < 1e-8L)) break;
        x -= step;
        if (step == 0) break;
      }
      z[j] = x;
    }
    return z;
  }

  // Gauss weight (Christoffel number) of the m-point rule at its zero x:
  // 1 / sum_{k < m} q_k(x)^2.
  long double christoffel(int m, long double x) const {
    std::vector<long double> q(static_cast<std::size_t>(m) + 1);
    orthonormal<long double>(m, x, q.data(), nullptr);
    long double s = 0;
    for (int k = 0; k < m; ++k) s += q[k] * q[k];
    return 1 / s;
  }
};

}  // namespace

// The interior Lobatto nodes are the Gauss nodes of (1 - x^2) w(x); the
// cardinal function of an interior node is (1 - x^2) / (1 - x_j^2) times
// that Gauss rule's, so its weight is the Gauss weight / (1 - x_j^2).  The
// two end weights then follow from the rule integrating 1 and x exactly.
Rule lobatto_rule(int n, int alpha, int beta) {
  INSEM_REQUIRE(n >= 2, "a Lobatto rule has at least 2 points (asked for " << n << ")");
  INSEM_REQUIRE(alpha >= 0 && beta >= 0, "Lobatto rules take integer weight exponents >= 0");
  const JacobiWeight outer{alpha, beta}, inner{alpha + 1, beta + 1};
  const int m = n - 2;
  Rule r;
  r.points.assign(static_cast<std::size_t>(n), 0.0);
  r.weights.assign(static_cast<std::size_t>(n), 0.0);
  // everything in extended precision, rounded to double at the end
  const std::vector<long double> z = inner.zeros(m);
  long double sum0 = 0, sum1 = 0;
  for (int j = 0; j < m; ++j) {
    const long double w = inner.christoffel(m, z[j]) / ((1 - z[j]) * (1 + z[j]));
    r.points[j + 1] = static_cast<double>(z[j]);
    r.weights[j + 1] = static_cast<double>(w);
    sum0 += w;
    sum1 += w * z[j];
  }
  r.points.front() = -1.0;
  r.points.back() = 1.0;
  const long double even = outer.mass<long double>() - sum0, odd = outer.first_moment<long double>() - sum1;
  r.weights.front() = static_cast<double>((even - odd) / 2);
  r.weights.back() = static_cast<double>((even + odd) / 2);
  return r;
}

Lagrange1D::Lagrange1D(std::vector<double> nodes) : nodes_(std::move(nodes)) {
  const int n = size();
  INSEM_REQUIRE(n >= 1, "an interpolation basis needs at least one node");
  // b_j by successive division by each node difference, k ascending: these
  // roundings reach the float differentiation matrices of the kernels, so
  // they are fixed.
  bary_.assign(static_cast<std::size_t>(n), 1.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      if (k == j) continue;
      INSEM_REQUIRE(nodes_[j] != nodes_[k], "interpolation nodes " << j << " and " << k << " coincide");
      bary_[j] /= nodes_[j] - nodes_[k];
    }
}

// Points closer to a node than rounding of the located reference
// coordinate takes the node's exact cardinal values.
int Lagrange1D::node_at(double t) const {
  for (int j = 0; j < size(); ++j)
    if (std::fabs(t - nodes_[j]) <= 1e-14 * (1.0 + std::fabs(nodes_[j]))) return j;
  return -1;
}

void Lagrange1D::values(double t, double* out) const {
  const int n = size();
  const int at = node_at(t);
  for (int j = 0; j < n; ++j) out[j] = 0.0;
  if (at >= 0) {
    out[at] = 1.0;
    return;
  }
  // second barycentric form: l_j = (b_j / (t - x_j)) / sum_k b_k / (t - x_k)
  double total = 0.0;
  for (int j = 0; j < n; ++j) total += out[j] = bary_[j] / (t - nodes_[j]);
  for (int j = 0; j < n; ++j) out[j] /= total;
}

// Row i of the differentiation matrix: off the diagonal l_j'(x_i) =
// (b_j / b_i) / (x_i - x_j); the diagonal makes the row sum vanish (so
// constants differentiate to zero), summed in ascending j.
void Lagrange1D::derivatives_at_node(int i, double* out) const {
  const int n = size();
  double row = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    out[j] = (bary_[j] / bary_[i]) / (nodes_[i] - nodes_[j]);
    row -= out[j];
  }
  out[i] = row;
}

void Lagrange1D::derivatives(double t, double* out) const {
  const int at = node_at(t);
  if (at >= 0) {
    derivatives_at_node(at, out);
    return;
  }
  // l_j'(t) = l_j(t) * d/dt log prod_{k != j} (t - x_k)
  const int n = size();
  values(t, out);
  for (int j = 0; j < n; ++j) {
    double logd = 0.0;
    for (int k = 0; k < n; ++k)
      if (k != j) logd += 1.0 / (t - nodes_[k]);
    out[j] *= logd;
  }
}

std::vector<double> Lagrange1D::values(double t) const {
  std::vector<double> v(static_cast<std::size_t>(size()));
  values(t, v.data());
  return v;
}

std::vector<double> Lagrange1D::derivatives(double t) const {
  std::vector<double> v(static_cast<std::size_t>(size()));
  derivatives(t, v.data());
  return v;
}

std::vector<double> Lagrange1D::derivative_matrix() const {
  const int n = size();
  std::vector<double> d(static_cast<std::size_t>(n) * n);
  for (int i = 0; i < n; ++i) derivatives_at_node(i, &d[static_cast<std::size_t>(i) * n]);
  return d;
}

}  // namespace insem_amd
