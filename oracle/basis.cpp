// CPU ORACLE -- test infrastructure only (never linked by the product).
//
// Lobatto-Legendre basis used by Trixi's DGSEM (LobattoLegendreBasis, [ext],
// reached from every elixir through `DGSEM(polydeg = ...)`, e.g.
// examples/euler/dry_air/global_circulation/elixir_energy_baroclinic_instability.jl:27).
// Restated from Kopriva, "Implementing Spectral Methods for PDEs" (2009),
// Algorithms 22-25 (Legendre recursion, LGL nodes by Newton iteration) and
// Algorithm 37 (derivative matrix from barycentric weights with the
// negative-sum diagonal).  Pinned in tests by closed-form nodes/weights for
// n = 4, 5, quadrature exactness, polynomial exactness of D and the SBP
// property M D + D^T M = diag(-1, 0, ..., 0, 1).
#include "counted.h"
#include "oracle.h"

#include <cmath>
#include <vector>

namespace {

// q(x) = L_{N+1}(x) - L_{N-1}(x), its derivative, and L_N(x)
void lgl_q_and_l(int N, double x, double* q, double* dq, double* LN) {
  double Lm2 = 1.0, Lm1 = x, dLm2 = 0.0, dLm1 = 1.0;
  double L = 0, dL = 0;
  for (int k = 2; k <= N; ++k) {
    L = ((2.0 * k - 1.0) * x * Lm1 - (k - 1.0) * Lm2) / k;
    dL = dLm2 + (2.0 * k - 1.0) * Lm1;
    Lm2 = Lm1; Lm1 = L; dLm2 = dLm1; dLm1 = dL;
  }
  // Lm1 = L_N, Lm2 = L_{N-1}
  int k = N + 1;
  double Lp1 = ((2.0 * k - 1.0) * x * Lm1 - (k - 1.0) * Lm2) / k;
  double dLp1 = dLm2 + (2.0 * k - 1.0) * Lm1;
  *q = Lp1 - Lm2;
  *dq = dLp1 - dLm2;
  *LN = Lm1;
}

double legendre(int N, double x) {
  if (N == 0) return 1.0;
  double a = 1.0, b = x;
  for (int k = 2; k <= N; ++k) {
    double c = ((2.0 * k - 1.0) * x * b - (k - 1.0) * a) / k;
    a = b; b = c;
  }
  return b;
}

}  // namespace

extern "C" void ora_lgl_nodes_weights(int n, double* nodes, double* weights) {
  const int N = n - 1;
  if (N == 1) {
    nodes[0] = -1.0; nodes[1] = 1.0; weights[0] = weights[1] = 1.0;
    return;
  }
  nodes[0] = -1.0; nodes[N] = 1.0;
  weights[0] = weights[N] = 2.0 / (N * (N + 1.0));
  for (int j = 1; j <= (N + 1) / 2 - 1; ++j) {
    double x = -std::cos((j + 0.25) * M_PI / N - 3.0 / (8.0 * N * M_PI * (j + 0.25)));
    for (int it = 0; it < 100; ++it) {
      double q, dq, LN;
      lgl_q_and_l(N, x, &q, &dq, &LN);
      double dx = -q / dq;
      x += dx;
      if (std::fabs(dx) <= 4e-16 * std::fabs(x)) break;
    }
    double q, dq, LN;
    lgl_q_and_l(N, x, &q, &dq, &LN);
    nodes[j] = x;
    nodes[N - j] = -x;
    weights[j] = weights[N - j] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
  if (N % 2 == 0) {
    double LN = legendre(N, 0.0);
    nodes[N / 2] = 0.0;
    weights[N / 2] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
}

extern "C" void ora_derivative_matrix(int n, const double* x, double* D) {
  std::vector<double> bw(n, 1.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      if (k != j) bw[j] *= (x[j] - x[k]);
  for (int j = 0; j < n; ++j) bw[j] = 1.0 / bw[j];
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      D[i * n + j] = (bw[j] / bw[i]) / (x[i] - x[j]);
      diag -= D[i * n + j];
    }
    D[i * n + i] = diag;
  }
}

// D_split = 2 D + diag(1/w_1, 0, ..., 0, -1/w_n): the flux-differencing
// operator whose diagonal vanishes by the SBP property.
extern "C" void ora_derivative_split(int n, const double* x, const double* w,
                                     double* Ds) {
  ora_derivative_matrix(n, x, Ds);
  for (int k = 0; k < n * n; ++k) Ds[k] *= 2.0;
  Ds[0] += 1.0 / w[0];
  Ds[n * n - 1] -= 1.0 / w[n - 1];
}

// Logarithmic mean and its inverse: the templates of counted.h (Ranocha et
// al. 2023, the form Trixi's ln_mean implements) in FP64
extern "C" double ora_ln_mean(double x, double y) { return ora_ln_mean_t(x, y); }

extern "C" double ora_inv_ln_mean(double x, double y) { return ora_inv_ln_mean_t(x, y); }
