// Lobatto-Legendre nodal basis for DGSEM (the basis Trixi builds from
// `DGSEM(polydeg = ...)`, e.g.
// examples/euler/dry_air/global_circulation/elixir_energy_baroclinic_instability.jl:27).
// Nodes are the roots of (1 - x^2) P'_N found by Newton iteration on
// P_{N+1} - P_{N-1} from Chebyshev-Gauss-Lobatto guesses; the derivative
// matrix uses barycentric weights with the negative-sum diagonal.
#include <cmath>
#include <vector>

#include "atmo_internal.hpp"

namespace atmo {

static void legendre_pair(int N, double x, double* LN, double* LNm1) {
  double a = 1.0, b = x;
  if (N == 0) { *LN = 1.0; *LNm1 = 0.0; return; }
  for (int k = 2; k <= N; ++k) {
    const double c = ((2.0 * k - 1.0) * x * b - (k - 1.0) * a) / k;
    a = b; b = c;
  }
  *LN = b; *LNm1 = a;
}

Basis::Basis(int n_) : n(n_), x(n_), w(n_), D(n_ * n_), Dsplit(n_ * n_), bary(n_) {
  const int N = n - 1;
  x[0] = -1.0; x[N] = 1.0;
  w[0] = w[N] = 2.0 / (N * (N + 1.0));
  for (int j = 1; j < (N + 1) / 2; ++j) {
    // Chebyshev-Gauss-Lobatto first guess, refined by Newton on
    // q = P_{N+1} - P_{N-1} (whose interior roots are the LGL nodes)
    double t = -std::cos(M_PI * j / N);
    for (int it = 0; it < 100; ++it) {
      // P_{N+1}, P_N, P_{N-1} and derivatives by the three-term recursion
      double p0 = 1.0, p1 = t, d0 = 0.0, d1 = 1.0;
      for (int k = 2; k <= N + 1; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        const double d2 = d0 + (2.0 * k - 1.0) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
      }
      // here p1 = P_{N+1}, p0 = P_N, d1 = P'_{N+1}, d0 = P'_N
      double PNm1, dPNm1;
      {
        double q0 = 1.0, q1 = t, e0 = 0.0, e1 = 1.0;
        for (int k = 2; k <= N - 1; ++k) {
          const double q2 = ((2.0 * k - 1.0) * t * q1 - (k - 1.0) * q0) / k;
          const double e2 = e0 + (2.0 * k - 1.0) * q1;
          q0 = q1; q1 = q2; e0 = e1; e1 = e2;
        }
        PNm1 = (N - 1 == 0) ? 1.0 : q1;
        dPNm1 = (N - 1 == 0) ? 0.0 : e1;
      }
      const double q = p1 - PNm1, dq = d1 - dPNm1;
      const double delta = -q / dq;
      t += delta;
      if (std::fabs(delta) <= 1e-15 * std::fabs(t) + 1e-300) break;
    }
    double LN, LNm1;
    legendre_pair(N, t, &LN, &LNm1);
    x[j] = t;
    x[N - j] = -t;
    w[j] = w[N - j] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
  if (N % 2 == 0 && N > 0) {
    double LN, LNm1;
    legendre_pair(N, 0.0, &LN, &LNm1);
    x[N / 2] = 0.0;
    w[N / 2] = 2.0 / (N * (N + 1.0) * LN * LN);
  }
  for (int j = 0; j < n; ++j) {
    double p = 1.0;
    for (int k = 0; k < n; ++k)
      if (k != j) p *= (x[j] - x[k]);
    bary[j] = 1.0 / p;
  }
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      D[i * n + j] = (bary[j] / bary[i]) / (x[i] - x[j]);
      s += D[i * n + j];
    }
    D[i * n + i] = -s;
  }
  for (int k = 0; k < n * n; ++k) Dsplit[k] = 2.0 * D[k];
  Dsplit[0] += 1.0 / w[0];
  Dsplit[n * n - 1] -= 1.0 / w[n - 1];
  // the diagonal vanishes analytically (SBP); store it as exact zero
  for (int i = 0; i < n; ++i) Dsplit[i * n + i] = 0.0;
}

}  // namespace atmo
