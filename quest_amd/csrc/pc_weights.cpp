// pc_weights.cpp -- host FP64 predictor-corrector weights.
//
// Replaces PredictorCorrector::Weights / WeightsBuilder
// (src/integrator.cpp:5-102).  The weights of a 22-step multistep predictor
// and corrector are the real part of the complex minimum-norm least-squares
// solution of  M w = exp(lambda t_f),  M = [exp(lambda t_j) | lambda exp(lambda
// t_j) (| lambda exp(lambda t_f))],  for n_lambda points lambda on a
// "semidisk" contour of the given radius.  The reference uses Eigen's
// JacobiSVD; this file solves with a one-sided (Hestenes) complex Jacobi SVD of
// M^H, which keeps high relative accuracy on the ~1e-11 singular values of
// this ill-conditioned (cond ~ 3e13) system.
#include <cmath>
#include <complex>
#include <vector>

#include "quest_internal.hpp"

namespace quest {
namespace {

using cd = std::complex<double>;

// integrator.cpp:5-14
cd semidisk(double t)
{
  const cd iu(0, 1);
  if(0 <= t && t < 1) return iu * t;
  if(1 <= t && t < M_PI + 1) return std::exp(iu * (t + M_PI_2 - 1));
  if(M_PI + 1 <= t && t < M_PI + 2) return iu * (t - (M_PI + 2));
  return 0;
}

// Minimum-norm least-squares solution of A x = b for A (rows x cols, rows <
// cols, row-major), by one-sided Jacobi on the columns of B = A^H.
std::vector<cd> min_norm_solve(const std::vector<cd> &A, int rows, int cols,
                               const std::vector<cd> &b)
{
  // B: cols x rows, column j of B = conj(row j of A); stored column-major.
  std::vector<cd> B(static_cast<size_t>(cols) * rows);
  for(int j = 0; j < rows; ++j)
    for(int i = 0; i < cols; ++i) B[static_cast<size_t>(j) * cols + i] = std::conj(A[static_cast<size_t>(j) * cols + i]);
  std::vector<cd> V(static_cast<size_t>(rows) * rows, 0.0);  // column-major
  for(int j = 0; j < rows; ++j) V[static_cast<size_t>(j) * rows + j] = 1.0;

  auto col = [&](std::vector<cd> &M, int ld, int j) { return M.data() + static_cast<size_t>(j) * ld; };
  const double eps = 1e-15;
  for(int sweep = 0; sweep < 60; ++sweep) {
    double off = 0;
    for(int p = 0; p < rows - 1; ++p) {
      for(int q = p + 1; q < rows; ++q) {
        cd *bp = col(B, cols, p), *bq = col(B, cols, q);
        double a = 0, bb = 0;
        cd g = 0;
        for(int i = 0; i < cols; ++i) {
          a += std::norm(bp[i]);
          bb += std::norm(bq[i]);
          g += std::conj(bp[i]) * bq[i];
        }
        const double ag = std::abs(g);
        if(ag == 0) continue;
        const double rel = ag / std::sqrt(a * bb);
        if(rel > off) off = rel;
        if(rel < eps) continue;
        const cd ph = g / ag;  // e^{i phi}
        const double zeta = (bb - a) / (2 * ag);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
        const double c = 1 / std::sqrt(1 + t * t), s = c * t;
        for(int i = 0; i < cols; ++i) {
          const cd xp = bp[i], xq = bq[i];
          bp[i] = c * xp - s * std::conj(ph) * xq;
          bq[i] = s * ph * xp + c * xq;
        }
        cd *vp = col(V, rows, p), *vq = col(V, rows, q);
        for(int i = 0; i < rows; ++i) {
          const cd xp = vp[i], xq = vq[i];
          vp[i] = c * xp - s * std::conj(ph) * xq;
          vq[i] = s * ph * xp + c * xq;
        }
      }
    }
    if(off < eps) break;
  }
  // B V = W with orthogonal columns w_j = sigma_j u_j; A = V Sigma U^H, so
  // x = sum_j w_j (v_j^H b) / sigma_j^2 over the numerically non-zero sigma_j.
  std::vector<double> sig2(rows);
  double smax = 0;
  for(int j = 0; j < rows; ++j) {
    double s2 = 0;
    const cd *w = col(B, cols, j);
    for(int i = 0; i < cols; ++i) s2 += std::norm(w[i]);
    sig2[j] = s2;
    smax = std::max(smax, std::sqrt(s2));
  }
  const double thresh = smax * rows * 2.220446049250313e-16;
  std::vector<cd> x(cols, 0.0);
  for(int j = 0; j < rows; ++j) {
    if(std::sqrt(sig2[j]) <= thresh) continue;
    const cd *v = col(V, rows, j);
    cd vb = 0;
    for(int i = 0; i < rows; ++i) vb += std::conj(v[i]) * b[i];
    const cd f = vb / sig2[j];
    const cd *w = col(B, cols, j);
    for(int i = 0; i < cols; ++i) x[i] += w[i] * f;
  }
  return x;
}

}  // namespace

// integrator.cpp:34-102
int pc_weights(int n_lambda, int n_time, double radius, double *ps, double *cs,
               double *future_coef)
{
  if(n_lambda < 1 || n_time < 2 || !ps || !cs || !future_coef) return -1;
  std::vector<cd> lambdas(n_lambda);
  for(int i = 0; i < n_lambda; ++i) {
    const double x = (M_PI + 2) * i / n_lambda;  // LinSpaced(n_lambda + 1, 0, pi + 2)[i]
    lambdas[i] = radius * semidisk(x);
  }
  std::vector<double> times(n_time);
  for(int j = 0; j < n_time; ++j) times[j] = -1 + 2.0 * j / (n_time - 1);
  const double timestep = 2.0 / (n_time - 1);
  const double future_time = 1 + timestep;

  std::vector<cd> rhs(n_lambda);
  for(int i = 0; i < n_lambda; ++i) rhs[i] = std::exp(lambdas[i] * future_time);

  const int pc = 2 * n_time, cc = 2 * n_time + 1;
  std::vector<cd> P(static_cast<size_t>(n_lambda) * pc), Cm(static_cast<size_t>(n_lambda) * cc);
  for(int i = 0; i < n_lambda; ++i) {
    for(int j = 0; j < n_time; ++j) {
      const cd e = std::exp(lambdas[i] * times[j]);
      P[static_cast<size_t>(i) * pc + j] = e;
      P[static_cast<size_t>(i) * pc + n_time + j] = e * lambdas[i];
      Cm[static_cast<size_t>(i) * cc + j] = e;
      Cm[static_cast<size_t>(i) * cc + n_time + j] = e * lambdas[i];
    }
    Cm[static_cast<size_t>(i) * cc + 2 * n_time] = lambdas[i] * rhs[i];
  }
  const std::vector<cd> wp = min_norm_solve(P, n_lambda, pc, rhs);
  const std::vector<cd> wc = min_norm_solve(Cm, n_lambda, cc, rhs);

  const double step_factor = (n_time - 1) / 2.0;
  for(int j = 0; j < n_time; ++j) {
    ps[j] = wp[j].real();
    ps[n_time + j] = wp[n_time + j].real() * step_factor;
    cs[j] = wc[j].real();
    cs[n_time + j] = wc[n_time + j].real() * step_factor;
  }
  *future_coef = wc[2 * n_time].real() * step_factor;
  return 0;
}

}  // namespace quest
