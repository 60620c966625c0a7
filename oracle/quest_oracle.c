/*
 * quest_oracle.c -- FP64 CPU restatement of QuEST's retarded-interaction step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see quest_oracle.h).  Built by
 * oracle/Makefile with -ffp-contract=off so that, like the reference's default
 * x86-64 Release build, no multiply-add is fused.
 *
 * Pins for each function live in tests/test_oracle.py:
 *   qo_lagrange_weights   closed-form identities (partition of unity, exact
 *                         value/derivative reproduction of polynomials of
 *                         degree <= order, delta property) + a numpy
 *                         polynomial-fit brute force
 *   qo_idx2coord          brute-force enumeration of the lower triangle
 *   qo_split_delay        hand-computed cases + numpy floor/modf restatement
 *   qo_rotating_coefs     analytic cases (d || z, dr || x), symmetry
 *                         G(dr) = G(-dr), omega -> 0 limit vs the lab form
 *   qo_lab_coefs          analytic cases
 *   qo_evaluate_history   N = 2 closed form, linearity, zero history => 0,
 *                         O(N^2) brute force on tiny N (numpy)
 *   qo_evaluate_observers qo_evaluate_history on every observer (both frames)
 *   qo_pulse_evaluate     closed-form Gaussian
 *   qo_liouville_rhs      the reference's known-answer test
 *                         (test/quantum_dot_test.cpp:13-43), restated
 *   qo_isfinite           the reference's test/history_test.cpp:13-53
 *   qo_solve              N = 1 analytic pulse-area limit (rho00 =
 *                         cos^2(theta/2)), free decay, mode 0 == mode 1
 *   qo_solve_pulses       one-pulse list == qo_solve bitwise; two equal
 *                         pulses == one of twice the amplitude; N = 1 with
 *                         pulse areas t1, t2: rho00 = cos^2((t1 + t2)/2);
 *                         empty list keeps the ground state
 */
#include "quest_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static inline cplx ldc(const double *p) { return CMPLX(p[0], p[1]); }
static inline void stc(double *p, cplx z)
{
  p[0] = creal(z);
  p[1] = cimag(z);
}

static double now_seconds(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* lagrange_set.cpp:17-41                                                     */
/* ------------------------------------------------------------------------ */
void qo_lagrange_weights(double x, int order, double dt, double *w)
{
  double *w0 = w, *w1 = w + (order + 1), *w2 = w + 2 * (order + 1);
  for(int basis_id = 0; basis_id <= order; ++basis_id) {
    double d0_product = 1, d1_sum = 0, d2_sum = 0;
    for(int m = 0; m <= order; ++m) {
      if(m == basis_id) continue;
      d0_product *= (x - m) / (double)(basis_id - m);
      d1_sum -= 1 / (x - m);
      d2_sum -= pow(x - m, -2);
    }
    w0[basis_id] = d0_product;
    w1[basis_id] = w0[basis_id] * d1_sum;
    w2[basis_id] = w0[basis_id] * d2_sum + pow(w1[basis_id], 2) / w0[basis_id];
  }
  for(int i = 0; i <= order; ++i) {
    w1[i] *= pow(dt, -1);
    w2[i] *= pow(dt, -2);
  }
}

/* ------------------------------------------------------------------------ */
/* history_interaction.cpp:72-97                                              */
/* ------------------------------------------------------------------------ */
void qo_idx2coord(int64_t idx, int64_t *row, int64_t *col)
{
  const int64_t r = (int64_t)floor((sqrt((double)(1 + 8 * idx)) + 1) / 2.0);
  *row = r;
  *col = idx - r * (r - 1) / 2;
}

static inline int64_t coord2idx(int64_t row, int64_t col)
{
  if(col > row) {
    int64_t t = row;
    row = col;
    col = t;
  }
  return row * (row - 1) / 2 + col;
}

/* separation() is r_obs - r_src (quantum_dot.cpp:21-24); the norm is
 * sqrt((dx^2 + dy^2) + dz^2) in FP64; the delay is norm / (c0 * dt) split by
 * modf (history_interaction.cpp:31-32,88-97). */
int qo_split_delay(const double *pos_src, const double *pos_obs, double c0,
                   double dt, double *frac)
{
  const double dx = pos_obs[0] - pos_src[0];
  const double dy = pos_obs[1] - pos_src[1];
  const double dz = pos_obs[2] - pos_src[2];
  const double norm = sqrt((dx * dx + dy * dy) + dz * dz);
  const double delay = norm / (c0 * dt);
  double idelay;
  *frac = modf(delay, &idelay);
  return (int)idelay;
}

/* ------------------------------------------------------------------------ */
/* green_function.h:32-55: the three spatial dyads, returned as real 3x3.     */
/* ------------------------------------------------------------------------ */
static void spatial_dyads(double c, const double *dr, double D[3][3][3])
{
  const double sq = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
  const double nrm = sqrt(sq);
  for(int i = 0; i < 3; ++i) {
    for(int j = 0; j < 3; ++j) {
      const double rr = dr[i] * dr[j] / sq;     /* rhat_dyadic */
      const double id = (i == j) ? 1.0 : 0.0;
      const double i3 = id - 3 * rr;            /* identity_minus_3rsq */
      const double i1 = id - rr;                /* identity_minus_rsq  */
      D[0][i][j] = i3 * pow(c, 2) / pow(nrm, 3);
      D[1][i][j] = i3 * c / sq;
      D[2][i][j] = i1 / nrm;
    }
  }
}

/* d_obs . (M d_src) for a complex 3x3 M (history_interaction.cpp:41-42) */
static cplx project(const cplx M[3][3], const double *d_obs, const double *d_src)
{
  cplx acc = 0;
  for(int i = 0; i < 3; ++i) {
    cplx row = 0;
    for(int j = 0; j < 3; ++j) row += M[i][j] * d_src[j];
    acc += d_obs[i] * row;
  }
  return acc;
}

/* rotating_green_function.cpp:3-24 */
void qo_rotating_coefs(const qo_params *p, const double *dr, const double *w,
                       const double *d_obs, const double *d_src, double *coef)
{
  const int M = p->order;
  const double *w0 = w, *w1 = w + (M + 1), *w2 = w + 2 * (M + 1);
  const double c = p->c0, omega = p->omega;
  const double k = p->mu0 / (4 * M_PI * p->hbar);
  double D[3][3][3];
  spatial_dyads(c, dr, D);
  const double nrm = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
  /* std::exp(-iu * omega * norm / c) = exp(i * ((-omega * norm) / c)) */
  const double ang = (-omega * nrm) / c;
  const cplx phase = CMPLX(cos(ang), sin(ang));
  for(int i = 0; i <= M; ++i) {
    const cplx a0 = w0[i];
    const cplx a1 = CMPLX(w1[i], omega * w0[i]);
    const cplx a2 = CMPLX(w2[i] - pow(omega, 2) * w0[i], 2.0 * omega * w1[i]);
    cplx G[3][3];
    for(int r = 0; r < 3; ++r)
      for(int s = 0; s < 3; ++s)
        G[r][s] = -k * phase * (D[0][r][s] * a0 + D[1][r][s] * a1 + D[2][r][s] * a2);
    stc(coef + 2 * i, project((const cplx(*)[3])G, d_obs, d_src));
  }
}

/* green_function.cpp:3-20 */
void qo_lab_coefs(const qo_params *p, const double *dr, const double *w,
                  const double *d_obs, const double *d_src, double *coef)
{
  const int M = p->order;
  const double k = p->mu0 / (4 * M_PI * p->hbar);
  double D[3][3][3];
  spatial_dyads(p->c0, dr, D);
  for(int i = 0; i <= M; ++i) {
    cplx G[3][3];
    for(int r = 0; r < 3; ++r)
      for(int s = 0; s < 3; ++s) {
        G[r][s] = 0;
        for(int term = 0; term < 3; ++term)
          G[r][s] += k * D[term][r][s] * w[term * (M + 1) + i];
      }
    stc(coef + 2 * i, project((const cplx(*)[3])G, d_obs, d_src));
  }
}

/* ------------------------------------------------------------------------ */
/* history_interaction.cpp:21-45                                              */
/* ------------------------------------------------------------------------ */
int qo_build_table(const qo_params *p, int kind, int n, const double *pos,
                   const double *dip, int32_t *delays, double *coef)
{
  const int M = p->order;
  const int64_t npairs = (int64_t)n * (n - 1) / 2;
#pragma omp parallel
  {
    double *w = (double *)malloc(sizeof(double) * 3 * (M + 1));
#pragma omp for schedule(static)
    for(int64_t pair = 0; pair < npairs; ++pair) {
      int64_t src, obs;
      qo_idx2coord(pair, &src, &obs);
      const double *ps = pos + 3 * src, *po = pos + 3 * obs;
      const double dr[3] = {po[0] - ps[0], po[1] - ps[1], po[2] - ps[2]};
      double frac;
      delays[pair] = qo_split_delay(ps, po, p->c0, p->dt, &frac);
      qo_lagrange_weights(frac, M, p->dt, w);
      if(kind == 0)
        qo_rotating_coefs(p, dr, w, dip + 3 * obs, dip + 3 * src,
                          coef + 2 * (M + 1) * pair);
      else
        qo_lab_coefs(p, dr, w, dip + 3 * obs, dip + 3 * src,
                     coef + 2 * (M + 1) * pair);
    }
    free(w);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* history_interaction.cpp:47-70                                              */
/* ------------------------------------------------------------------------ */
static inline const double *slot_ptr(const double *y, int n_slots, int window,
                                     int dot, int t)
{
  return y + 4 * ((int64_t)dot * n_slots + (t + window));
}

/* RotatingDyadic::polarization_prefactor -> rho01 (rotating_green_function.h:17-21);
 * Dyadic::polarization_prefactor -> 2 Re rho01 (green_function.h:23-27). */
static inline cplx prefactor(int kind, const double *sol)
{
  if(kind == 0) return CMPLX(sol[2], sol[3]);
  return CMPLX(2 * sol[2], 0.0);
}

void qo_evaluate_history(const qo_params *p, int kind, int n, int n_slots,
                         const int32_t *delays, const double *coef,
                         const double *y, int step, double *out)
{
  const int M = p->order, W = p->window;
  const int64_t npairs = (int64_t)n * (n - 1) / 2;
  cplx *res = (cplx *)calloc(n, sizeof(cplx));
  for(int64_t pair = 0; pair < npairs; ++pair) {
    int64_t src, obs;
    qo_idx2coord(pair, &src, &obs);
    const int s = step - delays[pair];
    for(int i = 0; i <= M; ++i) {
      if(s - i < -W) continue;
      const cplx cf = ldc(coef + 2 * ((M + 1) * pair + i));
      res[src] += prefactor(kind, slot_ptr(y, n_slots, W, (int)obs, s - i)) * cf;
      res[obs] += prefactor(kind, slot_ptr(y, n_slots, W, (int)src, s - i)) * cf;
    }
  }
  for(int i = 0; i < n; ++i) stc(out + 2 * i, res[i]);
  free(res);
}

/* The same evaluation for a subset of observers, each pair's delay and
 * coefficients formed on the fly exactly as qo_build_table forms them
 * (history_interaction.cpp:21-45, then :47-70 for one observer): no table, so
 * that a few thousand observers of a system of 1e5 dots fit in memory.  The
 * delay split and the coefficients are symmetric in the pair, so each
 * (observer, source) term is the table's term for that unordered pair; only
 * the order of the additions differs from qo_evaluate_history.
 * rho01: [n_t][n] complex samples at t = 0 .. n_t - 1 (the prehistory t < 0
 * is 0, main.cpp:32-36); needs step < n_t.  out: [n_obs] complex. */
void qo_evaluate_observers(const qo_params *p, int kind, int n, const double *pos,
                           const double *dip, int n_t, const double *rho01, int step,
                           int n_obs, const int32_t *obs, double *out)
{
  const int M = p->order;
  if(step < 0 || step >= n_t) return;
#pragma omp parallel
  {
    double *w = (double *)malloc(sizeof(double) * 3 * (M + 1));
    double *cf = (double *)malloc(sizeof(double) * 2 * (M + 1));
#pragma omp for schedule(dynamic, 4)
    for(int j = 0; j < n_obs; ++j) {
      const int o = obs[j];
      const double *po = pos + 3 * o;
      cplx acc = 0;
      for(int src = 0; src < n; ++src) {
        if(src == o) continue;
        const double *ps = pos + 3 * src;
        const double dr[3] = {po[0] - ps[0], po[1] - ps[1], po[2] - ps[2]};
        double frac;
        const int delay = qo_split_delay(ps, po, p->c0, p->dt, &frac);
        qo_lagrange_weights(frac, M, p->dt, w);
        if(kind == 0)
          qo_rotating_coefs(p, dr, w, dip + 3 * o, dip + 3 * src, cf);
        else
          qo_lab_coefs(p, dr, w, dip + 3 * o, dip + 3 * src, cf);
        const int s = step - delay;
        for(int i = 0; i <= M; ++i) {
          const int t = s - i;
          if(t < 0) continue;
          const double *r = rho01 + 2 * ((int64_t)t * n + src);
          const cplx pf = kind == 0 ? CMPLX(r[0], r[1]) : CMPLX(2 * r[0], 0.0);
          acc += pf * ldc(cf + 2 * i);
        }
      }
      stc(out + 2 * j, acc);
    }
    free(w);
    free(cf);
  }
}

/* ------------------------------------------------------------------------ */
/* pulse.cpp:15-20, pulse_interaction.cpp:9-20, math_utils.cpp:27-30          */
/* ------------------------------------------------------------------------ */
static inline double gaussian(double t) { return exp(-pow(t, 2) / 2); }

static cplx pulse_one(const qo_pulse *pl, double hbar, const double *r,
                      const double *d, double t)
{
  const double kr = (pl->k[0] * r[0] + pl->k[1] * r[1]) + pl->k[2] * r[2];
  const double arg = kr - pl->freq * (t - pl->delay);
  const double g = gaussian(arg / pl->width);
  const double e0 = (pl->amplitude / 2 * pl->pol[0]) * g;
  const double e1 = (pl->amplitude / 2 * pl->pol[1]) * g;
  const double e2 = (pl->amplitude / 2 * pl->pol[2]) * g;
  return ((e0 * d[0] + e1 * d[1]) + e2 * d[2]) / hbar;
}

void qo_pulse_evaluate(const qo_pulse *pulse, double hbar, int n,
                       const double *pos, const double *dip, double t,
                       double *out)
{
  for(int i = 0; i < n; ++i)
    stc(out + 2 * i, pulse_one(pulse, hbar, pos + 3 * i, dip + 3 * i, t));
}

/* ------------------------------------------------------------------------ */
/* quantum_dot.cpp:10-19                                                      */
/* ------------------------------------------------------------------------ */
static inline void rhs(const cplx *rho, cplx rabi, double laser_freq,
                       double freq, double t1, double t2, cplx *out)
{
  const cplx iu = CMPLX(0.0, 1.0);
  out[0] = -iu * (rabi * conj(rho[1]) - conj(rabi) * rho[1]) - (rho[0] - 1.0) / t1;
  out[1] = -iu * (rabi * (1.0 - 2.0 * rho[0]) + rho[1] * (laser_freq - freq)) -
           rho[1] / t2;
}

void qo_liouville_rhs(const double *rho, const double *rabi, double laser_freq,
                      double freq, double t1, double t2, double *out)
{
  cplx r[2] = {ldc(rho), ldc(rho + 2)}, o[2];
  rhs(r, ldc(rabi), laser_freq, freq, t1, t2, o);
  stc(out, o[0]);
  stc(out + 2, o[1]);
}

/* history.cpp:11-15 */
int qo_isfinite(const double *sol)
{
  return isfinite(sol[0]) && isfinite(sol[1]) && isfinite(sol[2]) &&
         isfinite(sol[3]);
}

/* ------------------------------------------------------------------------ */
/* integrator.cpp:104-207 with main.cpp:25-55                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
  const qo_params *p;
  int kind, n, n_slots, n_steps;
  const double *pos, *dip, *freq, *t1, *t2;
  const qo_pulse *pulse;     /* [n_pulses] */
  int n_pulses;
  const double *ps, *cs;
  double fc;
  double *y, *yp;            /* [n][n_slots][2] complex */
  int32_t *delays;
  double *coef;
  /* mode 1 */
  int64_t *near_start;       /* [n+1] */
  int32_t *near_src;
  int64_t *near_pair;
  cplx *past, *pulse_now, *hist;
} qo_run;

static inline cplx *Y(qo_run *R, int dot, int t, int k)
{
  return (cplx *)(R->y + 4 * ((int64_t)dot * R->n_slots + (t + R->p->window)) + 2 * k);
}
static inline cplx *YP(qo_run *R, int dot, int t, int k)
{
  return (cplx *)(R->yp + 4 * ((int64_t)dot * R->n_slots + (t + R->p->window)) + 2 * k);
}

static void predictor(qo_run *R, int step)
{
  const int W = R->p->window;
  const int start = step - W;
  const double dt = R->p->dt;
#pragma omp parallel for schedule(static)
  for(int i = 0; i < R->n; ++i) {
    for(int k = 0; k < 2; ++k) *Y(R, i, step, k) = 0;
    for(int h = 0; h < W; ++h)
      for(int k = 0; k < 2; ++k)
        *Y(R, i, step, k) += *Y(R, i, start + h, k) * R->ps[h] +
                             *YP(R, i, start + h, k) * R->ps[W + h] * dt;
  }
}

static void corrector(qo_run *R, int step)
{
  const int W = R->p->window;
  const int start = step - W;
  const double dt = R->p->dt;
#pragma omp parallel for schedule(static)
  for(int i = 0; i < R->n; ++i) {
    for(int k = 0; k < 2; ++k) *Y(R, i, step, k) = R->fc * *YP(R, i, step, k) * dt;
    for(int h = 0; h < W; ++h)
      for(int k = 0; k < 2; ++k)
        *Y(R, i, step, k) += *Y(R, i, start + h, k) * R->cs[h] +
                             *YP(R, i, start + h, k) * R->cs[W + h] * dt;
  }
}

static void apply_rhs(qo_run *R, int step, const cplx *field)
{
#pragma omp parallel for schedule(static)
  for(int i = 0; i < R->n; ++i) {
    cplx rho[2] = {*Y(R, i, step, 0), *Y(R, i, step, 1)}, o[2];
    rhs(rho, field[i], R->p->omega, R->freq[i], R->t1[i], R->t2[i], o);
    *YP(R, i, step, 0) = o[0];
    *YP(R, i, step, 1) = o[1];
  }
}

/* The pulse interactions of dot i at time t, accumulated from zero in list
 * order (integrator.cpp:173-179 sums its interactions with std::accumulate;
 * main.cpp:43-46 lists one pulse interaction, several add the same way). */
static cplx pulses_at(const qo_run *R, int i, double t)
{
  cplx f = 0;
  for(int q = 0; q < R->n_pulses; ++q)
    f = f + pulse_one(R->pulse + q, R->p->hbar, R->pos + 3 * i, R->dip + 3 * i, t);
  return f;
}

/* mode 0: Integrator::evaluator (integrator.cpp:171-185) with the pulse
 * interaction first and the history interaction second (main.cpp:43-46). */
static void evaluator_reference(qo_run *R, int step, cplx *field)
{
  const double t = step * R->p->dt;
  for(int i = 0; i < R->n; ++i) R->pulse_now[i] = pulses_at(R, i, t);
  qo_evaluate_history(R->p, R->kind, R->n, R->n_slots, R->delays, R->coef, R->y,
                      step, (double *)R->hist);
  for(int i = 0; i < R->n; ++i) field[i] = (0 + R->pulse_now[i]) + R->hist[i];
}

/* mode 1 helpers: the pairs with delay >= 1 read only finished steps
 * (s - i <= step - 1), so their sum is the same for all 11 evaluations of a
 * step; the zero-delay pairs read the current iterate through tap 0. */
static void build_near_lists(qo_run *R)
{
  const int n = R->n;
  int64_t *count = (int64_t *)calloc(n + 1, sizeof(int64_t));
  const int64_t npairs = (int64_t)n * (n - 1) / 2;
  for(int64_t pair = 0; pair < npairs; ++pair) {
    if(R->delays[pair] != 0) continue;
    int64_t src, obs;
    qo_idx2coord(pair, &src, &obs);
    count[src + 1]++;
    count[obs + 1]++;
  }
  for(int i = 0; i < n; ++i) count[i + 1] += count[i];
  R->near_start = count;
  R->near_src = (int32_t *)malloc(sizeof(int32_t) * (count[n] + 1));
  R->near_pair = (int64_t *)malloc(sizeof(int64_t) * (count[n] + 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * n);
  for(int i = 0; i < n; ++i) fill[i] = count[i];
  for(int64_t pair = 0; pair < npairs; ++pair) {
    if(R->delays[pair] != 0) continue;
    int64_t src, obs;
    qo_idx2coord(pair, &src, &obs);
    R->near_src[fill[src]] = (int32_t)obs;
    R->near_pair[fill[src]++] = pair;
    R->near_src[fill[obs]] = (int32_t)src;
    R->near_pair[fill[obs]++] = pair;
  }
  free(fill);
}

static inline cplx pair_sum(qo_run *R, int64_t pair, int other, int step)
{
  const int M = R->p->order, W = R->p->window;
  const int s = step - R->delays[pair];
  cplx acc = 0;
  for(int i = 0; i <= M; ++i) {
    if(s - i < -W) continue;
    const cplx cf = ldc(R->coef + 2 * ((M + 1) * pair + i));
    acc += prefactor(R->kind, (const double *)Y(R, other, s - i, 0)) * cf;
  }
  return acc;
}

static void past_sums(qo_run *R, int step)
{
  const int n = R->n;
#pragma omp parallel for schedule(dynamic, 16)
  for(int a = 0; a < n; ++a) {
    cplx acc = 0;
    for(int b = 0; b < n; ++b) {
      if(b == a) continue;
      const int64_t pair = coord2idx(a, b);
      if(R->delays[pair] == 0) continue;
      acc += pair_sum(R, pair, b, step);
    }
    R->past[a] = acc;
  }
}

static void evaluator_split(qo_run *R, int step, cplx *field)
{
#pragma omp parallel for schedule(static)
  for(int a = 0; a < R->n; ++a) {
    cplx near = 0;
    for(int64_t e = R->near_start[a]; e < R->near_start[a + 1]; ++e)
      near += pair_sum(R, R->near_pair[e], R->near_src[e], step);
    field[a] = R->pulse_now[a] + (R->past[a] + near);
  }
}

static int all_finite(qo_run *R, int step)
{
  for(int i = 0; i < R->n; ++i) {
    if(!qo_isfinite((const double *)Y(R, i, step, 0)) ||
       !qo_isfinite((const double *)YP(R, i, step, 0)))
      return 0;
  }
  return 1;
}

static int run_alloc(qo_run *R, const qo_params *p, int kind, int n,
                     const double *pos, const double *dip, const double *freq,
                     const double *t1, const double *t2, const qo_pulse *pulse,
                     int n_pulses, const double *ps, const double *cs, double fc,
                     int n_steps)
{
  memset(R, 0, sizeof(*R));
  R->p = p;
  R->kind = kind;
  R->n = n;
  R->n_steps = n_steps;
  R->n_slots = p->window + n_steps;
  R->pos = pos;
  R->dip = dip;
  R->freq = freq;
  R->t1 = t1;
  R->t2 = t2;
  R->pulse = pulse;
  R->n_pulses = pulse ? n_pulses : 0;
  R->ps = ps;
  R->cs = cs;
  R->fc = fc;
  const int64_t nv = 4 * (int64_t)n * R->n_slots;
  R->y = (double *)calloc(nv, sizeof(double));
  R->yp = (double *)calloc(nv, sizeof(double));
  const int64_t npairs = (int64_t)n * (n - 1) / 2;
  R->delays = (int32_t *)malloc(sizeof(int32_t) * (npairs + 1));
  R->coef = (double *)malloc(sizeof(double) * 2 * (p->order + 1) * (npairs + 1));
  R->past = (cplx *)calloc(n, sizeof(cplx));
  R->pulse_now = (cplx *)calloc(n, sizeof(cplx));
  R->hist = (cplx *)calloc(n, sizeof(cplx));
  if(!R->y || !R->yp || !R->delays || !R->coef) return -1;
  /* main.cpp:32-36: ground state for t in [-window, 0]; derivatives 0 */
  for(int t = -p->window; t <= 0; ++t)
    for(int i = 0; i < n; ++i) *Y(R, i, t, 0) = 1.0;
  return 0;
}

static void run_free(qo_run *R)
{
  free(R->y);
  free(R->yp);
  free(R->delays);
  free(R->coef);
  free(R->past);
  free(R->pulse_now);
  free(R->hist);
  free(R->near_start);
  free(R->near_src);
  free(R->near_pair);
}

int qo_solve(const qo_params *p, int kind, int n, const double *pos,
             const double *dip, const double *freq, const double *t1,
             const double *t2, const qo_pulse *pulse, const double *ps,
             const double *cs, double future_coef, int n_steps, int mode,
             int record_every, double *rec, double *final_y, double *final_yp,
             int *bad_step, double *seconds)
{
  return qo_solve_pulses(p, kind, n, pos, dip, freq, t1, t2, pulse, pulse ? 1 : 0,
                         ps, cs, future_coef, n_steps, mode, record_every, rec,
                         final_y, final_yp, bad_step, seconds);
}

int qo_solve_pulses(const qo_params *p, int kind, int n, const double *pos,
                    const double *dip, const double *freq, const double *t1,
                    const double *t2, const qo_pulse *pulses, int n_pulses,
                    const double *ps, const double *cs, double future_coef,
                    int n_steps, int mode, int record_every, double *rec,
                    double *final_y, double *final_yp, int *bad_step,
                    double *seconds)
{
  qo_run R;
  if(run_alloc(&R, p, kind, n, pos, dip, freq, t1, t2, pulses, n_pulses, ps, cs,
               future_coef, n_steps) != 0) {
    run_free(&R);
    return -1;
  }
  qo_build_table(p, kind, n, pos, dip, R.delays, R.coef);
  if(mode == 1) build_near_lists(&R);
  cplx *field = (cplx *)calloc(n, sizeof(cplx));
  int status = 0;
  if(bad_step) *bad_step = -1;
  const double t0 = now_seconds();
  const int NUM_CORRECTOR_STEPS = 10; /* integrator.cpp:3 */
  int last = -1;
  for(int step = 0; step < n_steps; ++step) {
    if(mode == 1) {
      past_sums(&R, step);
      const double t = step * p->dt;
      for(int i = 0; i < n; ++i) R.pulse_now[i] = pulses_at(&R, i, t);
    }
    predictor(&R, step);
    if(mode == 0) evaluator_reference(&R, step, field);
    else evaluator_split(&R, step, field);
    apply_rhs(&R, step, field);
    for(int m = 0; m < NUM_CORRECTOR_STEPS; ++m) {
      corrector(&R, step);
      if(mode == 0) evaluator_reference(&R, step, field);
      else evaluator_split(&R, step, field);
      apply_rhs(&R, step, field);
    }
    last = step;
    if(rec && record_every > 0 && (step + 1) % record_every == 0) {
      double *dst = rec + 4 * (int64_t)n * ((step + 1) / record_every - 1);
      for(int i = 0; i < n; ++i) {
        stc(dst + 4 * i, *Y(&R, i, step, 0));
        stc(dst + 4 * i + 2, *Y(&R, i, step, 1));
      }
    }
    if(!all_finite(&R, step)) {
      if(bad_step) *bad_step = step;
      status = 1;
      break;
    }
  }
  if(seconds) *seconds = now_seconds() - t0;
  if(last >= 0) {
    for(int i = 0; i < n; ++i) {
      if(final_y) {
        stc(final_y + 4 * i, *Y(&R, i, last, 0));
        stc(final_y + 4 * i + 2, *Y(&R, i, last, 1));
      }
      if(final_yp) {
        stc(final_yp + 4 * i, *YP(&R, i, last, 0));
        stc(final_yp + 4 * i + 2, *YP(&R, i, last, 1));
      }
    }
  }
  free(field);
  run_free(&R);
  return status;
}

/* Terms (pair, tap, direction) the evaluations of steps [start, start + count)
 * add: mode 0, 11 reference evaluations per step, each adding both directions
 * of every (pair, tap) that history_interaction.cpp:57 does not skip; mode 1,
 * the past sum once per step over the pairs of delay >= 1 plus the zero-delay
 * pairs at each of the 11 evaluations (pair_sum's skip).  The same condition
 * as the evaluation loops, counted in a separate pass. */
static int64_t count_taps(const qo_run *R, int mode, int start, int count)
{
  const int M = R->p->order, W = R->p->window;
  const int64_t npairs = (int64_t)R->n * (R->n - 1) / 2;
  int64_t taps = 0;
  for(int step = start; step < start + count; ++step)
    for(int64_t pair = 0; pair < npairs; ++pair) {
      const int s = step - R->delays[pair];
      int64_t live = 0;
      for(int i = 0; i <= M; ++i)
        if(s - i >= -W) ++live;
      const int evals = (mode == 0 || R->delays[pair] == 0) ? 11 : 1;
      taps += 2 * live * evals;
    }
  return taps;
}

/* A smooth synthetic history in every slot t in [-window, start): rho01 =
 * A(t) e^{i phi_i}, A(t) = 0.2 (1 - e^{-(t + window + 1) / 60}), phi_i = 0.7 i,
 * rho00 = 1 - |rho01|^2, and the time derivatives (per ps) of the same
 * formulas.  No slot that steps >= start read is zero, so with start beyond
 * the largest delay plus the order the reference's skip of taps before the
 * prehistory (history_interaction.cpp:57) never applies. */
static void fill_synthetic_history(qo_run *R, int start)
{
  const int W = R->p->window;
  const double dt = R->p->dt;
  for(int i = 0; i < R->n; ++i) {
    const cplx ph = cexp(I * (0.7 * i));
    for(int t = -W; t < start; ++t) {
      const double e = exp(-(t + W + 1) / 60.0);
      const double a = 0.2 * (1.0 - e), da = 0.2 * e / (60.0 * dt);
      *Y(R, i, t, 0) = 1.0 - a * a;
      *Y(R, i, t, 1) = a * ph;
      *YP(R, i, t, 0) = -2.0 * a * da;
      *YP(R, i, t, 1) = da * ph;
    }
  }
}

double qo_time_steps(const qo_params *p, int n, const double *pos, const double *dip,
                     const double *freq, const double *t1, const double *t2,
                     const qo_pulse *pulse, const double *ps, const double *cs,
                     double future_coef, int mode, int start_step, int n_steps,
                     double *setup_seconds, double *step_seconds, int64_t *taps)
{
  if(start_step < 0 || n_steps < 1) return NAN;
  qo_run R;
  if(run_alloc(&R, p, 0, n, pos, dip, freq, t1, t2, pulse, pulse ? 1 : 0, ps,
               cs, future_coef, start_step + n_steps) != 0) {
    run_free(&R);
    return NAN;
  }
  if(start_step > 0) fill_synthetic_history(&R, start_step);
  const double t0 = now_seconds();
  qo_build_table(p, 0, n, pos, dip, R.delays, R.coef);
  if(mode == 1) build_near_lists(&R);
  const double t1s = now_seconds();
  cplx *field = (cplx *)calloc(n, sizeof(cplx));
  for(int step = start_step; step < start_step + n_steps; ++step) {
    if(mode == 1) {
      past_sums(&R, step);
      for(int i = 0; i < n; ++i) R.pulse_now[i] = pulses_at(&R, i, step * p->dt);
    }
    predictor(&R, step);
    if(mode == 0) evaluator_reference(&R, step, field);
    else evaluator_split(&R, step, field);
    apply_rhs(&R, step, field);
    for(int m = 0; m < 10; ++m) {
      corrector(&R, step);
      if(mode == 0) evaluator_reference(&R, step, field);
      else evaluator_split(&R, step, field);
      apply_rhs(&R, step, field);
    }
  }
  const double t2s = now_seconds();
  double checksum = 0;
  for(int i = 0; i < n; ++i) checksum += creal(*Y(&R, i, start_step + n_steps - 1, 0));
  if(setup_seconds) *setup_seconds = t1s - t0;
  if(step_seconds) *step_seconds = t2s - t1s;
  if(taps) *taps = count_taps(&R, mode, start_step, n_steps);
  free(field);
  run_free(&R);
  return checksum;
}

void qo_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
