/*
 * quest_oracle.h -- FP64 CPU restatement of QuEST's retarded-interaction step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product in quest_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; no product code path links or calls it.
 *
 * Every function names the reference lines it restates (paths relative to the
 * QuEST source tree).  Pins (what checks each function against something other
 * than itself) are listed in tests/test_oracle.py; a function whose only check
 * is agreement with another restatement says "parity unpinned" below.
 *
 * Layout conventions (all arrays row-major, complex = interleaved re,im):
 *   pos, dip         [n][3] doubles
 *   delays           [P] int32, P = n(n-1)/2, lower-triangle pair order
 *   coef             [P][order+1] complex
 *   history y, yp    [n][window + n_steps][2] complex: slot t + window holds
 *                    time index t (t in [-window, n_steps)), element 0 = rho00,
 *                    element 1 = rho01 (History::HistoryArray, history.h:13-14)
 */
#ifndef QUEST_ORACLE_H
#define QUEST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double c0;       /* speed of light, um/ps          (configuration.h:16) */
  double hbar;     /* meV ps                          */
  double mu0;      /* meV ps^2 / (e^2 um)             */
  double omega;    /* laser (rotating-frame) angular frequency, rad/ps */
  double dt;       /* timestep, ps                    */
  int order;       /* Lagrange interpolation order    */
  int window;      /* PC history window (main.cpp:30 uses 22) */
} qo_params;

typedef struct {
  double amplitude, delay, width, freq;  /* pulse.h:22 */
  double k[3];                           /* wavevector  */
  double pol[3];                         /* polarisation, used as given */
} qo_pulse;

/* lagrange_set.cpp:17-41 -- weights[3][order+1] at fractional delay x. */
void qo_lagrange_weights(double x, int order, double dt, double *w);

/* history_interaction.cpp:80-86 (64-bit: the reference's int form overflows
 * 8*idx for n >= 23171). */
void qo_idx2coord(int64_t idx, int64_t *row, int64_t *col);

/* history_interaction.cpp:29-35,88-97: integer delay and fraction of the pair
 * (src, obs): |r_obs - r_src| / (c0 dt), FP64, no contraction. */
int qo_split_delay(const double *pos_src, const double *pos_obs, double c0,
                   double dt, double *frac);

/* rotating_green_function.cpp:3-24 + green_function.h:32-55 +
 * history_interaction.cpp:40-43: coef[i] = d_obs . G_i . d_src. */
void qo_rotating_coefs(const qo_params *p, const double *dr, const double *w,
                       const double *d_obs, const double *d_src, double *coef);

/* green_function.cpp:3-20 (lab-frame Dyadic; polarization prefactor
 * 2 Re rho01, green_function.h:23-27). */
void qo_lab_coefs(const qo_params *p, const double *dr, const double *w,
                  const double *d_obs, const double *d_src, double *coef);

/* history_interaction.cpp:21-45.  kind 0 = rotating, 1 = lab. */
int qo_build_table(const qo_params *p, int kind, int n, const double *pos,
                   const double *dip, int32_t *delays, double *coef);

/* history_interaction.cpp:47-70.  rho01 is read from y (element 1 of each
 * slot); out[n] complex. kind selects the polarization prefactor. */
void qo_evaluate_history(const qo_params *p, int kind, int n, int n_slots,
                         const int32_t *delays, const double *coef,
                         const double *y, int step, double *out);

/* qo_evaluate_history for the observers obs[n_obs] only, coefficients formed
 * on the fly (no table).  rho01: [n_t][n] complex, t = 0 .. n_t - 1. */
void qo_evaluate_observers(const qo_params *p, int kind, int n, const double *pos,
                           const double *dip, int n_t, const double *rho01, int step,
                           int n_obs, const int32_t *obs, double *out);

/* pulse.cpp:15-20 + pulse_interaction.cpp:9-20 + math_utils.cpp:27-30. */
void qo_pulse_evaluate(const qo_pulse *pulse, double hbar, int n,
                       const double *pos, const double *dip, double t,
                       double *out);

/* quantum_dot.cpp:10-19. rho[2] complex, rabi complex, out[2] complex. */
void qo_liouville_rhs(const double *rho, const double *rabi, double laser_freq,
                      double freq, double t1, double t2, double *out);

/* history.cpp:11-15 */
int qo_isfinite(const double *sol);

/* Whole run: integrator.cpp:104-185 with main.cpp:25-55 wiring
 * (pulse interaction first, then history; ground-state prehistory).
 *   ps, cs     [2][window] (row 0: value weights, row 1: derivative weights
 *              already multiplied by the step factor, integrator.cpp:95-101)
 *   mode 0     the reference method: pair table and 11 full scatter sums per
 *              step, serial.
 *   mode 1     the same arithmetic as a gather: per-target sums, the past-only
 *              part (pairs with delay >= 1) once per step and the zero-delay
 *              pairs at every evaluation; OpenMP over targets.
 *   rec        [n_steps / record_every][n][2] complex (rho00, rho01) of steps
 *              record_every-1, 2*record_every-1, ...; may be NULL.
 *   final_y/yp [n][2] complex state and derivative of the last step; may be NULL.
 * Returns 0, or 1 with *bad_step set when a non-finite value is produced
 * (integrator.cpp:187-207: the first step whose state is not finite). */
int qo_solve(const qo_params *p, int kind, int n, const double *pos,
             const double *dip, const double *freq, const double *t1,
             const double *t2, const qo_pulse *pulse, const double *ps,
             const double *cs, double future_coef, int n_steps, int mode,
             int record_every, double *rec, double *final_y, double *final_yp,
             int *bad_step, double *seconds);

/* qo_solve with a list of pulse interactions (n_pulses >= 0), summed in list
 * order ahead of the history interaction (integrator.cpp:173-179). */
int qo_solve_pulses(const qo_params *p, int kind, int n, const double *pos,
                    const double *dip, const double *freq, const double *t1,
                    const double *t2, const qo_pulse *pulses, int n_pulses,
                    const double *ps, const double *cs, double future_coef,
                    int n_steps, int mode, int record_every, double *rec,
                    double *final_y, double *final_yp, int *bad_step,
                    double *seconds);

/* The step loop only, for timing: builds the table (and, mode 1, the
 * zero-delay lists; their time goes to *setup_seconds), fills every slot
 * before start_step with a smooth non-zero synthetic history (start_step > 0),
 * then runs steps [start_step, start_step + n_steps) with the pulse on:
 * mode 0 the reference method (predictor + 11 scatter sums, serial), mode 1
 * the split gather (OpenMP).  With start_step beyond the largest delay plus
 * the order, no tap is skipped.  *taps: the (pair, tap, direction) terms the
 * timed evaluations added.  Returns a checksum (NAN on failure). */
double qo_time_steps(const qo_params *p, int n, const double *pos, const double *dip,
                     const double *freq, const double *t1, const double *t2,
                     const qo_pulse *pulse, const double *ps, const double *cs,
                     double future_coef, int mode, int start_step, int n_steps,
                     double *setup_seconds, double *step_seconds, int64_t *taps);

/* Number of OpenMP threads the oracle uses (1 = the reference's serial run). */
void qo_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
