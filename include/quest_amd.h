/*
 * quest_amd.h -- C ABI of the MI355X (gfx950) retarded-interaction step.
 *
 * This is the drop-in boundary for QuEST's direct pairwise history
 * interaction and the predictor-corrector loop around it.  No C++ or HIP types
 * cross it: plain pointers, sizes and status codes.  Complex numbers are
 * interleaved (re, im) doubles.  Dot arrays are in the caller's order; the
 * library reorders internally and converts back.
 *
 * Every entry point cites the reference interface it replaces (paths relative
 * to the QuEST source tree).  The C++ mirror of the reference classes that
 * sits on top of this ABI is quest_amd/host/quest.hpp.
 *
 * Threading: one context owns one HIP stream; a context is not thread-safe,
 * distinct contexts are independent.
 */
#ifndef QUEST_AMD_H
#define QUEST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The C++ mirror rethrows them as the reference's exception
 * types: QUEST_ERR_DOMAIN -> std::domain_error("unbounded history values at
 * or before step N") (integrator.cpp:200-207); everything else ->
 * std::runtime_error with quest_last_error(). */
#define QUEST_OK 0
#define QUEST_ERR_ARG 1      /* invalid argument                              */
#define QUEST_ERR_HIP 2      /* HIP runtime error (message says which call)   */
#define QUEST_ERR_DOMAIN 3   /* non-finite state; see quest_bad_step          */
#define QUEST_ERR_STATE 4    /* call out of sequence for this context         */
#define QUEST_ERR_NOMEM 5    /* device or host allocation failed              */

/* Arithmetic of the O(N^2) pair sum.  FP64 everywhere, or FP32 pair
 * arithmetic with FP64 accumulation per target (the Bloch/PC update, the
 * delay split of the zero-delay pairs and all per-dot work stay FP64). */
#define QUEST_PREC_FP64 0
#define QUEST_PREC_MIXED 1

/* Green's function of the history interaction.
 * ROTATING: GreenFunction::RotatingDyadic (rotating_green_function.cpp:3-24),
 *           prefactor rho01 (rotating_green_function.h:17-21) -- what main.cpp
 *           wires (main.cpp:40-46).
 * LAB:      GreenFunction::Dyadic (green_function.cpp:3-20), prefactor
 *           2 Re rho01 (green_function.h:23-27). */
#define QUEST_DYADIC_ROTATING 0
#define QUEST_DYADIC_LAB 1

typedef struct quest_ctx quest_ctx;

/* Configuration values the path reads (configuration.h:12-19) plus the
 * choices main.cpp hard-codes (history window 22, main.cpp:30). */
typedef struct {
  double c0;           /* config.c0, um/ps                                  */
  double hbar;         /* config.hbar, meV ps                               */
  double mu0;          /* config.mu0                                        */
  double laser_omega;  /* config.omega, rad/ps (rotating frame)             */
  double dt;           /* config.dt, ps                                     */
  int interp_order;    /* config.interpolation_order, 1..6                  */
  int history_window;  /* PC window; main.cpp:30 uses 22 (2..64)            */
  int precision;       /* QUEST_PREC_*                                      */
  int dyadic;          /* QUEST_DYADIC_*                                    */
  int device;          /* HIP device ordinal                                */
  int source_chunk;    /* sources per pair-kernel work item; 0 = default
                          (the largest system in 48 (mixed) or 24 (FP64)
                          equal chunks -- 8 in both with
                          QUEST_PAIR_TIME_BLOCKED, taken when the method is
                          set -- of at least 128 / 256 sources, halved until
                          the work list fills the GPU); see quest_autotune  */
  int item_claim;      /* how the pair kernel's waves take work items:
                          QUEST_CLAIM_AUTO (static until quest_autotune,
                          which times both and keeps the faster),
                          QUEST_CLAIM_STATIC (a fixed stride over an
                          XCD-aware list: each (system, source chunk) pair
                          on one XCD)
                          or QUEST_CLAIM_DYNAMIC (a device counter; balances
                          items of uneven size, in list order)               */
  int source_loop;     /* mixed precision's pair loop over an item's sources:
                          QUEST_LOOP_AUTO (pipelined until quest_autotune
                          times both), QUEST_LOOP_PLAIN, or
                          QUEST_LOOP_PIPELINED (the next source's delay split
                          and interpolant gathers issued before this
                          source's arithmetic, so their cache latency
                          overlaps it).  Same arithmetic, same results.  FP64
                          contexts use the plain loop.                       */
  int accumulation;    /* mixed precision's per-target sum over a block of
                          64 sources before it is folded into the FP64 sum:
                          QUEST_ACC_BLOCK (plain FP32 block sums; the
                          default) or QUEST_ACC_COMPENSATED (Kahan-compensated
                          FP32 block sums: each addition's rounding error is
                          carried in a second FP32 sum and the pair is folded
                          into FP64 per block, so the block sum's error no
                          longer grows with the block length).  FP64
                          contexts ignore it.                                */
  int block_steps;     /* steps per pass of QUEST_PAIR_TIME_BLOCKED: 0 = auto
                          (16 when the ordered pairs of delay below 16 are
                          under 2 % of all -- large systems -- else 8), 8 or
                          16.  The other pair methods ignore it.            */
} quest_params;

#define QUEST_CLAIM_AUTO 0
#define QUEST_CLAIM_STATIC 1
#define QUEST_CLAIM_DYNAMIC 2
#define QUEST_ACC_BLOCK 0
#define QUEST_ACC_COMPENSATED 1
#define QUEST_LOOP_AUTO 0
#define QUEST_LOOP_PLAIN 1
#define QUEST_LOOP_PIPELINED 2

/* Pulse (pulse.h:10-24): amplitude, delay, width (dimensionless), freq,
 * wavevector, polarization.  The polarization is used as given, as on the
 * reference's read_pulse_config path (pulse.cpp:29-46). */
typedef struct {
  double amplitude, delay, width, freq;
  double wavevector[3];
  double polarization[3];
} quest_pulse;

/* ---------------------------------------------------------------- context */

/* Replaces the construction of HistoryInteraction (history_interaction.cpp:
 * 5-19; the coefficient table of :21-45 is not built -- the kernel evaluates
 * the Green's function per pair on the fly) and of the DotVector it reads
 * (quantum_dot.h:21-48: pos, freq, damping (T1, T2), dipole).
 * pos, dip: [n][3]; freq, t1, t2: [n].  freq/t1/t2 may be NULL when only
 * quest_hist_evaluate is used. */
int quest_ctx_create(const quest_params *params, int n_dots, const double *pos,
                     const double *dip, const double *freq, const double *t1,
                     const double *t2, quest_ctx **out);

/* Many independent systems in one context (no coupling between systems):
 * system s has n_dots[s] dots; pos, dip, freq, t1, t2 are the systems'
 * arrays concatenated in system order, and every per-dot array in and out of
 * the other entry points uses the same order.  Each pair-sum launch covers
 * all systems (one persistent kernel over a work list of (system, target
 * tile, source chunk) items) and quest_run replays one captured graph per
 * step for all of them.  No counterpart in the reference, which runs one
 * system per process (main.cpp:17-63); for callers with many small systems. */
int quest_batch_create(const quest_params *params, int n_systems, const int *n_dots,
                       const double *pos, const double *dip, const double *freq,
                       const double *t1, const double *t2, quest_ctx **out);

void quest_ctx_destroy(quest_ctx *ctx);

/* Message of the last failed call on ctx (or of the last failed
 * quest_ctx_create on this thread when ctx is NULL). */
const char *quest_last_error(const quest_ctx *ctx);

/* ------------------------------------------ HistoryInteraction::evaluate */

/* Replaces HistoryInteraction::evaluate(step) (history_interaction.cpp:47-70):
 * out[n] (complex) = the retarded field projected on each dot's dipole, from
 * the history of rho01.  rho01 [n] (complex) is each dot's current rho01 at
 * `step` -- the only history slot the integrator changes between calls
 * (integrator.cpp:142-185).  Values at earlier steps are the ones passed with
 * the last call for that step; steps before 0 are the ground-state prehistory
 * (rho01 = 0, main.cpp:32-36).  The first call on a context is at step 0;
 * `step` must not decrease and may grow by at most one per call (else
 * QUEST_ERR_STATE).  rho01 and out hold n complex values each. */
int quest_hist_evaluate(quest_ctx *ctx, int step, const double *rho01,
                        double *out);

/* How the past pair sum is formed.
 * ON_THE_FLY: Green's function and interpolant evaluated per pair per step
 *             (default; FP64 or mixed).
 * TABLE:      the reference's own method -- a per-pair coefficient table
 *             built once (HistoryInteraction::build_coefficient_table,
 *             history_interaction.cpp:21-45) and streamed every step against a
 *             ring of rho01 samples (evaluate, :47-70).  FP64 contexts only;
 *             memory (order + 1) x 16 B + 2 B per ordered pair of each system
 *             (1.4 GB at N = 4096, 44 GB at N = 23170), checked against free
 *             device memory (QUEST_ERR_NOMEM).  Bound by HBM bandwidth. */
#define QUEST_PAIR_ON_THE_FLY 0
#define QUEST_PAIR_TABLE 1
/* TIME_BLOCKED: the same sum as ON_THE_FLY (either precision), reordered in
 *             time.  A pair of integer delay n reads, at step t, steps up to
 *             t - n only, so at every step t0 that is a multiple of K (8 or
 *             16, quest_params.block_steps) one pass over the pairs of delay
 *             >= K forms their sums for the K steps t0 .. t0 + K - 1: the
 *             pair's weights (geometry, Green's function, Lagrange weights)
 *             once, then K windows of the source's rho01 samples.  The pairs
 *             of delay 1 .. 3 are summed step by step from a short list, those
 *             of delay 4 .. 7 from a list in 4-step windows every 4 steps,
 *             with K = 16 those of delay 8 .. 15 from a list in 8-step windows
 *             every 8 steps, the zero-delay pairs as in the other methods.
 *             Memory: a sample ring of n x (the largest delay + order + K, to
 *             a power of two, + K + order) complex values and K slices of the
 *             partial sums.  Both quest_run and quest_hist_evaluate drive it;
 *             quest_run replays graphs of K steps. */
#define QUEST_PAIR_TIME_BLOCKED 2

/* Selects the pair method; call before the first quest_run or
 * quest_hist_evaluate (QUEST_ERR_STATE afterwards).  Replaces the choice the
 * reference makes once and for all in HistoryInteraction's constructor
 * (history_interaction.cpp:5-19 builds the table). */
int quest_set_pair_method(quest_ctx *ctx, int method);

/* --------------------------------------------- PulseInteraction::evaluate */

/* Sets (or, with NULL, clears) the pulse of every system: Pulse
 * (pulse.h:10-24). */
int quest_set_pulse(quest_ctx *ctx, const quest_pulse *pulse);

/* One pulse per system of a batch context (pulses: [n_systems]). */
int quest_set_pulses(quest_ctx *ctx, const quest_pulse *pulses);

/* Several pulses acting on one system (0 <= system < n_systems; 0 <=
 * n_pulses <= QUEST_MAX_PULSES; n_pulses = 0 clears): the pulse field is the
 * sum of the pulses' fields, added in list order, as the reference's
 * integrator sums its list of interactions (integrator.cpp:173-179) when
 * main.cpp's one PulseInteraction (main.cpp:43-46) is replaced by several;
 * README.md:64 speaks of "pulse(s)" while read_pulse_config reads one
 * (pulse.cpp:37-46). */
#define QUEST_MAX_PULSES 8
int quest_set_pulse_list(quest_ctx *ctx, int system, int n_pulses, const quest_pulse *pulses);

/* Replaces PulseInteraction::evaluate(step) (pulse_interaction.cpp:9-20):
 * out[n] (complex, imaginary part 0). */
int quest_pulse_evaluate(quest_ctx *ctx, int step, double *out);

/* ------------------------------------------- PredictorCorrector::Weights */

/* Replaces PredictorCorrector::Weights(n_lambda, n_time, radius)
 * (integrator.cpp:5-102): host FP64 complex minimum-norm least squares
 * (one-sided Jacobi SVD).  ps, cs: [2][n_time]; row 1 and *future_coef carry
 * the step factor (n_time - 1) / 2.  Pure host code; needs no GPU. */
int quest_pc_weights(int n_lambda, int n_time, double radius, double *ps,
                     double *cs, double *future_coef);

/* Sets the weights the device integrator uses (n_time must equal the
 * context's history_window).  Without this call quest_run uses
 * quest_pc_weights(18, window, 3.15) (main.cpp:49). */
int quest_set_pc_weights(quest_ctx *ctx, int n_time, const double *ps,
                         const double *cs, double future_coef);

/* -------------------------------------- PredictorCorrector::Integrator */

/* Replaces Integrator::solve (integrator.cpp:120-127) for n_steps steps
 * starting at the context's current step, with the interactions main.cpp
 * wires (pulse first, then history; main.cpp:43-46) and the ground-state
 * prehistory (main.cpp:32-36).  Everything stays on the device.  On a
 * non-finite state returns QUEST_ERR_DOMAIN and quest_bad_step gives the
 * first bad step (integrator.cpp:187-207). */
int quest_run(quest_ctx *ctx, int n_steps);

/* Allocation-free variant of quest_run.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory once a run is under way
 * (tight serving loops, harnesses that forbid hipMalloc between steps).  It
 * calls no hipMalloc, hipFree or other allocating routine: the same kernels as
 * quest_run are launched one by one (no graph is instantiated), and the pair
 * kernel's partial sums -- (source chunks + 1) x n x 16 B, e.g. 49 x 65536 x
 * 16 B = 51.4 MB for mixed precision at the default split (48 chunks of 1366
 * sources), 25 x 65536 x 16 B = 26.2 MB for FP64 (24 chunks) -- go to a
 * static device array of QUEST_SCRATCH_BYTES that each GPU the library runs
 * on holds until the process ends, instead of the context's buffer.  Results
 * are those of quest_run, bit for bit.  QUEST_ERR_ARG if the partial sums do
 * not fit (at the default split: mixed above ~85k dots, FP64 above ~167k;
 * longer chunks via quest_params.source_chunk shrink them; with
 * QUEST_PAIR_TIME_BLOCKED, whose default split is 8 chunks, there are K = 8
 * or 16 slices of 9 rows: above ~58k dots with 8-step passes, ~29k with
 * 16-step ones, fewer where the split is halved to fill the GPU).  With
 * QUEST_PAIR_TIME_BLOCKED, a call that starts inside a pass whose partial
 * sums the previous quest_run left in the context's buffer forms that pass
 * again in the static array first (and quest_run does the reverse).  Every
 * call on one GPU shares that GPU's array:
 * calls that overlap in time on one GPU (from several host threads) race and
 * must be serialised by the caller; a call has finished with the array when
 * it returns.  The pair kernel is not timed (quest_set_timing). */
#define QUEST_SCRATCH_BYTES (64u << 20)
int quest_run_noalloc(quest_ctx *ctx, int n_steps);

/* Step the next quest_run/quest_hist_evaluate call works on. */
int quest_current_step(const quest_ctx *ctx, int *step);

/* First step with a non-finite state, or -1. */
int quest_bad_step(const quest_ctx *ctx, int *step);

/* State (rho00, rho01) and derivative of every dot at the last finished step:
 * the slot history[i][step][0] and [1] (history.h:13-14).  rho, drho: [n][2]
 * complex; drho may be NULL.  rho00 and its derivative come back with
 * imaginary part 0: from the real ground state the reference's imaginary
 * part stays an exact +-0 (its RHS is -Re z - Im rho00 / T1, z = Omega
 * conj(rho01) - conj(Omega) rho01, whose real part is the difference of two
 * identically rounded products), so the device holds rho00 real. */
int quest_get_state(quest_ctx *ctx, double *rho, double *drho);

/* ------------------------------------------------------ diagnostics */

/* Integer delay floor(|r_obs - r_src| / (c0 dt)) of every unordered pair in
 * the reference's pair order (history_interaction.cpp:25-34, idx2coord
 * :80-86; the pair index as a 64-bit integer), computed on the device by the
 * delay split of the context's pair kernel: the FP64 split (QUEST_PREC_FP64)
 * or the mixed path's FP32 estimate with the FP64 split near integers
 * (QUEST_PREC_MIXED), so the two precisions' classifications can be
 * compared pair by pair.  delays: [n(n-1)/2]. */
int quest_pair_delays(quest_ctx *ctx, int32_t *delays);

/* The same for the pairs [first, first + count) of that order (delays:
 * [count]); QUEST_ERR_ARG outside [0, n(n-1)/2). */
int quest_pair_delays_range(quest_ctx *ctx, int64_t first, int64_t count, int32_t *delays);

/* Enables HIP-event timing of the pair kernel (on the context's stream). */
int quest_set_timing(quest_ctx *ctx, int enable);

/* Accumulated pair-kernel time (ms) and launch count since timing was
 * enabled or last reset; reset != 0 clears them after reading.  With
 * QUEST_PAIR_TIME_BLOCKED a launch is one pass of K steps. */
int quest_pair_kernel_time(quest_ctx *ctx, double *ms, int64_t *launches,
                           int reset);

/* The same for the list kernels of QUEST_PAIR_TIME_BLOCKED (pairs of delay
 * 1 .. 3 every step, 4 .. 7 every 4 steps, 8 .. 15 every 8 steps with
 * 16-step passes), added up; launches counts the steps they served.  Zero
 * for the other methods. */
int quest_short_kernel_time(quest_ctx *ctx, double *ms, int64_t *launches,
                            int reset);

/* Times the pair kernel (`launches` launches each, after a warm-up) and keeps
 * the fastest configuration, in two stages: (1) source-chunk lengths (the
 * default scaled by 1/4 to 4, and the largest system in 8, 16, 24, 32 or 48
 * equal chunks) x a persistent grid of all resident blocks per CU, one fewer
 * and two fewer, with static item claiming and the default source loop; (2)
 * at the three fastest of those, dynamic claiming (unless item_claim fixes
 * it) and, in mixed precision, the other source loop (unless source_loop
 * fixes it).  Returns the chunk length in *chunk_len.  Changes no simulation
 * state.  No counterpart in the reference, which has no tunable kernel. */
int quest_autotune(quest_ctx *ctx, int launches, int *chunk_len);

/* Integer facts about the context.  info: [20] int64 = n_dots, ring slots,
 * delay bound, zero-delay pairs, work items, persistent grid, state slots,
 * precision, systems, source-chunk length, chunks of the largest system,
 * step graph captured (0/1), block size of the one-launch predictor-corrector
 * step (0: one launch per iteration, used when a connected group of
 * zero-delay partners exceeds 256 dots), size of the largest such group,
 * pair-kernel items claimed dynamically (1) or by a static stride (0),
 * pipelined (1) or plain (0) source loop of the mixed-precision pair kernel,
 * pair method (QUEST_PAIR_*), steps per pass of the time-blocked method (0
 * for the others), its listed pairs of delay 1 .. 7, its sample-ring
 * slots. */
int quest_ctx_info(const quest_ctx *ctx, int64_t *info);

/* Blocks until all work queued on the context's stream is done. */
int quest_sync(quest_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif
