// quest_internal.hpp -- internal declarations shared by the C ABI and the
// HIP kernels.  Not part of the public interface (include/quest_amd.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/quest_amd.h"

namespace quest {

constexpr int kMaxOrder = 6;       // interpolation orders 1..6 are instantiated
constexpr int kMaxWindow = 64;     // PC history window limit
constexpr int kBlock = 256;        // threads per block of the per-dot and pair kernels
constexpr int kWave = 64;          // wavefront size; a pair-kernel work item is one wave of targets
// The step kernel counts its finished blocks in kStepGroups counters (block b
// in counter b % kStepGroups), each on its own 128-B line after the first
// line of the step counter array, and the last block of each group counts
// the group at step[3]: blocks that finish together spread their atomics
// over many lines instead of queueing on one.
constexpr int kStepGroups = 64;
constexpr int kCounterStride = 32;  // ints per 128-B line
constexpr int kStepCounterInts = kCounterStride * (1 + kStepGroups);
constexpr int kNumCorrectorSteps = 10;  // integrator.cpp:3
// steps of one pass of the time-blocked pair sum: 8, or 16 where the pairs of
// delay below 16 are few (quest_params.block_steps)
constexpr int kBlockStepsShort = 8;
constexpr int kBlockStepsLong = 16;
constexpr double kLongPassShortFraction = 0.02;
// the pairs of delay below K: 1 .. 3 on the short list, summed at every step;
// 4 .. 7 on a window list summed every 4 steps for 4 steps; with 16-step
// passes 8 .. 15 on a window list summed every 8 steps for 8 steps (a pair of
// delay n >= w has its samples for w steps)
constexpr int kShortListEnd = 4;
constexpr int kWindowLists = 2;
// e^{i theta s} tables: 2^10 entries each for the low and middle ten bits of
// the step, 2^11 for the high bits (steps below 2^31)
constexpr int kPhaseLevel = 1024;
constexpr int kPhaseEntries = 4 * kPhaseLevel;
// static device scratch of quest_run_noalloc (QUEST_SCRATCH_BYTES)
constexpr size_t kScratchBytes = QUEST_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

int pc_weights(int n_lambda, int n_time, double radius, double *ps, double *cs,
               double *future_coef);

// Scalars every kernel of a context shares (all FP64; kernels convert).
struct Consts {
  int n;              // dots
  int order;          // M
  int kstride;        // complex entries per ring slot (M + 1)
  int ring_slots;     // power of two > max delay
  int ring_mask;
  int grid;           // blocks of the persistent pair kernel (a multiple of 8)
  int tiles;          // target tiles of kBlock dots (all systems)
  int state_slots;    // state ring: power of two > max(window, order + 2)
  int state_mask;
  double c, dt, cdt, icdt, omega, theta;   // theta = omega dt
  double k_green;     // mu0 / (4 pi hbar)
  double hbar, laser_omega;
  double two_idt2;    // 2 / dt^2
  // Both pair sums work in units of L = c0 dt (positions relative to the
  // centre of all systems' boxes, divided by L, so a distance is the delay
  // tau) and divide their terms by unit_scale = two_idt2 / L (applied once per
  // target, in FP64); there w^2 dt^2 / 2 becomes oo_l = theta^2 / 2
  double oo_l, unit_scale;
  // FP32 sin/cos of theta x, |x| <= 1/2 (sincos_theta): sin_c[0..1] the
  // Chebyshev-economized u - u^3/6 + u^5/120 over |u| <= theta/2 (two terms),
  // cos_c = -theta^2/2, theta^4/24
  double sin_c[2], cos_c[2];
  // FP64 sin/cos of theta x, |x| <= 1/2, theta <= 0.33 (sincos_centred64):
  // Taylor series in x^2 with theta folded in, highest first: trig64[0..4] =
  // (-1)^k theta^(2k+1) / (2k+1)!, k = 4..0; trig64[5..9] = (-1)^k theta^(2k) / (2k)!,
  // k = 5..1 (the constant 1 is the last addend); truncation below 1e-16 for
  // |theta x| <= 0.165
  double trig64[10];
  // 1: the rotation e^{-i theta x} of the pair term uses those short series
  // (theta small enough for rounding-level accuracy); 0: library sincos
  int trig_poly;
  // error bound of the FP32 delay estimate: eps = eps_rel tau + eps_abs
  double eps_rel, eps_abs;
  // monomial coefficients of the Lagrange basis: vinv[k][m] = [x^k] L_m(x)
  double vinv[(kMaxOrder + 1) * (kMaxOrder + 1)];
  // the same about x = 1/2 (the ring's basis): vinv_c[j][m] = [(x - 1/2)^j] L_m(x)
  double vinv_c[(kMaxOrder + 1) * (kMaxOrder + 1)];
  double half_phase_re, half_phase_im;  // e^{-i theta / 2} (the ring's phase, rotating frame)
  int dyadic;         // QUEST_DYADIC_*
  int n_systems;      // independent dot systems (pairs only within a system)
  int n_work;         // pair-kernel work items (system, target tile, source chunk)
  int dyn_items;      // 1: pair-kernel waves claim items from a device counter (else a static stride)
  int pipelined;      // 1: the FP32 pair loop stages the next source ahead of this one's arithmetic
  int compensated;    // 1: the FP32 pair loop's block sums are Kahan-compensated
  int max_chunks;     // largest number of source chunks of one system (rows of `partial`)
  // coefficient-table mode (the reference's precomputed per-pair coefficients)
  int pair_table;     // 1: the pair sum streams `table` instead of evaluating on the fly
  int samp_slots;     // rho01 sample ring: power of two > max delay + order + 1
  int samp_mask;
  long long table_plane;  // table entries per coefficient plane (sum over systems of n_s^2)
  // time-blocked mode (QUEST_PAIR_TIME_BLOCKED): the pair sums of block_steps
  // steps from one pass over the pairs of delay >= block_steps, a short list
  // for delays 1 .. block_steps - 1, and a rho01 sample ring
  int pair_blocked;   // 1: time-blocked pair sums
  int block_steps;    // K, the steps of one pass (a launch at every step that is a multiple of K): 8 or 16
  int hist_slots;     // sample ring: power of two > max delay + order
  int hist_mask;
  int hist_row;       // entries per source row: hist_slots + K + order - 1, rounded up to even
  int win_steps[kWindowLists];  // window lists: 4 (delays 4 .. 7), and 8 with 16-step passes (8 .. 15), else 0
};

struct PulseParams {
  int on;
  double amplitude, delay, width, freq;
  double k[3], pol[3];
};

// Device buffers of one context (sorted dot order).
struct DeviceState {
  double4 *pos64 = nullptr, *dip64 = nullptr;   // (x, y, z, 0)
  float4 *pd32 = nullptr;              // [n][2]: position (pos - centre) / (c0 dt), dipole (FP32)
  double4 *pd64 = nullptr;             // the same in FP64 (FP64 contexts)
  double *freq = nullptr, *t1 = nullptr, *t2 = nullptr;
  void *ring = nullptr;                // coefficient planes, kstride x [n][ring_slots] complex
  // rho00 is held real: from the real ground state its imaginary part stays an
  // exact +-0 (the RHS's imaginary part is -Re(om conj(r1) - conj(om) r1) -
  // Im(r00) / T1, whose first term is the difference of two identically
  // rounded products), so its ring, iterate and corrector base are doubles
  double *y0 = nullptr;                // state ring [state_slots][n]: rho00
  double2 *y1 = nullptr;               // state ring [state_slots][n]: rho01
  double *d0 = nullptr;                // derivative ring of rho00
  double2 *d1 = nullptr;               // derivative ring of rho01
  double *cur0 = nullptr;              // iterate double buffer [2][n] (rho00)
  double2 *cur1 = nullptr;             // iterate double buffer [2][n] (rho01)
  double *yc0 = nullptr;               // corrector base [n] (rho00)
  double2 *yc1 = nullptr;              // corrector base [n] (rho01)
  double2 *field = nullptr;            // fixed part of the field at this step [n]
  double2 *partial = nullptr;          // [max_chunks][n]
  double2 *phase = nullptr;            // [kPhaseEntries]: e^{i theta l}, l < 2^10; e^{i theta 2^10 m},
                                       // m < 2^10; e^{i theta 2^20 h}, h < 2^11 (phase_at)
  double2 *out = nullptr;              // [n] staging of per-call results
  double2 *io = nullptr;               // [2n] per-call mode: rho01 in caller order, then results in it
  int *near_start = nullptr;           // [n + 1]
  int *near_src = nullptr;             // [near]
  int *near_tgt = nullptr;             // [near]
  int *near_n = nullptr;               // [near] FP64 integer delay
  double2 *near_coef = nullptr;        // [near][order + 1]
  double *pcw = nullptr;               // ps[2W] then cs[2W]
  int *bad_step = nullptr;             // first non-finite step (INT_MAX = none)
  int *counts = nullptr;               // scratch [n + 1]
  int *perm_dev = nullptr;             // inverse permutation (original -> sorted)
  int *sys_of = nullptr;               // [n] system of each dot
  int2 *sys_range = nullptr;           // [systems] dot range [begin, end)
  int *sys_chunks = nullptr;           // [systems] source chunks (partial rows used)
  PulseParams *pulses = nullptr;       // [systems][QUEST_MAX_PULSES], used ones first
  int4 *work = nullptr;                // [n_work] (target begin, target end, source begin, source end)
  int *work_row = nullptr;             // [n_work] partial row written by the item
  int *step = nullptr;                 // [kStepCounterInts]: the step being integrated (graph replays
                                       // read it), the pair kernels' claim counter ([1]), the step
                                       // kernel's finished groups ([3]), then its finished blocks per
                                       // group, one counter per 128-B line (all counters zero between
                                       // launches)
  // coefficient-table mode only
  double2 *samp = nullptr;             // rho01 sample ring [n][samp_slots] (finished steps)
  double2 *table = nullptr;            // [order + 1][table_plane]: system s, source b, target a at
                                       // sys_toff[s] + (b - begin_s) n_s + (a - begin_s)
  int16_t *table_n = nullptr;          // [table_plane] integer delay (0 for excluded pairs)
  long long *sys_toff = nullptr;       // [systems] first table entry of each system
  // time-blocked mode only
  void *hist = nullptr;                // rho01 sample ring [n][hist_row] (double2 FP64, float2 mixed)
  void *rot = nullptr;                 // [hist_slots] e^{-i theta (n + 1/2)} (double2 FP64, float2 mixed)
  int *short_start = nullptr;          // [n + 1] short list (delays 1 .. 3), CSR by target
  int *short_src = nullptr;            // [short pairs]
  double2 *short_sum = nullptr;        // [n] the short list's sum at the current step
  int *win_start[kWindowLists] = {};   // [n + 1] window list l (k.win_steps[l] != 0), CSR by target
  int *win_src[kWindowLists] = {};     // [its pairs]
  double2 *win_sum[kWindowLists] = {}; // [win_steps[l]][n] its sums for the steps from its last launch
  // one-launch predictor-corrector step: thread -> dot (-1 idle), each group of
  // zero-delay partners within one block, all groups in the leading blocks
  int *step_map = nullptr;
  // [near]: thread of the near pair's source within the block of the one-launch
  // step (groups of partners lie whole in one block), for the LDS exchange
  uint8_t *near_lane = nullptr;
};

// Device address of the static scratch array on `device`, which must be the
// current device; looked up once per device (hipGetSymbolAddress) and cached.
hipError_t scratch_address(int device, double2 **ptr);

// ---------------------------------------------------------------- launchers
// All launch on `s` and return the hipError_t of the launch.
hipError_t launch_fill_ground_state(const Consts &k, DeviceState &d, hipStream_t s);
hipError_t launch_clear_ring(const Consts &k, DeviceState &d, size_t ring_bytes, hipStream_t s);
hipError_t launch_clear(void *p, size_t bytes, hipStream_t s);  // bytes: a multiple of 16
// Coefficient table of one system (dots [range.x, range.y)), entries from toff.
hipError_t launch_table_build(const Consts &k, DeviceState &d, int2 range, long long toff, hipStream_t s);
// Kernels of the step read the step number from d.step (so that one captured
// graph serves every step); launch_set_step / launch_advance_step write it.
hipError_t launch_set_step(DeviceState &d, int step, hipStream_t s);
hipError_t launch_advance_step(DeviceState &d, hipStream_t s);
hipError_t launch_pair_past(const Consts &k, DeviceState &d, int precision, hipStream_t s);
hipError_t launch_pc_begin(const Consts &k, DeviceState &d, int window, hipStream_t s);
// The whole step of every dot (predictor + 11 iterations) in one launch of
// `blocks` blocks of `block` (<= kBlock) threads over d.step_map; the first
// `sync_blocks` blocks hold every dot with zero-delay partners.  Advances
// the device step counter at its end.
hipError_t launch_pc_step(const Consts &k, DeviceState &d, int precision, double future_coef, int window,
                          int block, int blocks, int sync_blocks, hipStream_t s);
hipError_t launch_pc_iter(const Consts &k, DeviceState &d, int precision, double future_coef,
                          int iter, hipStream_t s);
hipError_t launch_hist_commit(const Consts &k, DeviceState &d, int precision, int step, hipStream_t s);
hipError_t launch_hist_reduce(const Consts &k, DeviceState &d, hipStream_t s);
hipError_t launch_hist_near(const Consts &k, DeviceState &d, hipStream_t s);
// Per-call mode: caller order <-> sorted order on the device (d.io, d.perm_dev):
// cur1[sorted] = io[caller]; io[n + caller] = out[sorted].
hipError_t launch_permute_in(const Consts &k, DeviceState &d, hipStream_t s);
hipError_t launch_permute_out(const Consts &k, DeviceState &d, hipStream_t s);
hipError_t launch_pulse(const Consts &k, DeviceState &d, hipStream_t s);
// Occupancy of the pair kernel (blocks per CU) for sizing the persistent grid.
int pair_blocks_per_cu(const Consts &k, int precision);
// Lists of the sources of each target whose pair delay (the precision's own
// split) lies in [lo, hi), b != a: counts[a] (FILL false) or, with start
// given, src[start[a] ..] in source order and tgt (may be null) (FILL true).
// The zero-delay list is [0, 1); the time-blocked short list [1, K).
hipError_t launch_near_count(const Consts &k, DeviceState &d, int precision, hipStream_t s);
hipError_t launch_near_fill(const Consts &k, DeviceState &d, int precision, hipStream_t s);
hipError_t launch_delay_list(const Consts &k, DeviceState &d, int precision, int lo, int hi, bool fill,
                             const int *start, int *src, int *tgt, int *counts, hipStream_t s);
// Time-blocked mode: the short list's sum at the current step into d.short_sum
// (list < 0), or window list `list`'s sums for the k.win_steps[list] steps
// from the current one into d.win_sum[list].
hipError_t launch_pair_list(const Consts &k, DeviceState &d, int precision, int list, hipStream_t s);
hipError_t launch_near_coef(const Consts &k, DeviceState &d, int near_total, hipStream_t s);
hipError_t launch_pair_delays(const Consts &k, DeviceState &d, int precision, int32_t *out, int64_t first,
                              int64_t count, hipStream_t s);

}  // namespace quest
