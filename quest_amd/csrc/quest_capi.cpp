// quest_capi.cpp -- the C ABI (include/quest_amd.h) over the gfx950 kernels.
//
// A context owns the device copy of one or more independent dot systems
// (each in Morton order, concatenated), the interpolant ring of the pair sum,
// the predictor-corrector state ring and one HIP stream.  Two ways to drive
// it, one per context:
//   per-call  quest_hist_evaluate   -- HistoryInteraction::evaluate replacement
//   whole-run quest_run             -- Integrator::solve replacement, all steps
//                                      on the device, one captured graph per step
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "quest_internal.hpp"

using namespace quest;

struct quest_ctx {
  quest_params P{};
  int n = 0;
  int n_sys = 0;
  std::vector<int> sys_begin;  // [n_sys + 1] dot offsets of the systems
  std::vector<int> perm;       // perm[sorted] = caller index
  std::vector<int> inv;        // inv[caller] = sorted index
  Consts k{};
  DeviceState d{};
  hipStream_t stream = nullptr;
  size_t ring_bytes = 0;
  int dmax_bound = 0;
  int near_total = 0;
  int mode = 0;           // 0 unset, 1 per-call, 2 whole-run
  int cur_step = 0;       // whole-run: next step; per-call: step of the current values (-1 = none)
  int bad = -1;
  std::vector<double> pcw;  // ps[2W], cs[2W]
  double fc = 0;
  std::vector<PulseParams> pulses;
  std::string err;
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double pair_ms = 0;
  int64_t pair_launches = 0;
  double2 *h_stage = nullptr;  // pinned [2n]
  int chunk_len = 0;           // sources per pair-kernel work item
  int partial_rows = 0;        // rows allocated in d.partial
  int cu_count = 0;
  int blocks_per_cu = 1;
  int pair_method = QUEST_PAIR_ON_THE_FLY;
  int step_block = 0;          // block size of the fused step kernel (0: per-iteration kernels)
  int step_blocks = 0;         // its grid
  int step_sync_blocks = 0;    // leading blocks holding the zero-delay groups
  int max_group = 0;           // largest connected group of zero-delay partners
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double2 *scratch = nullptr;  // this device's static scratch array (quest_run_noalloc)
  int short_total = 0;         // time-blocked mode: pairs of delay 1 .. 7 (short list and the 4-step window list)
  int slice_buf = 0;           // time-blocked mode: the current block's partial sums are in the
                               // context's buffer (0) or the static scratch (1)
  size_t hist_bytes = 0;       // time-blocked mode: the sample ring
  std::vector<int> ev_kind;    // per timed event pair: 0 pair kernel, 1 short-list kernel
  double short_ms = 0;
  int64_t short_launches = 0;
};

namespace {

thread_local std::string g_create_error;

int fail(quest_ctx *c, int code, const std::string &msg)
{
  if(c) c->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPC(ctx, call)                                                                 \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if(e_ != hipSuccess)                                                                \
      return fail((ctx), QUEST_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while(0)

int next_pow2(int v)
{
  int p = 1;
  while(p < v) p <<= 1;
  return p;
}

// Morton (Z-order) permutation of dots [begin, end): 21 bits per axis over the
// system's bounding box, stable for equal keys.  Returns caller indices.
std::vector<int> morton_order(int begin, int end, const double *pos)
{
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for(int i = begin; i < end; ++i)
    for(int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], pos[3 * i + a]);
      hi[a] = std::max(hi[a], pos[3 * i + a]);
    }
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  const int n = end - begin;
  std::vector<uint64_t> key(n);
  for(int i = 0; i < n; ++i) {
    uint64_t q[3];
    for(int a = 0; a < 3; ++a) {
      const double span = hi[a] - lo[a];
      const double t = span > 0 ? (pos[3 * (begin + i) + a] - lo[a]) / span : 0.0;
      q[a] = static_cast<uint64_t>(std::min(std::max(t, 0.0), 1.0) * 2097151.0);
    }
    key[i] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  for(int &o : order) o += begin;
  return order;
}

// vinv[k][m] = coefficient of x^k in the Lagrange basis polynomial L_m(x) on
// the nodes 0..M (lagrange_set.cpp:20-29 expanded in monomials).
void lagrange_monomials(int M, double *vinv)
{
  for(int i = 0; i < (kMaxOrder + 1) * (kMaxOrder + 1); ++i) vinv[i] = 0;
  for(int m = 0; m <= M; ++m) {
    std::vector<double> poly(1, 1.0);  // coefficients, ascending powers
    double den = 1;
    for(int j = 0; j <= M; ++j) {
      if(j == m) continue;
      std::vector<double> next(poly.size() + 1, 0.0);
      for(size_t p = 0; p < poly.size(); ++p) {
        next[p + 1] += poly[p];
        next[p] -= j * poly[p];
      }
      poly.swap(next);
      den *= (m - j);
    }
    for(int kk = 0; kk <= M; ++kk) vinv[kk * (kMaxOrder + 1) + m] = poly[kk] / den;
  }
}

// vinv_c[j][m] = sum_{k >= j} C(k, j) (1/2)^(k - j) vinv[k][m]: the Lagrange
// basis as polynomials in (x - 1/2).
void centre_basis(int M, const double *vinv, double *vinv_c)
{
  for(int i = 0; i < (kMaxOrder + 1) * (kMaxOrder + 1); ++i) vinv_c[i] = 0;
  for(int j = 0; j <= M; ++j)
    for(int m = 0; m <= M; ++m) {
      double acc = 0;
      for(int kk = j; kk <= M; ++kk) {
        double binom = 1;
        for(int t = 0; t < j; ++t) binom = binom * (kk - t) / (t + 1);
        acc += binom * std::ldexp(1.0, -(kk - j)) * vinv[kk * (kMaxOrder + 1) + m];
      }
      vinv_c[j * (kMaxOrder + 1) + m] = acc;
    }
}

void drop_graph(quest_ctx *c)
{
  if(c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if(c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
}

// Work items (one per wave) of the pair kernel for a given chunk length.
int count_work(const quest_ctx *c, int chunk_len)
{
  int64_t items = 0;
  for(int s = 0; s < c->n_sys; ++s) {
    const int ns = c->sys_begin[s + 1] - c->sys_begin[s];
    if(ns < 2) continue;
    items += static_cast<int64_t>((ns + kWave - 1) / kWave) * ((ns + chunk_len - 1) / chunk_len);
  }
  return static_cast<int>(std::min<int64_t>(items, INT_MAX));
}

// Work list of the pair kernel for a source-chunk length.  Items are one wave
// of targets (kWave dots of one system) against one chunk of chunk_len of the
// system's sources; they come in groups of up to four wave tiles with the same
// chunk (the waves of one block; for systems of >= 256 dots a 256-target
// tile).  The partial row an item writes is its chunk's index.
//   Dynamic claiming: tile groups in order, the chunks of a tile group
// consecutive; wave w of the grid takes item w first, then the waves claim
// the items after the first grid's worth in list order.
//   Static stride (block i's waves take items 4i..4i+3, then 4(i + grid)..,
// and block i runs on XCD i mod 8 under round-robin dispatch): the unit of
// placement is a (system, chunk) pair -- all tile groups of the system against
// that chunk's sources.  Units go to XCDs longest first, each to the XCD with
// the least work so far (for one system: chunk q to XCD q mod 8), and each
// XCD's blocks run its units one after another, chunk-major, so that all of
// an XCD's resident blocks sweep the same source chunk at about the same time
// and read its ring rows from that XCD's L2 together.  Unused slots hold
// empty items (one target, dot 0, no sources) whose partial row is a spare row
// past the real ones.  (A batch spread over all XCDs instead hit L2 21 % of the
// time and fetched 3.4 GB per launch.)
int build_work(quest_ctx *c, int chunk_len)
{
  constexpr int kPer = kBlock / kWave;
  constexpr int kXcds = 8;
  std::vector<int4> work;
  std::vector<int> row;
  std::vector<int> chunks(c->n_sys, 0);
  int max_chunks = 1;
  for(int s = 0; s < c->n_sys; ++s) {
    const int ns = c->sys_begin[s + 1] - c->sys_begin[s];
    if(ns < 2) continue;
    chunks[s] = (ns + chunk_len - 1) / chunk_len;
    max_chunks = std::max(max_chunks, chunks[s]);
  }
  // items of tile group g, chunk q of system s
  auto emit = [&](std::vector<int4> &w, std::vector<int> &r, int s, int g, int q) {
    const int b = c->sys_begin[s], e = c->sys_begin[s + 1], wtiles = (e - b + kWave - 1) / kWave;
    for(int t = g; t < std::min(wtiles, g + kPer); ++t) {
      w.push_back(make_int4(b + t * kWave, std::min(e, b + (t + 1) * kWave), b + q * chunk_len,
                            std::min(e, b + (q + 1) * chunk_len)));
      r.push_back(q);
    }
  };
  const int resident = std::max(kXcds, c->cu_count * c->blocks_per_cu / kXcds * kXcds);
  int grid = 0;
  if(c->k.dyn_items) {
    for(int s = 0; s < c->n_sys; ++s) {
      const int ns = c->sys_begin[s + 1] - c->sys_begin[s];
      if(ns < 2) continue;
      const int wtiles = (ns + kWave - 1) / kWave;
      for(int g = 0; g < wtiles; g += kPer)
        for(int q = 0; q < chunks[s]; ++q) emit(work, row, s, g, q);
    }
    // persistent grid: one full wave of resident blocks (a multiple of 8), at
    // most the blocks the items need, rounded up to 8
    const int blocks8 = ((static_cast<int>(work.size()) + kPer - 1) / kPer + kXcds - 1) / kXcds * kXcds;
    grid = std::max(kXcds, std::min(resident, blocks8));
  } else {
    // (system, chunk) units to XCDs, longest first (pair work ~ wave tiles x sources)
    struct Unit {
      int s, q;
      long long cost;
    };
    std::vector<Unit> units;
    int n_groups = 0;
    for(int s = 0; s < c->n_sys; ++s) {
      const int b = c->sys_begin[s], e = c->sys_begin[s + 1], ns = e - b;
      if(ns < 2) continue;
      const int wtiles = (ns + kWave - 1) / kWave;
      n_groups += (wtiles + kPer - 1) / kPer * chunks[s];
      for(int q = 0; q < chunks[s]; ++q)
        units.push_back({s, q, static_cast<long long>(wtiles) * (std::min(e, b + (q + 1) * chunk_len) - (b + q * chunk_len))});
    }
    std::stable_sort(units.begin(), units.end(), [](const Unit &x, const Unit &y) { return x.cost > y.cost; });
    long long load[kXcds] = {0};
    std::vector<std::vector<Unit>> on_xcd(kXcds);
    for(const Unit &u : units) {
      const int x = static_cast<int>(std::min_element(load, load + kXcds) - load);
      on_xcd[x].push_back(u);
      load[x] += u.cost;
    }
    const int blocks8 = (n_groups + kXcds - 1) / kXcds * kXcds;
    grid = std::max(kXcds, std::min(resident, blocks8));
    const int per_xcd = grid / kXcds;  // blocks of one XCD: x, x + 8, x + 16, ...
    // each XCD's tile groups, unit after unit (chunk-major)
    struct Group {
      int s, g, q;  // system, first wave tile, chunk
    };
    std::vector<std::vector<Group>> seq(kXcds);
    size_t rounds = 0;
    for(int x = 0; x < kXcds; ++x) {
      for(const Unit &u : on_xcd[x]) {
        const int wtiles = (c->sys_begin[u.s + 1] - c->sys_begin[u.s] + kWave - 1) / kWave;
        for(int g = 0; g < wtiles; g += kPer) seq[x].push_back({u.s, g, u.q});
      }
      rounds = std::max(rounds, (seq[x].size() + per_xcd - 1) / per_xcd);
    }
    if(static_cast<size_t>(kPer) * grid * rounds > static_cast<size_t>(INT_MAX))
      return fail(c, QUEST_ERR_ARG, "pair-kernel work list too long");
    const size_t slots = static_cast<size_t>(kPer) * grid * rounds;
    work.assign(slots, make_int4(0, 1, 0, 0));  // empty item
    row.assign(slots, max_chunks);              // the spare row
    std::vector<int4> w;
    std::vector<int> r;
    for(int x = 0; x < kXcds; ++x)
      for(size_t k = 0; k < seq[x].size(); ++k) {
        const size_t blk = x + static_cast<size_t>(kXcds) * (k % per_xcd), rnd = k / per_xcd;
        const size_t base = static_cast<size_t>(kPer) * (blk + static_cast<size_t>(grid) * rnd);
        w.clear();
        r.clear();
        emit(w, r, seq[x][k].s, seq[x][k].g, seq[x][k].q);
        for(size_t j = 0; j < w.size(); ++j) {
          work[base + j] = w[j];
          row[base + j] = r[j];
        }
      }
  }
  DeviceState &d = c->d;
  drop_graph(c);  // the graph's pair kernel holds the old list
  if(d.work) HIPC(c, hipFree(d.work));
  if(d.work_row) HIPC(c, hipFree(d.work_row));
  d.work = nullptr;
  d.work_row = nullptr;
  const size_t nw = std::max<size_t>(work.size(), 1);
  HIPC(c, hipMalloc(&d.work, sizeof(int4) * nw));
  HIPC(c, hipMalloc(&d.work_row, sizeof(int) * nw));
  if(!work.empty()) {
    HIPC(c, hipMemcpyAsync(d.work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipMemcpyAsync(d.work_row, row.data(), sizeof(int) * row.size(), hipMemcpyHostToDevice, c->stream));
  }
  HIPC(c, hipMemcpyAsync(d.sys_chunks, chunks.data(), sizeof(int) * c->n_sys, hipMemcpyHostToDevice, c->stream));
  // + the spare row of empty items; the time-blocked method keeps one slice
  // of rows per step of its pass
  const int rows = (max_chunks + 1) * (c->k.pair_blocked ? c->k.block_steps : 1);
  if(rows > c->partial_rows) {
    if(d.partial) HIPC(c, hipFree(d.partial));
    d.partial = nullptr;
    HIPC(c, hipMalloc(&d.partial, sizeof(double2) * static_cast<size_t>(c->n) * rows));
    c->partial_rows = rows;
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->chunk_len = chunk_len;
  c->k.n_work = static_cast<int>(work.size());
  c->k.max_chunks = max_chunks;
  c->k.grid = grid;
  return QUEST_OK;
}

// Default chunk length: the largest system cut into 48 (mixed) or 24 (FP64)
// equal source ranges -- a multiple of the 8 XCDs, so that the static list
// gives every XCD the same share of a large system -- of at least 128 (mixed)
// or 256 (FP64) sources, halved while the work list has fewer items than
// resident waves.  Measured on MI355X at N = 16384-131072 against 8 and 16
// chunks: mixed 1-11 % faster at N >= 65536 and equal below; FP64 2-15 %
// faster except at N = 32768 and 131072 (7 % slower there; quest_autotune
// finds those).  A chunk count that is not a multiple of 8 (N = 49152 cut
// into 4096-source chunks: 12) left half the XCDs with half the work: 1.5x.
// The time-blocked pass takes 8 chunks in both precisions: each chunk adds a
// row of partial sums per step, written by the pass and read back by the step
// kernel, and against 16 and 48 chunks (N = 16384-131072, both precisions,
// one box) 8 was the best on average (N = 65536: 0.9 % faster mixed, 2.3 %
// FP64, than 48; the worst case N = 131072 FP64 2.7 % slower than 48).
int default_chunk(const quest_ctx *c)
{
  int nmax = 2;
  for(int s = 0; s < c->n_sys; ++s) nmax = std::max(nmax, c->sys_begin[s + 1] - c->sys_begin[s]);
  const bool mixed = c->P.precision == QUEST_PREC_MIXED;
  const bool blocked = c->pair_method == QUEST_PAIR_TIME_BLOCKED;
  const int count = blocked ? 8 : mixed ? 48 : 24, min_len = mixed ? 128 : 256;
  int len = std::max(min_len, (nmax + count - 1) / count);
  const int resident = c->cu_count * c->blocks_per_cu * (kBlock / kWave);  // waves
  while(len > 64 && count_work(c, len) < resident) len = std::max(64, (len + 1) / 2);
  return len;
}

// Persistent grid of the pair kernels: the resident blocks per CU, one fewer
// for the on-the-fly kernels (measured at N = 32768-131072: the FP64 kernel's
// gathers and the pipelined FP32 loop share each CU's L1 with fewer waves and
// run up to 11 % faster; quest_autotune also times the full grid and two
// fewer).
int default_blocks_per_cu(const quest_ctx *c)
{
  const int full = pair_blocks_per_cu(c->k, c->P.precision);
  return c->pair_method == QUEST_PAIR_ON_THE_FLY && full >= 4 ? full - 1 : full;  // blocked, table: full
}

// The three factor tables of e^{i theta s} (phase_at on the device): each
// entry from a long-double argument reduction, rounded once to FP64.  Fixed
// size, uploaded once at create: no later call allocates for the phase.
int upload_phase(quest_ctx *c)
{
  std::vector<double2> h(kPhaseEntries);
  const long double theta = static_cast<long double>(c->k.omega) * static_cast<long double>(c->k.dt);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  auto put = [&](int at, long double mult) {
    const long double a = fmodl(theta * mult, two_pi);
    h[at] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  };
  for(int j = 0; j < kPhaseLevel; ++j) {
    put(j, static_cast<long double>(j));
    put(kPhaseLevel + j, static_cast<long double>(j) * kPhaseLevel);
  }
  for(int j = 0; j < 2 * kPhaseLevel; ++j)
    put(2 * kPhaseLevel + j, static_cast<long double>(j) * kPhaseLevel * kPhaseLevel);
  HIPC(c, hipMemcpyAsync(c->d.phase, h.data(), sizeof(double2) * kPhaseEntries, hipMemcpyHostToDevice,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return QUEST_OK;
}

int start_state(quest_ctx *c)
{
  HIPC(c, launch_fill_ground_state(c->k, c->d, c->stream));
  HIPC(c, launch_clear_ring(c->k, c->d, c->ring_bytes, c->stream));
  if(c->d.hist) HIPC(c, launch_clear(c->d.hist, c->hist_bytes, c->stream));  // rho01 = 0 before step 0
  c->bad = -1;
  return QUEST_OK;
}

// The predictor-corrector weights on the device: main.cpp:49's (18, window,
// 3.15) from create on, or the caller's from quest_set_pc_weights.
int upload_pc(quest_ctx *c)
{
  const int W = c->P.history_window;
  HIPC(c, hipMemcpyAsync(c->d.pcw, c->pcw.data(), sizeof(double) * 4 * W, hipMemcpyHostToDevice,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  drop_graph(c);  // future_coef is a kernel argument of the graph
  return QUEST_OK;
}

int upload_pulses(quest_ctx *c)
{
  HIPC(c, hipMemcpyAsync(c->d.pulses, c->pulses.data(), sizeof(PulseParams) * c->pulses.size(),
                         hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return QUEST_OK;
}

hipEvent_t pool_event(quest_ctx *c)
{
  if(c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if(hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

int collect_timing(quest_ctx *c)
{
  for(size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float ms = 0;
    HIPC(c, hipEventElapsedTime(&ms, c->ev_pool[i], c->ev_pool[i + 1]));
    if(i / 2 < c->ev_kind.size() && c->ev_kind[i / 2] == 1) {
      c->short_ms += ms;
      c->short_launches += 1;
    } else {
      c->pair_ms += ms;
      c->pair_launches += 1;
    }
  }
  c->ev_used = 0;
  c->ev_kind.clear();
  return QUEST_OK;
}

// The pair kernel, bracketed by HIP events on the context's stream when
// timing (`timed`).  `d` is the context's device state, or a copy of it whose
// partial sums live elsewhere (quest_run_noalloc).
int pair_sum(quest_ctx *c, DeviceState &d, bool timed)
{
  if(c->k.n_work == 0) return QUEST_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timed = timed && c->timing;
  if(timed) {
    e0 = pool_event(c);
    e1 = pool_event(c);
    if(!e0 || !e1) return fail(c, QUEST_ERR_HIP, "hipEventCreate failed");
    c->ev_kind.push_back(0);
    HIPC(c, hipEventRecord(e0, c->stream));
  }
  HIPC(c, launch_pair_past(c->k, d, c->P.precision, c->stream));
  if(timed) HIPC(c, hipEventRecord(e1, c->stream));
  return QUEST_OK;
}
int pair_sum(quest_ctx *c) { return pair_sum(c, c->d, true); }

// Time-blocked method: the list sums of step `step` -- each window list's
// for its w steps at the multiples of w (w = 8 with 16-step passes, and 4),
// the short list's at every step -- bracketed by HIP events when timing.
int list_sums(quest_ctx *c, DeviceState &d, bool timed, int step)
{
  hipEvent_t e0 = nullptr, e1 = nullptr;
  timed = timed && c->timing;
  if(timed) {
    e0 = pool_event(c);
    e1 = pool_event(c);
    if(!e0 || !e1) return fail(c, QUEST_ERR_HIP, "hipEventCreate failed");
    c->ev_kind.push_back(1);
    HIPC(c, hipEventRecord(e0, c->stream));
  }
  for(int l = kWindowLists - 1; l >= 0; --l)
    if(c->k.win_steps[l] && step % c->k.win_steps[l] == 0)
      HIPC(c, launch_pair_list(c->k, d, c->P.precision, l, c->stream));
  HIPC(c, launch_pair_list(c->k, d, c->P.precision, -1, c->stream));
  if(timed) HIPC(c, hipEventRecord(e1, c->stream));
  return QUEST_OK;
}

// One whole step: pair sum, predictor, 10 x (evaluate, correct), final
// evaluation and commit, step counter advance.  `step` is the step being
// queued: the time-blocked method forms its pair sums at the steps that are
// multiples of K and the short list's sum at every step.
int enqueue_step(quest_ctx *c, DeviceState &d, bool timed, int step)
{
  int rc = QUEST_OK;
  if(!c->k.pair_blocked || step % c->k.block_steps == 0) rc = pair_sum(c, d, timed);
  if(rc == QUEST_OK && c->k.pair_blocked) rc = list_sums(c, d, timed, step);
  if(rc != QUEST_OK) return rc;
  if(c->step_block > 0) {
    HIPC(c, launch_pc_step(c->k, d, c->P.precision, c->fc, c->P.history_window, c->step_block,
                           c->step_blocks, c->step_sync_blocks, c->stream));
    return QUEST_OK;  // the step kernel advanced the step counter
  }
  HIPC(c, launch_pc_begin(c->k, d, c->P.history_window, c->stream));
  for(int it = 0; it <= kNumCorrectorSteps; ++it)
    HIPC(c, launch_pc_iter(c->k, d, c->P.precision, c->fc, it, c->stream));
  HIPC(c, launch_advance_step(d, c->stream));
  return QUEST_OK;
}

// Steps captured per graph: one (replayed for every step; the kernels read
// the step number from the device counter), or K for the time-blocked method
// (replayed from steps that are multiples of K).
int graph_steps(const quest_ctx *c) { return c->k.pair_blocked ? c->k.block_steps : 1; }

// The graph of graph_steps(c) steps (kernel launches only; no memsets).
int ensure_graph(quest_ctx *c)
{
  if(c->graph_exec) return QUEST_OK;
  HIPC(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const bool was_timing = c->timing;
  c->timing = false;
  int rc = QUEST_OK;
  for(int i = 0; i < graph_steps(c) && rc == QUEST_OK; ++i) rc = enqueue_step(c, c->d, false, i);
  c->timing = was_timing;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if(rc != QUEST_OK) {
    if(g) (void)hipGraphDestroy(g);
    return rc;
  }
  if(e != hipSuccess) return fail(c, QUEST_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  c->graph = g;
  HIPC(c, hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  return QUEST_OK;
}

void free_device(quest_ctx *c)
{
  drop_graph(c);
  DeviceState &d = c->d;
  void *ptrs[] = {d.pos64, d.dip64, d.pd32, d.pd64, d.freq, d.t1, d.t2, d.ring, d.y0, d.y1,
                  d.d0, d.d1, d.cur0, d.cur1, d.yc0, d.yc1, d.field, d.partial, d.phase, d.out,
                  d.near_start, d.near_src, d.near_tgt, d.near_n, d.near_coef, d.pcw,
                  d.bad_step, d.counts, d.perm_dev, d.io, d.sys_of, d.sys_range, d.sys_chunks,
                  d.pulses, d.work, d.work_row, d.step, d.samp, d.table, d.table_n, d.sys_toff,
                  d.step_map, d.near_lane, d.hist, d.rot, d.short_start, d.short_src, d.short_sum};
  for(void *p : ptrs)
    if(p) (void)hipFree(p);
  for(int l = 0; l < kWindowLists; ++l)
    for(void *p : {static_cast<void *>(d.win_start[l]), static_cast<void *>(d.win_src[l]),
                   static_cast<void *>(d.win_sum[l])})
      if(p) (void)hipFree(p);
  d = DeviceState{};
  if(c->h_stage) (void)hipHostFree(c->h_stage);
  c->h_stage = nullptr;
  for(hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  c->ev_pool.clear();
  if(c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

// Packs the dots into blocks of the one-launch step kernel so that every
// connected group of zero-delay partners (the only coupling inside a step)
// lies in one block.  Groups of two or more come first, whole, in order of
// their first dot, members ascending; a group that does not fit the rest of a
// block starts the next one.  Those blocks (step_sync_blocks) synchronise
// between iterations.  The dots without partners follow in ascending order
// (so the map stays close to the identity and the per-dot loads coalesce).
// The block is the smallest of 64, 128, 256 threads that holds the largest
// group; a larger group leaves the per-iteration kernels in charge
// (step_block = 0).
int pack_step_blocks(quest_ctx *c, const std::vector<int> &start, const std::vector<int> &near_src)
{
  const int n = c->n;
  std::vector<int> root(n);
  for(int i = 0; i < n; ++i) root[i] = i;
  auto find = [&root](int x) {
    while(root[x] != x) {
      root[x] = root[root[x]];
      x = root[x];
    }
    return x;
  };
  for(int a = 0; a < n; ++a)
    for(int e = start[a]; e < start[a + 1]; ++e) {
      const int b = near_src[e];
      if(b < 0 || b >= n) return fail(c, QUEST_ERR_HIP, "zero-delay list holds an invalid source index");
      const int ra = find(a), rb = find(b);
      if(ra != rb) root[std::max(ra, rb)] = std::min(ra, rb);  // the group's root is its first dot
    }
  std::vector<int> size(n, 0), first(n + 1, 0), members(n);
  for(int i = 0; i < n; ++i) ++size[find(i)];
  for(int i = 0; i < n; ++i) first[i + 1] = first[i] + size[i];
  std::vector<int> fill(first.begin(), first.end() - 1);
  for(int i = 0; i < n; ++i) members[fill[find(i)]++] = i;
  c->max_group = n > 0 ? *std::max_element(size.begin(), size.end()) : 0;
  c->step_block = 0;
  c->step_blocks = 0;
  int block = 0;
  for(int b : {64, 128, kBlock})
    if(c->max_group <= b) {
      block = b;
      break;
    }
  if(block == 0 || n == 0) return QUEST_OK;
  std::vector<int> map;
  map.reserve(static_cast<size_t>(n) + n / 4 + block);
  for(int r = 0; r < n; ++r) {
    if(size[r] < 2) continue;
    const size_t used = map.size() % block;
    if(used != 0 && used + size[r] > static_cast<size_t>(block)) map.resize(map.size() + block - used, -1);
    map.insert(map.end(), members.begin() + first[r], members.begin() + first[r] + size[r]);
  }
  const int sync_blocks = static_cast<int>((map.size() + block - 1) / block);
  for(int r = 0; r < n; ++r)
    if(size[r] == 1) map.push_back(r);
  if(map.size() % block) map.resize(map.size() + block - map.size() % block, -1);
  const size_t blocks = map.size() / block;
  // each near pair's source as a thread of its target's block
  std::vector<int> slot_of(n, -1);
  for(size_t t = 0; t < map.size(); ++t)
    if(map[t] >= 0) slot_of[map[t]] = static_cast<int>(t);
  std::vector<uint8_t> lane(near_src.size());
  for(int a = 0; a < n; ++a)
    for(int e = start[a]; e < start[a + 1]; ++e) {
      const int ta = slot_of[a], tb = slot_of[near_src[e]];
      if(ta / block != tb / block) return fail(c, QUEST_ERR_HIP, "zero-delay partners split over two blocks");
      lane[e] = static_cast<uint8_t>(tb % block);
    }
  DeviceState &d = c->d;
  HIPC(c, hipMalloc(reinterpret_cast<void **>(&d.step_map), sizeof(int) * map.size()));
  HIPC(c, hipMemcpyAsync(d.step_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, c->stream));
  if(!lane.empty()) {
    HIPC(c, hipMalloc(reinterpret_cast<void **>(&d.near_lane), lane.size()));
    HIPC(c, hipMemcpyAsync(d.near_lane, lane.data(), lane.size(), hipMemcpyHostToDevice, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  c->step_block = block;
  c->step_blocks = static_cast<int>(blocks);
  c->step_sync_blocks = sync_blocks;
  return QUEST_OK;
}

int create(quest_ctx *c, const quest_params *p, int n_sys, const int *sizes, const double *pos,
           const double *dip, const double *freq, const double *t1, const double *t2)
{
  c->P = *p;
  c->n_sys = n_sys;
  c->sys_begin.assign(n_sys + 1, 0);
  for(int s = 0; s < n_sys; ++s) c->sys_begin[s + 1] = c->sys_begin[s] + sizes[s];
  const int n = c->sys_begin[n_sys];
  c->n = n;
  HIPC(c, hipSetDevice(p->device));
  HIPC(c, scratch_address(p->device, &c->scratch));  // once per device, with it current
  HIPC(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPC(c, hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, p->device));

  // Morton order inside every system; systems stay contiguous
  c->perm.clear();
  c->perm.reserve(n);
  for(int s = 0; s < n_sys; ++s) {
    const std::vector<int> o = morton_order(c->sys_begin[s], c->sys_begin[s + 1], pos);
    c->perm.insert(c->perm.end(), o.begin(), o.end());
  }
  c->inv.assign(n, 0);
  for(int s = 0; s < n; ++s) c->inv[c->perm[s]] = s;

  // the largest bounding-box diagonal of a system bounds every pair's delay
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  double diag = 0;
  for(int s = 0; s < n_sys; ++s) {
    if(c->sys_begin[s + 1] == c->sys_begin[s]) continue;
    double slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for(int i = c->sys_begin[s]; i < c->sys_begin[s + 1]; ++i)
      for(int a = 0; a < 3; ++a) {
        slo[a] = std::min(slo[a], pos[3 * i + a]);
        shi[a] = std::max(shi[a], pos[3 * i + a]);
      }
    double dd = 0;
    for(int a = 0; a < 3; ++a) {
      dd += (shi[a] - slo[a]) * (shi[a] - slo[a]);
      lo[a] = std::min(lo[a], slo[a]);
      hi[a] = std::max(hi[a], shi[a]);
    }
    diag = std::max(diag, std::sqrt(dd));
  }
  const double cdt = p->c0 * p->dt;
  // delays below 1e6 steps: also keeps the FP32 delay split's rounding
  // constant (exact below 2^22 steps) valid
  if(!(diag / cdt < 1e6)) return fail(c, QUEST_ERR_ARG, "system extent / (c0 dt) too large");
  c->dmax_bound = static_cast<int>(std::floor(diag / cdt)) + 2;

  Consts &k = c->k;
  k.n = n;
  k.order = p->interp_order;
  k.kstride = p->interp_order + 1;  // coefficients per ring entry
  k.ring_slots = next_pow2(c->dmax_bound + 1);
  k.ring_mask = k.ring_slots - 1;
  k.tiles = (n + kBlock - 1) / kBlock;
  k.state_slots = next_pow2(std::max(p->history_window + 1, p->interp_order + 3));
  k.state_mask = k.state_slots - 1;
  k.c = p->c0;
  k.dt = p->dt;
  k.cdt = cdt;
  k.icdt = 1.0 / cdt;
  k.omega = p->laser_omega;
  k.theta = p->laser_omega * p->dt;
  k.k_green = p->mu0 / (4 * M_PI * p->hbar);
  k.hbar = p->hbar;
  k.laser_omega = p->laser_omega;
  k.two_idt2 = 2 / (p->dt * p->dt);
  k.unit_scale = k.two_idt2 / cdt;
  // In these units (lengths in cdt, the term divided by unit_scale) the pair
  // constants are c^2 -> 1/2, w c -> theta/2, c/dt -> 1/2, 2w/dt -> theta and
  // w^2 -> theta^2/2 (oo_l); the pair term is written with those values.
  k.oo_l = 0.5 * k.theta * k.theta;
  {
    // FP32 rotation e^{-i theta xc}, |xc| <= 1/2, u = theta xc, |u| <= U = theta/2:
    // sin u = u - u^3/6 + u^5/120 - ... with u^5 economized over [-U, U]
    // (T5: t^5 ~ (20 t^3 - 5 t)/16, t = u/U), error <= U^5/1920 + U^7/5040;
    // cos u = 1 - u^2/2 + u^4/24, error <= U^6/720.  Both stay below 3e-8 (half
    // an FP32 unit of a unit term) for theta <= 0.28.  The FP64 series
    // (sincos_centred64, the same u, terms to u^9 and u^10) has errors <=
    // U^11/11! and U^12/12!, below 1e-16 for theta <= 0.33.  Beyond these the
    // pair kernels call the library sincos.
    const double t = k.theta, t2 = t * t, U = 0.5 * std::fabs(t), U2 = U * U;
    k.sin_c[0] = t * (1.0 - 5.0 * U2 * U2 / 1920.0);
    k.sin_c[1] = t * t2 * (-1.0 / 6.0 + U2 / 96.0);
    k.cos_c[0] = -t2 / 2.0;
    k.cos_c[1] = t2 * t2 / 24.0;
    // the FP64 series' coefficients in x: sin: (-1)^k t^(2k+1)/(2k+1)!, k = 4..0;
    // cos: (-1)^k t^(2k)/(2k)!, k = 5..1
    for(int j = 0; j < 5; ++j) {
      const int ks = 4 - j, kc = 5 - j;
      double a = t, b = 1.0;  // t^(2 ks + 1) / (2 ks + 1)!, t^(2 kc) / (2 kc)!
      for(int i = 2; i <= 2 * ks + 1; ++i) a = a * t / i;
      for(int i = 1; i <= 2 * kc; ++i) b = b * t / i;
      k.trig64[j] = (ks % 2 ? -a : a);
      k.trig64[5 + j] = (kc % 2 ? -b : b);
    }
    k.trig_poly = std::fabs(t) <= (p->precision == QUEST_PREC_FP64 ? 0.33 : 0.28) ? 1 : 0;
  }
  lagrange_monomials(p->interp_order, k.vinv);
  centre_basis(p->interp_order, k.vinv, k.vinv_c);
  k.half_phase_re = std::cos(-0.5 * k.theta);
  k.half_phase_im = std::sin(-0.5 * k.theta);
  k.dyadic = p->dyadic;
  k.n_systems = n_sys;
  k.block_steps = 1;
  k.compensated = p->precision == QUEST_PREC_MIXED && p->accumulation == QUEST_ACC_COMPENSATED;
  // FP32 positions are stored relative to the centre of all systems' boxes,
  // in units of c0 dt (so a pair's FP32 distance is its delay tau); each
  // coordinate is off by at most half an ulp of the largest scaled centred
  // coordinate, so a pair's tau by at most sqrt(3) ulp.  The arithmetic
  // (squares, sum, rsqrt, one product) adds at most ~3e-7 relative.  eps_rel
  // and eps_abs carry a factor two of margin.
  double ctr[3] = {0, 0, 0}, half = 0;
  if(diag > 0 || n > 0)
    for(int a = 0; a < 3; ++a) {
      ctr[a] = 0.5 * (lo[a] + hi[a]);
      half = std::max(half, 0.5 * (hi[a] - lo[a]));
    }
  const double half_l = half / cdt;
  const double ulp = half_l > 0 ? std::ldexp(1.0, std::ilogb(half_l) - 23) : 0x1p-126;
  k.eps_rel = 8e-7;
  k.eps_abs = 2.0 * std::sqrt(3.0) * ulp + 2e-6;

  // sorted host copies
  std::vector<double4> p64(n), d64(n);
  std::vector<float4> pd32(2 * static_cast<size_t>(n));  // (position, dipole) per dot
  std::vector<double4> pd64(p->precision == QUEST_PREC_FP64 ? 2 * static_cast<size_t>(n) : 0);
  std::vector<double> f(n, 0.0), a1(n, 1.0), a2(n, 1.0);
  std::vector<int> sys_of(n);
  std::vector<int2> ranges(n_sys);
  for(int s = 0; s < n_sys; ++s) {
    ranges[s] = make_int2(c->sys_begin[s], c->sys_begin[s + 1]);
    for(int i = c->sys_begin[s]; i < c->sys_begin[s + 1]; ++i) sys_of[i] = s;
  }
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    p64[s] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
    d64[s] = make_double4(dip[3 * i], dip[3 * i + 1], dip[3 * i + 2], 0.0);
    pd32[2 * s] = make_float4(static_cast<float>((p64[s].x - ctr[0]) / cdt),
                              static_cast<float>((p64[s].y - ctr[1]) / cdt),
                              static_cast<float>((p64[s].z - ctr[2]) / cdt), 0.0f);
    pd32[2 * s + 1] = make_float4(static_cast<float>(d64[s].x), static_cast<float>(d64[s].y),
                                  static_cast<float>(d64[s].z), 0.0f);
    if(!pd64.empty()) {
      pd64[2 * s] = make_double4((p64[s].x - ctr[0]) / cdt, (p64[s].y - ctr[1]) / cdt, (p64[s].z - ctr[2]) / cdt, 0.0);
      pd64[2 * s + 1] = d64[s];
    }
    if(freq) f[s] = freq[i];
    if(t1) a1[s] = t1[i];
    if(t2) a2[s] = t2[i];
  }

  DeviceState &d = c->d;
  const size_t nn = static_cast<size_t>(n);
  const size_t r2size = p->precision == QUEST_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  c->ring_bytes = nn * k.ring_slots * k.kstride * r2size;
  // the FP32 pair kernel addresses a 16-B coefficient plane with 32-bit byte offsets
  if(p->precision == QUEST_PREC_MIXED && nn * k.ring_slots * sizeof(float4) > 0xFFFFFFFFull)
    return fail(c, QUEST_ERR_ARG,
                "mixed precision: dots x ring slots (" + std::to_string(nn) + " x " + std::to_string(k.ring_slots) +
                    ") exceeds 2^28; use QUEST_PREC_FP64 or split the system");
  const size_t ss = static_cast<size_t>(k.state_slots) * nn;
  auto alloc = [&](void **ptr, size_t bytes) -> hipError_t { return hipMalloc(ptr, std::max<size_t>(bytes, 16)); };
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pos64), sizeof(double4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.dip64), sizeof(double4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pd32), sizeof(float4) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pd64), sizeof(double4) * pd64.size()));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.freq), sizeof(double) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.t1), sizeof(double) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.t2), sizeof(double) * nn));
  {
    hipError_t e = alloc(&d.ring, c->ring_bytes);
    if(e != hipSuccess)
      return fail(c, QUEST_ERR_NOMEM, "ring allocation of " + std::to_string(c->ring_bytes) + " bytes failed");
  }
  HIPC(c, alloc(reinterpret_cast<void **>(&d.y0), sizeof(double) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.y1), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.d0), sizeof(double) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.d1), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.cur0), sizeof(double) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.cur1), sizeof(double2) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.yc0), sizeof(double) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.yc1), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.field), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.out), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_start), sizeof(int) * (nn + 1)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.counts), sizeof(int) * (nn + 1)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pcw), sizeof(double) * 4 * kMaxWindow));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.bad_step), sizeof(int)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.step), kStepCounterInts * sizeof(int)));
  HIPC(c, launch_clear(d.step, kStepCounterInts * sizeof(int), c->stream));  // step and every counter
  HIPC(c, alloc(reinterpret_cast<void **>(&d.perm_dev), sizeof(int) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.io), sizeof(double2) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.sys_of), sizeof(int) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.sys_range), sizeof(int2) * n_sys));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.sys_chunks), sizeof(int) * n_sys));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pulses), sizeof(PulseParams) * n_sys * QUEST_MAX_PULSES));
  HIPC(c, hipHostMalloc(reinterpret_cast<void **>(&c->h_stage), sizeof(double2) * nn * 2));

  HIPC(c, hipMemcpyAsync(d.pos64, p64.data(), sizeof(double4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.dip64, d64.data(), sizeof(double4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.pd32, pd32.data(), sizeof(float4) * 2 * nn, hipMemcpyHostToDevice, c->stream));
  if(!pd64.empty())
    HIPC(c, hipMemcpyAsync(d.pd64, pd64.data(), sizeof(double4) * pd64.size(), hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.freq, f.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.t1, a1.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.t2, a2.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.perm_dev, c->inv.data(), sizeof(int) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.sys_of, sys_of.data(), sizeof(int) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.sys_range, ranges.data(), sizeof(int2) * n_sys, hipMemcpyHostToDevice, c->stream));
  c->pulses.assign(static_cast<size_t>(n_sys) * QUEST_MAX_PULSES, PulseParams{});
  int rc = upload_pulses(c);
  if(rc != QUEST_OK) return rc;

  // FP32 source loop: pipelined unless asked otherwise (quest_autotune times
  // both; measured with the default grid and chunks: faster or equal at every
  // size from N = 4096 and the 256-system ensemble to N = 131072)
  k.pipelined = p->precision == QUEST_PREC_MIXED && p->source_loop != QUEST_LOOP_PLAIN;
  // pair-kernel work list sized for one wave of resident blocks; static
  // claiming unless asked otherwise (quest_autotune times both)
  c->blocks_per_cu = default_blocks_per_cu(c);
  k.dyn_items = p->item_claim == QUEST_CLAIM_DYNAMIC;
  rc = build_work(c, p->source_chunk > 0 ? p->source_chunk : default_chunk(c));
  if(rc != QUEST_OK) return rc;

  // zero-delay list: count, scan on the host, fill in source order, coefficients
  HIPC(c, launch_near_count(k, d, p->precision, c->stream));
  std::vector<int> counts(n);
  HIPC(c, hipMemcpyAsync(counts.data(), d.counts, sizeof(int) * nn, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  std::vector<int> start(n + 1, 0);
  for(int i = 0; i < n; ++i) start[i + 1] = start[i] + counts[i];
  c->near_total = start[n];
  const size_t ne = static_cast<size_t>(std::max(c->near_total, 1));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_src), sizeof(int) * ne));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_tgt), sizeof(int) * ne));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_n), sizeof(int) * ne));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_coef), sizeof(double2) * ne * (k.order + 1)));
  HIPC(c, hipMemcpyAsync(d.near_start, start.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice, c->stream));
  HIPC(c, launch_near_fill(k, d, p->precision, c->stream));
  HIPC(c, launch_near_coef(k, d, c->near_total, c->stream));
  std::vector<int> near_src(ne);
  HIPC(c, hipMemcpyAsync(near_src.data(), d.near_src, sizeof(int) * ne, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  rc = pack_step_blocks(c, start, near_src);
  if(rc != QUEST_OK) return rc;
  HIPC(c, alloc(reinterpret_cast<void **>(&d.phase), sizeof(double2) * kPhaseEntries));
  rc = upload_phase(c);
  if(rc != QUEST_OK) return rc;
  {
    const int W = p->history_window;
    c->pcw.assign(4 * W, 0.0);
    if(pc_weights(18, W, 3.15, c->pcw.data(), c->pcw.data() + 2 * W, &c->fc) != 0)
      return fail(c, QUEST_ERR_ARG, "pc weight solve failed");
    rc = upload_pc(c);
    if(rc != QUEST_OK) return rc;
  }
  return start_state(c);
}

int check_params(const quest_params *p, int n)
{
  if(!p) return fail(nullptr, QUEST_ERR_ARG, "params is NULL");
  if(n < 1) return fail(nullptr, QUEST_ERR_ARG, "n_dots must be >= 1");
  if(p->interp_order < 1 || p->interp_order > kMaxOrder)
    return fail(nullptr, QUEST_ERR_ARG, "interp_order must be in 1..6");
  if(p->history_window < 2 || p->history_window > kMaxWindow)
    return fail(nullptr, QUEST_ERR_ARG, "history_window must be in 2..64");
  if(p->precision != QUEST_PREC_FP64 && p->precision != QUEST_PREC_MIXED)
    return fail(nullptr, QUEST_ERR_ARG, "unknown precision");
  if(p->dyadic != QUEST_DYADIC_ROTATING && p->dyadic != QUEST_DYADIC_LAB)
    return fail(nullptr, QUEST_ERR_ARG, "unknown dyadic");
  if(!(p->dt > 0) || !(p->c0 > 0) || !(p->hbar > 0) || !std::isfinite(p->mu0) ||
     !std::isfinite(p->laser_omega))
    return fail(nullptr, QUEST_ERR_ARG, "constants must be finite, dt, c0, hbar > 0");
  if(p->source_chunk < 0) return fail(nullptr, QUEST_ERR_ARG, "source_chunk must be >= 0");
  if(p->item_claim < QUEST_CLAIM_AUTO || p->item_claim > QUEST_CLAIM_DYNAMIC)
    return fail(nullptr, QUEST_ERR_ARG, "item_claim must be QUEST_CLAIM_AUTO, _STATIC or _DYNAMIC");
  if(p->source_loop < QUEST_LOOP_AUTO || p->source_loop > QUEST_LOOP_PIPELINED)
    return fail(nullptr, QUEST_ERR_ARG, "source_loop must be QUEST_LOOP_AUTO, _PLAIN or _PIPELINED");
  if(p->accumulation != QUEST_ACC_BLOCK && p->accumulation != QUEST_ACC_COMPENSATED)
    return fail(nullptr, QUEST_ERR_ARG, "accumulation must be QUEST_ACC_BLOCK or QUEST_ACC_COMPENSATED");
  if(p->block_steps != 0 && p->block_steps != kBlockStepsShort && p->block_steps != kBlockStepsLong)
    return fail(nullptr, QUEST_ERR_ARG, "block_steps must be 0 (auto), 8 or 16");
  return QUEST_OK;
}

PulseParams to_params(const quest_pulse *pl)
{
  PulseParams q{};
  if(!pl) return q;
  q.on = 1;
  q.amplitude = pl->amplitude;
  q.delay = pl->delay;
  q.width = pl->width;
  q.freq = pl->freq;
  for(int a = 0; a < 3; ++a) {
    q.k[a] = pl->wavevector[a];
    q.pol[a] = pl->polarization[a];
  }
  return q;
}

void free_table(quest_ctx *c)
{
  DeviceState &d = c->d;
  for(void *p : {static_cast<void *>(d.samp), static_cast<void *>(d.table), static_cast<void *>(d.table_n),
                 static_cast<void *>(d.sys_toff)})
    if(p) (void)hipFree(p);
  d.samp = nullptr;
  d.table = nullptr;
  d.table_n = nullptr;
  d.sys_toff = nullptr;
  c->k.pair_table = 0;
  c->k.table_plane = 0;
}

// Coefficient-table mode: the reference's per-pair coefficient table
// (history_interaction.cpp:21-45) built on the device for every system, and
// a ring of rho01 samples for the table's gather.
int enable_table(quest_ctx *c)
{
  const int M = c->P.interp_order;
  if(c->P.precision != QUEST_PREC_FP64)
    return fail(c, QUEST_ERR_ARG, "the coefficient table is FP64: create the context with QUEST_PREC_FP64");
  if(c->dmax_bound + M + 2 > 32767)
    return fail(c, QUEST_ERR_ARG, "delays exceed the table's 16-bit delay field");
  std::vector<long long> toff(c->n_sys);
  long long plane = 0;
  for(int s = 0; s < c->n_sys; ++s) {
    const long long ns = c->sys_begin[s + 1] - c->sys_begin[s];
    toff[s] = plane;
    plane += ns * ns;
  }
  const size_t coef_bytes = static_cast<size_t>(plane) * (M + 1) * sizeof(double2);
  const size_t n_bytes = (static_cast<size_t>(plane) * sizeof(int16_t) + 15) / 16 * 16;
  Consts &k = c->k;
  k.samp_slots = 1;
  while(k.samp_slots < c->dmax_bound + M + 2) k.samp_slots <<= 1;
  k.samp_mask = k.samp_slots - 1;
  const size_t samp_bytes = static_cast<size_t>(c->n) * k.samp_slots * sizeof(double2);
  size_t free_b = 0, total_b = 0;
  HIPC(c, hipMemGetInfo(&free_b, &total_b));
  const size_t need = coef_bytes + n_bytes + samp_bytes;
  if(need + (size_t(1) << 30) > free_b)
    return fail(c, QUEST_ERR_NOMEM, "coefficient table needs " + std::to_string(need >> 20) + " MiB, " +
                                        std::to_string(free_b >> 20) + " MiB free");
  DeviceState &d = c->d;
  if(hipMalloc(&d.table, coef_bytes) != hipSuccess || hipMalloc(&d.table_n, n_bytes) != hipSuccess ||
     hipMalloc(&d.samp, samp_bytes) != hipSuccess ||
     hipMalloc(&d.sys_toff, sizeof(long long) * c->n_sys) != hipSuccess) {
    free_table(c);
    return fail(c, QUEST_ERR_NOMEM, "coefficient table allocation failed");
  }
  HIPC(c, hipMemcpyAsync(d.sys_toff, toff.data(), sizeof(long long) * c->n_sys, hipMemcpyHostToDevice,
                         c->stream));
  k.table_plane = plane;
  for(int s = 0; s < c->n_sys; ++s)
    HIPC(c, launch_table_build(k, d, make_int2(c->sys_begin[s], c->sys_begin[s + 1]), toff[s], c->stream));
  HIPC(c, launch_clear(d.samp, samp_bytes, c->stream));  // rho01 = 0 before step 0 (main.cpp:32-36)
  HIPC(c, hipStreamSynchronize(c->stream));
  k.pair_table = 1;
  return QUEST_OK;
}

void free_blocked(quest_ctx *c)
{
  DeviceState &d = c->d;
  for(void *p : {d.hist, d.rot, static_cast<void *>(d.short_start), static_cast<void *>(d.short_src),
                 static_cast<void *>(d.short_sum)})
    if(p) (void)hipFree(p);
  for(int l = 0; l < kWindowLists; ++l) {
    for(void *p : {static_cast<void *>(d.win_start[l]), static_cast<void *>(d.win_src[l]),
                   static_cast<void *>(d.win_sum[l])})
      if(p) (void)hipFree(p);
    d.win_start[l] = d.win_src[l] = nullptr;
    d.win_sum[l] = nullptr;
    c->k.win_steps[l] = 0;
  }
  d.hist = d.rot = nullptr;
  d.short_start = d.short_src = nullptr;
  d.short_sum = nullptr;
  c->k.pair_blocked = 0;
  c->k.block_steps = 1;
  c->short_total = 0;
  c->hist_bytes = 0;
}

// A list of every target's sources whose delay lies in [lo, hi) (the
// precision's split), CSR by target: *start [n + 1] and *src on the device,
// the number of pairs in *total.
int build_list(quest_ctx *c, int lo, int hi, int **start_dev, int **src_dev, int *total)
{
  DeviceState &d = c->d;
  const size_t nn = static_cast<size_t>(c->n);
  HIPC(c, launch_delay_list(c->k, d, c->P.precision, lo, hi, false, nullptr, nullptr, nullptr, d.counts, c->stream));
  std::vector<int> counts(nn);
  HIPC(c, hipMemcpyAsync(counts.data(), d.counts, sizeof(int) * nn, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  std::vector<int> start(nn + 1, 0);
  for(size_t i = 0; i < nn; ++i) {
    if(start[i] > INT_MAX - counts[i]) return fail(c, QUEST_ERR_ARG, "time-blocked method: list beyond 2^31 pairs");
    start[i + 1] = start[i] + counts[i];
  }
  *total = start[nn];
  if(hipMalloc(reinterpret_cast<void **>(start_dev), sizeof(int) * (nn + 1)) != hipSuccess ||
     hipMalloc(reinterpret_cast<void **>(src_dev), sizeof(int) * std::max(*total, 1)) != hipSuccess)
    return fail(c, QUEST_ERR_NOMEM, "time-blocked method: list allocation failed");
  HIPC(c, hipMemcpyAsync(*start_dev, start.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice, c->stream));
  HIPC(c, launch_delay_list(c->k, d, c->P.precision, lo, hi, true, *start_dev, *src_dev, nullptr, d.counts,
                            c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  return QUEST_OK;
}

// Time-blocked mode: the steps per pass K, the rho01 sample ring, the
// rotation table e^{-i theta (n + 1/2)}, the short list (pairs of delay
// 1 .. 3), the 4-step window list (4 .. 7) and with K = 16 the 8-step one
// (8 .. 15), all by the precision's own split, as the pair kernels classify
// them.  The ring holds
// more than the largest delay + order + K steps, so that the pass of the
// current block can be formed again from it at any step of the block
// (refill_block).
int enable_blocked(quest_ctx *c)
{
  Consts &k = c->k;
  DeviceState &d = c->d;
  const int M = c->P.interp_order;
  const bool f64 = c->P.precision == QUEST_PREC_FP64;
  const size_t r2 = f64 ? sizeof(double2) : sizeof(float2);
  const size_t nn = static_cast<size_t>(c->n);
  // steps per pass: the caller's, else 16 where the pairs of delay below 16
  // (the lists of a 16-step pass) are under 2 % of all ordered pairs
  std::vector<int> counts(nn);
  auto count_short = [&](int hi) -> int64_t {
    if(launch_delay_list(k, d, c->P.precision, 1, hi, false, nullptr, nullptr, nullptr, d.counts, c->stream) !=
           hipSuccess ||
       hipMemcpyAsync(counts.data(), d.counts, sizeof(int) * nn, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
       hipStreamSynchronize(c->stream) != hipSuccess)
      return -1;
    int64_t total = 0;
    for(size_t i = 0; i < nn; ++i) total += counts[i];
    return total;
  };
  int K = c->P.block_steps;
  if(K == 0) {
    const int64_t short16 = count_short(kBlockStepsLong);
    if(short16 < 0) return fail(c, QUEST_ERR_HIP, "time-blocked method: short-list scan failed");
    double ordered = 0;
    for(int s = 0; s < c->n_sys; ++s) {
      const double ns = c->sys_begin[s + 1] - c->sys_begin[s];
      ordered += ns * (ns - 1);
    }
    K = static_cast<double>(short16) <= kLongPassShortFraction * ordered ? kBlockStepsLong : kBlockStepsShort;
  }
  k.block_steps = K;
  k.win_steps[0] = kShortListEnd;
  k.win_steps[1] = K == kBlockStepsLong ? kBlockStepsShort : 0;
  k.hist_slots = next_pow2(c->dmax_bound + M + K + 1);
  k.hist_mask = k.hist_slots - 1;
  k.hist_row = (k.hist_slots + K + M - 1 + 1) / 2 * 2;
  // the pair kernel addresses a window with a 32-bit byte offset
  if(nn * k.hist_row * r2 > 0xFFFFFFFFull)
    return fail(c, QUEST_ERR_ARG, "time-blocked method: dots x sample-ring row (" + std::to_string(nn) + " x " +
                                      std::to_string(k.hist_row) + ") exceeds 4 GiB of samples");
  c->hist_bytes = (nn * k.hist_row * r2 + 15) / 16 * 16;
  if(hipMalloc(&d.hist, c->hist_bytes) != hipSuccess || hipMalloc(&d.rot, k.hist_slots * r2) != hipSuccess ||
     hipMalloc(reinterpret_cast<void **>(&d.short_sum), sizeof(double2) * std::max<size_t>(nn, 1)) != hipSuccess ||
     (k.win_steps[0] && hipMalloc(reinterpret_cast<void **>(&d.win_sum[0]),
                                  sizeof(double2) * k.win_steps[0] * std::max<size_t>(nn, 1)) != hipSuccess) ||
     (k.win_steps[1] && hipMalloc(reinterpret_cast<void **>(&d.win_sum[1]),
                                  sizeof(double2) * k.win_steps[1] * std::max<size_t>(nn, 1)) != hipSuccess)) {
    free_blocked(c);
    return fail(c, QUEST_ERR_NOMEM, "time-blocked method: allocation failed");
  }
  {
    // e^{-i theta (j + 1/2)}, each from a long-double argument reduction
    const long double theta = static_cast<long double>(k.omega) * static_cast<long double>(k.dt);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<double2> h(k.hist_slots);
    std::vector<float2> hf(k.hist_slots);
    for(int j = 0; j < k.hist_slots; ++j) {
      const long double a = fmodl(-theta * (static_cast<long double>(j) + 0.5L), two_pi);
      h[j] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
      hf[j] = make_float2(static_cast<float>(cosl(a)), static_cast<float>(sinl(a)));
    }
    if(f64) HIPC(c, hipMemcpyAsync(d.rot, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, c->stream));
    else HIPC(c, hipMemcpyAsync(d.rot, hf.data(), sizeof(float2) * hf.size(), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));  // h, hf go out of scope
  }
  // the short list (delays 1 .. 3), the window lists (4 .. 7; 8 .. 15 with
  // 16-step passes)
  int short_pairs = 0, win_pairs[kWindowLists] = {};
  int rc = build_list(c, 1, kShortListEnd, &d.short_start, &d.short_src, &short_pairs);
  if(rc == QUEST_OK)
    rc = build_list(c, kShortListEnd, kBlockStepsShort, &d.win_start[0], &d.win_src[0], &win_pairs[0]);
  if(rc == QUEST_OK && k.win_steps[1])
    rc = build_list(c, kBlockStepsShort, K, &d.win_start[1], &d.win_src[1], &win_pairs[1]);
  c->short_total = short_pairs + win_pairs[0];
  if(rc != QUEST_OK) {
    const std::string msg = c->err;
    free_blocked(c);
    return fail(c, rc, msg);
  }
  HIPC(c, launch_clear(d.hist, c->hist_bytes, c->stream));  // rho01 = 0 before step 0 (main.cpp:32-36)
  HIPC(c, hipStreamSynchronize(c->stream));
  k.pair_blocked = 1;
  return QUEST_OK;
}

// Time-blocked method, at a step inside a block: forms the current block's
// pass again (into d's partial sums), e.g. after the partial sums moved to
// another buffer or the work split changed.  The block started at
// K floor(step / K); the sample ring still holds everything its pass reads.
int refill_block(quest_ctx *c, DeviceState &d, int step)
{
  if(!c->k.pair_blocked || step % c->k.block_steps == 0) return QUEST_OK;
  HIPC(c, launch_set_step(d, step - step % c->k.block_steps, c->stream));
  const int rc = pair_sum(c, d, false);
  if(rc != QUEST_OK) return rc;
  HIPC(c, launch_set_step(d, step, c->stream));
  return QUEST_OK;
}

// Pulse list of system s on the host copy: the given pulses first, the rest off.
void put_pulses(quest_ctx *c, int s, int n, const quest_pulse *pl)
{
  PulseParams *row = c->pulses.data() + static_cast<size_t>(s) * QUEST_MAX_PULSES;
  for(int p = 0; p < QUEST_MAX_PULSES; ++p) row[p] = p < n ? to_params(pl + p) : PulseParams{};
}

int create_checked(const quest_params *params, int n_sys, const int *sizes, const double *pos,
                   const double *dip, const double *freq, const double *t1, const double *t2,
                   quest_ctx **out)
{
  if(!out) return fail(nullptr, QUEST_ERR_ARG, "out is NULL");
  *out = nullptr;
  int64_t n = 0;
  for(int s = 0; s < n_sys; ++s) {
    if(sizes[s] < 0) return fail(nullptr, QUEST_ERR_ARG, "system sizes must be >= 0");
    n += sizes[s];
  }
  if(n > INT_MAX / 2) return fail(nullptr, QUEST_ERR_ARG, "too many dots");
  int rc = check_params(params, static_cast<int>(n));
  if(rc != QUEST_OK) return rc;
  if(!pos || !dip) return fail(nullptr, QUEST_ERR_ARG, "pos and dip are required");
  for(int64_t i = 0; i < 3 * n; ++i)
    if(!std::isfinite(pos[i])) return fail(nullptr, QUEST_ERR_ARG, "non-finite position");
  quest_ctx *c = new(std::nothrow) quest_ctx;
  if(!c) return fail(nullptr, QUEST_ERR_NOMEM, "context allocation failed");
  rc = create(c, params, n_sys, sizes, pos, dip, freq, t1, t2);
  if(rc != QUEST_OK) {
    g_create_error = c->err;
    free_device(c);
    delete c;
    return rc;
  }
  *out = c;
  return QUEST_OK;
}

// Entry checks of quest_run and quest_run_noalloc; the first call starts the
// whole-run sequence at step 0.
int run_begin(quest_ctx *c, int n_steps)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(n_steps < 0) return fail(c, QUEST_ERR_ARG, "n_steps must be >= 0");
  if(c->mode == 1) return fail(c, QUEST_ERR_STATE, "context is driven by quest_hist_evaluate");
  HIPC(c, hipSetDevice(c->P.device));
  if(c->mode == 0) {
    c->mode = 2;
    c->cur_step = 0;
    HIPC(c, launch_set_step(c->d, 0, c->stream));
  }
  if(c->bad >= 0) return fail(c, QUEST_ERR_DOMAIN, "unbounded history values at or before step " + std::to_string(c->bad));
  if(n_steps > INT_MAX - 1 - c->cur_step) return fail(c, QUEST_ERR_ARG, "steps beyond 2^31 - 1");
  return QUEST_OK;
}

// After the steps are queued: waits for them, advances the host step and
// reports the first non-finite step (integrator.cpp:187-207).
int run_end(quest_ctx *c, int n_steps)
{
  int bad = INT_MAX;
  HIPC(c, hipMemcpyAsync(&bad, c->d.bad_step, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->cur_step += n_steps;
  if(bad != INT_MAX) {
    c->bad = bad;
    return fail(c, QUEST_ERR_DOMAIN, "unbounded history values at or before step " + std::to_string(bad));
  }
  return QUEST_OK;
}

}  // namespace

extern "C" {

int quest_ctx_create(const quest_params *params, int n_dots, const double *pos,
                     const double *dip, const double *freq, const double *t1,
                     const double *t2, quest_ctx **out)
{
  const int sizes[1] = {n_dots};
  return create_checked(params, 1, sizes, pos, dip, freq, t1, t2, out);
}

int quest_batch_create(const quest_params *params, int n_systems, const int *n_dots,
                       const double *pos, const double *dip, const double *freq,
                       const double *t1, const double *t2, quest_ctx **out)
{
  if(n_systems < 1 || !n_dots) return fail(nullptr, QUEST_ERR_ARG, "n_systems >= 1 and n_dots required");
  return create_checked(params, n_systems, n_dots, pos, dip, freq, t1, t2, out);
}

void quest_ctx_destroy(quest_ctx *ctx)
{
  if(!ctx) return;
  (void)hipSetDevice(ctx->P.device);
  if(ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  free_device(ctx);
  delete ctx;
}

const char *quest_last_error(const quest_ctx *ctx)
{
  return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int quest_hist_evaluate(quest_ctx *c, int step, const double *rho01, double *out)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!rho01 || !out) return fail(c, QUEST_ERR_ARG, "rho01 and out are required");
  if(step < 0) return fail(c, QUEST_ERR_ARG, "step must be >= 0");
  if(c->mode == 2) return fail(c, QUEST_ERR_STATE, "context is driven by quest_run");
  HIPC(c, hipSetDevice(c->P.device));
  if(c->mode == 0) {
    c->mode = 1;
    c->cur_step = -1;
  }
  if(c->cur_step < 0 && step != 0)
    return fail(c, QUEST_ERR_STATE, "the first evaluation must be at step 0 (every step is evaluated in order)");
  if(c->cur_step >= 0 && step < c->cur_step)
    return fail(c, QUEST_ERR_STATE, "step went backwards");
  if(c->cur_step >= 0 && step > c->cur_step + 1)
    return fail(c, QUEST_ERR_STATE, "step skipped: every step must be evaluated");
  if(step != c->cur_step) {
    if(c->cur_step >= 0) HIPC(c, launch_hist_commit(c->k, c->d, c->P.precision, c->cur_step, c->stream));
    HIPC(c, launch_set_step(c->d, step, c->stream));
    int rc = (!c->k.pair_blocked || step % c->k.block_steps == 0) ? pair_sum(c) : QUEST_OK;
    if(rc == QUEST_OK && c->k.pair_blocked) rc = list_sums(c, c->d, true, step);
    if(rc != QUEST_OK) return rc;
    HIPC(c, launch_hist_reduce(c->k, c->d, c->stream));
    c->cur_step = step;
  }
  // caller order crosses PCIe as is (sequential copies through the pinned
  // stage); the permutation to and from the sorted order runs on the device
  const int n = c->n;
  std::memcpy(c->h_stage, rho01, sizeof(double2) * n);
  HIPC(c, hipMemcpyAsync(c->d.io, c->h_stage, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream));
  HIPC(c, launch_permute_in(c->k, c->d, c->stream));
  HIPC(c, launch_hist_near(c->k, c->d, c->stream));
  HIPC(c, launch_permute_out(c->k, c->d, c->stream));
  HIPC(c, hipMemcpyAsync(c->h_stage + n, c->d.io + n, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->h_stage + n, sizeof(double2) * n);
  if(c->timing) return collect_timing(c);
  return QUEST_OK;
}

int quest_set_pulse(quest_ctx *c, const quest_pulse *pl)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  HIPC(c, hipSetDevice(c->P.device));
  for(int s = 0; s < c->n_sys; ++s) put_pulses(c, s, pl ? 1 : 0, pl);
  return upload_pulses(c);
}

int quest_set_pulses(quest_ctx *c, const quest_pulse *pulses)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!pulses) return fail(c, QUEST_ERR_ARG, "pulses is required (one per system)");
  HIPC(c, hipSetDevice(c->P.device));
  for(int s = 0; s < c->n_sys; ++s) put_pulses(c, s, 1, pulses + s);
  return upload_pulses(c);
}

int quest_set_pulse_list(quest_ctx *c, int system, int n_pulses, const quest_pulse *pulses)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(system < 0 || system >= c->n_sys)
    return fail(c, QUEST_ERR_ARG, "system must be in [0, " + std::to_string(c->n_sys) + ")");
  if(n_pulses < 0 || n_pulses > QUEST_MAX_PULSES)
    return fail(c, QUEST_ERR_ARG, "n_pulses must be in [0, " + std::to_string(QUEST_MAX_PULSES) + "]");
  if(n_pulses > 0 && !pulses) return fail(c, QUEST_ERR_ARG, "pulses is required when n_pulses > 0");
  HIPC(c, hipSetDevice(c->P.device));
  put_pulses(c, system, n_pulses, pulses);
  return upload_pulses(c);
}

int quest_set_pair_method(quest_ctx *c, int method)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(method != QUEST_PAIR_ON_THE_FLY && method != QUEST_PAIR_TABLE && method != QUEST_PAIR_TIME_BLOCKED)
    return fail(c, QUEST_ERR_ARG,
                "method must be QUEST_PAIR_ON_THE_FLY, QUEST_PAIR_TABLE or QUEST_PAIR_TIME_BLOCKED");
  if(method == c->pair_method) return QUEST_OK;
  if(c->mode != 0) return fail(c, QUEST_ERR_STATE, "choose the pair method before the first step");
  HIPC(c, hipSetDevice(c->P.device));
  drop_graph(c);
  free_table(c);
  free_blocked(c);
  c->pair_method = QUEST_PAIR_ON_THE_FLY;
  if(method == QUEST_PAIR_TABLE) {
    const int rc = enable_table(c);
    if(rc != QUEST_OK) return rc;
  } else if(method == QUEST_PAIR_TIME_BLOCKED) {
    const int rc = enable_blocked(c);
    if(rc != QUEST_OK) return rc;
  }
  c->pair_method = method;
  c->blocks_per_cu = default_blocks_per_cu(c);
  // the new method's default chunk unless the caller fixed one (a length
  // quest_autotune chose was timed with the other method's kernel)
  return build_work(c, c->P.source_chunk > 0 ? c->P.source_chunk : default_chunk(c));
}

int quest_pulse_evaluate(quest_ctx *c, int step, double *out)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!out) return fail(c, QUEST_ERR_ARG, "out is required");
  if(c->mode == 2) return fail(c, QUEST_ERR_STATE, "context is driven by quest_run");
  HIPC(c, hipSetDevice(c->P.device));
  // the device step counter belongs to the per-call sequence: set, use, restore
  HIPC(c, launch_set_step(c->d, step, c->stream));
  HIPC(c, launch_pulse(c->k, c->d, c->stream));
  if(c->mode == 1 && c->cur_step >= 0) HIPC(c, launch_set_step(c->d, c->cur_step, c->stream));
  // back to caller order on the device, then one sequential copy
  const int n = c->n;
  HIPC(c, launch_permute_out(c->k, c->d, c->stream));
  HIPC(c, hipMemcpyAsync(c->h_stage + n, c->d.io + n, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->h_stage + n, sizeof(double2) * n);
  return QUEST_OK;
}

int quest_pc_weights(int n_lambda, int n_time, double radius, double *ps, double *cs,
                     double *future_coef)
{
  if(n_lambda < 1 || n_time < 2 || n_time > kMaxWindow || !ps || !cs || !future_coef)
    return fail(nullptr, QUEST_ERR_ARG, "invalid pc weight arguments");
  return pc_weights(n_lambda, n_time, radius, ps, cs, future_coef) == 0
             ? QUEST_OK
             : fail(nullptr, QUEST_ERR_ARG, "pc weight solve failed");
}

int quest_set_pc_weights(quest_ctx *c, int n_time, const double *ps, const double *cs,
                         double future_coef)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(n_time != c->P.history_window || !ps || !cs)
    return fail(c, QUEST_ERR_ARG, "n_time must equal history_window; ps, cs required");
  const int W = n_time;
  c->pcw.assign(ps, ps + 2 * W);
  c->pcw.insert(c->pcw.end(), cs, cs + 2 * W);
  c->fc = future_coef;
  HIPC(c, hipSetDevice(c->P.device));
  return upload_pc(c);
}

int quest_run(quest_ctx *c, int n_steps)
{
  int rc = run_begin(c, n_steps);
  if(rc != QUEST_OK) return rc;
  int step = c->cur_step;
  const int end = step + n_steps;
  if(c->slice_buf != 0 && n_steps > 0) {  // the block's partial sums are in the static scratch
    rc = refill_block(c, c->d, step);
    if(rc != QUEST_OK) return rc;
    c->slice_buf = 0;
  }
  if(c->timing) {
    // kernel by kernel, so that HIP events can bracket every pair-kernel launch
    for(; step < end; ++step) {
      rc = enqueue_step(c, c->d, true, step);
      if(rc != QUEST_OK) return rc;
    }
  } else {
    // steps one by one up to a graph boundary, whole graphs, then the rest
    const int G = graph_steps(c);
    for(; step < end && step % G != 0; ++step) {
      rc = enqueue_step(c, c->d, false, step);
      if(rc != QUEST_OK) return rc;
    }
    if(end - step >= G) {
      rc = ensure_graph(c);
      if(rc != QUEST_OK) return rc;
      for(; end - step >= G; step += G) HIPC(c, hipGraphLaunch(c->graph_exec, c->stream));
    }
    for(; step < end; ++step) {
      rc = enqueue_step(c, c->d, false, step);
      if(rc != QUEST_OK) return rc;
    }
  }
  if(c->timing) {
    HIPC(c, hipStreamSynchronize(c->stream));
    rc = collect_timing(c);
    if(rc != QUEST_OK) return rc;
  }
  return run_end(c, n_steps);
}

int quest_run_noalloc(quest_ctx *c, int n_steps)
{
  if(c && !c->scratch) return fail(c, QUEST_ERR_HIP, "the static scratch array has no device address");
  // the pair kernel's partial rows (source chunks, plus the spare row of empty items)
  // (time-blocked method: one slice of them per step of the pass)
  const size_t rows = c ? static_cast<size_t>(c->k.max_chunks + 1) * (c->k.pair_blocked ? c->k.block_steps : 1) : 0;
  if(c && sizeof(double2) * static_cast<size_t>(c->n) * rows > kScratchBytes)
    return fail(c, QUEST_ERR_ARG,
                "quest_run_noalloc: the pair sums need " + std::to_string(rows) + " x " + std::to_string(c->n) +
                    " x 16 B, more than the static scratch of " + std::to_string(kScratchBytes >> 20) +
                    " MiB; use quest_run or fewer source chunks");
  int rc = run_begin(c, n_steps);
  if(rc != QUEST_OK) return rc;
  DeviceState d = c->d;
  d.partial = c->scratch;
  int step = c->cur_step;
  if(c->slice_buf != 1 && n_steps > 0) {  // the block's partial sums are in the context's buffer
    rc = refill_block(c, d, step);
    if(rc != QUEST_OK) return rc;
    c->slice_buf = 1;
  }
  for(int i = 0; i < n_steps; ++i, ++step) {
    rc = enqueue_step(c, d, false, step);
    if(rc != QUEST_OK) return rc;
  }
  return run_end(c, n_steps);
}

int quest_autotune(quest_ctx *c, int launches, int *chunk_len)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(launches < 1) launches = 3;
  HIPC(c, hipSetDevice(c->P.device));
  if(c->k.n_work == 0) {
    if(chunk_len) *chunk_len = c->chunk_len;
    return QUEST_OK;
  }
  int nmax = 2;
  for(int s = 0; s < c->n_sys; ++s) nmax = std::max(nmax, c->sys_begin[s + 1] - c->sys_begin[s]);
  // chunk lengths: the default scaled by 1/4 .. 4, and the largest system cut
  // into 8, 16, 24, 32 or 48 equal source ranges (multiples of the 8 XCDs)
  const int base = default_chunk(c);
  std::vector<int> cand;
  auto add_len = [&](int len) {
    len = std::max(32, std::min(len, nmax));
    if(std::find(cand.begin(), cand.end(), len) == cand.end()) cand.push_back(len);
  };
  for(int f : {-2, -1, 0, 1, 2}) add_len(f >= 0 ? base << f : base >> -f);
  for(int cnt : {8, 16, 24, 32, 48})
    if(nmax / cnt >= 64) add_len((nmax + cnt - 1) / cnt);
  // The pair kernel reads the ring and the step counter and writes only the
  // partial sums, which every step recomputes from scratch, so timing it
  // between steps changes no state.  The counter is set to the next step (the
  // value it holds between steps; before the first step it may not be set yet).
  HIPC(c, launch_set_step(c->d, std::max(c->cur_step, 0), c->stream));
  const bool mixed = c->P.precision == QUEST_PREC_MIXED;
  struct Events {  // destroyed on every return path
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Events()
    {
      if(e0) (void)hipEventDestroy(e0);
      if(e1) (void)hipEventDestroy(e1);
    }
  } ev;
  HIPC(c, hipEventCreate(&ev.e0));
  HIPC(c, hipEventCreate(&ev.e1));
  struct Cfg {
    int len, dyn, pf, occ;  // occ: resident blocks per CU left out of the persistent grid
    float ms;
  };
  // occupancy is a candidate dimension: at fewer waves per SIMD the FP64
  // kernel's gathers share the L1 with fewer waves (measured up to 11 % faster
  // with 4 of 6 blocks per CU at N = 98304, slower at N = 131072), and the
  // grid's last round of items fills differently
  auto time_cfg = [&](Cfg &g) -> int {
    c->k.dyn_items = g.dyn;
    c->k.pipelined = g.pf;
    c->blocks_per_cu = std::max(1, pair_blocks_per_cu(c->k, c->P.precision) - g.occ);
    int rc = build_work(c, g.len);  // the list's layout depends on the claiming mode and the grid
    if(rc != QUEST_OK) return rc;
    HIPC(c, launch_pair_past(c->k, c->d, c->P.precision, c->stream));  // warm-up
    HIPC(c, hipEventRecord(ev.e0, c->stream));
    for(int i = 0; i < launches; ++i) HIPC(c, launch_pair_past(c->k, c->d, c->P.precision, c->stream));
    HIPC(c, hipEventRecord(ev.e1, c->stream));
    HIPC(c, hipEventSynchronize(ev.e1));
    HIPC(c, hipEventElapsedTime(&g.ms, ev.e0, ev.e1));
    return QUEST_OK;
  };
  // claiming and source loop: the caller's choice, else static and (mixed) the
  // default loop in the first stage
  const bool dyn_fixed = c->P.item_claim != QUEST_CLAIM_AUTO;
  const bool pf_fixed = !mixed || c->P.source_loop != QUEST_LOOP_AUTO || c->k.pair_blocked;
  const int dyn0 = c->P.item_claim == QUEST_CLAIM_DYNAMIC ? 1 : 0;
  const int pf0 = mixed ? (c->P.source_loop == QUEST_LOOP_AUTO ? c->k.pipelined
                                                               : c->P.source_loop == QUEST_LOOP_PIPELINED)
                        : 0;
  // stage 1: chunk length x occupancy (all resident blocks per CU, one and two fewer)
  std::vector<Cfg> tried;
  for(int len : cand)
    for(int occ = 0; occ <= 2; ++occ) {
      c->k.dyn_items = dyn0;
      c->k.pipelined = pf0;
      if(occ > 0 && pair_blocks_per_cu(c->k, c->P.precision) - occ < 2) continue;
      Cfg g{len, dyn0, pf0, occ, 0.0f};
      int rc = time_cfg(g);
      if(rc != QUEST_OK) return rc;
      tried.push_back(g);
    }
  std::stable_sort(tried.begin(), tried.end(), [](const Cfg &x, const Cfg &y) { return x.ms < y.ms; });
  // stage 2: the other claiming mode and source loop at the three fastest
  const size_t top = std::min<size_t>(3, tried.size());
  Cfg best = tried[0];
  for(size_t t = 0; t < top; ++t)
    for(int dyn : {0, 1})
      for(int pf : {0, 1}) {
        if((dyn_fixed && dyn != dyn0) || (pf_fixed && pf != pf0) || (dyn == dyn0 && pf == pf0)) continue;
        Cfg g{tried[t].len, dyn, pf, tried[t].occ, 0.0f};
        c->k.dyn_items = dyn;
        c->k.pipelined = pf;
        if(g.occ > 0 && pair_blocks_per_cu(c->k, c->P.precision) - g.occ < 2) continue;
        int rc = time_cfg(g);
        if(rc != QUEST_OK) return rc;
        if(g.ms < best.ms) best = g;
      }
  c->k.dyn_items = best.dyn;
  c->k.pipelined = best.pf;
  c->blocks_per_cu = std::max(1, pair_blocks_per_cu(c->k, c->P.precision) - best.occ);
  int rc = build_work(c, best.len);
  if(rc != QUEST_OK) return rc;
  if(c->k.pair_blocked && c->mode != 0 && c->cur_step >= 0) {
    // the timed passes overwrote the current block's partial sums: per-call
    // mode has formed the block of its current step, whole-run mode that of
    // the step before its next one unless that starts a block
    if(c->mode == 1 && c->cur_step % c->k.block_steps == 0) {
      HIPC(c, launch_set_step(c->d, c->cur_step, c->stream));
      rc = pair_sum(c, c->d, false);
    } else {
      rc = refill_block(c, c->d, c->cur_step);
    }
    if(rc != QUEST_OK) return rc;
    c->slice_buf = 0;
  }
  if(c->mode == 1 && c->cur_step >= 0) {
    // per-call mode keeps the current step's pair sum in `field`: recompute it
    if(!c->k.pair_blocked) {
      rc = pair_sum(c);
      if(rc != QUEST_OK) return rc;
    }
    HIPC(c, launch_hist_reduce(c->k, c->d, c->stream));
  }
  HIPC(c, hipStreamSynchronize(c->stream));
  if(chunk_len) *chunk_len = best.len;
  return QUEST_OK;
}

int quest_current_step(const quest_ctx *c, int *step)
{
  if(!c || !step) return QUEST_ERR_ARG;
  *step = c->cur_step;
  return QUEST_OK;
}

int quest_bad_step(const quest_ctx *c, int *step)
{
  if(!c || !step) return QUEST_ERR_ARG;
  *step = c->bad;
  return QUEST_OK;
}

int quest_get_state(quest_ctx *c, double *rho, double *drho)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!rho) return fail(c, QUEST_ERR_ARG, "rho is required");
  if(c->mode != 2 || c->cur_step < 1) return fail(c, QUEST_ERR_STATE, "no finished step (quest_run)");
  HIPC(c, hipSetDevice(c->P.device));
  const int n = c->n;
  const size_t off = static_cast<size_t>((c->cur_step - 1) & c->k.state_mask) * n;
  // rho00 and its derivative are real on the device (imaginary part 0)
  std::vector<double> h0(2 * static_cast<size_t>(n));
  std::vector<double2> h1(2 * static_cast<size_t>(n));
  HIPC(c, hipMemcpyAsync(h0.data(), c->d.y0 + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipMemcpyAsync(h0.data() + n, c->d.d0 + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipMemcpyAsync(h1.data(), c->d.y1 + off, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipMemcpyAsync(h1.data() + n, c->d.d1 + off, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    rho[4 * i + 0] = h0[s];
    rho[4 * i + 1] = 0.0;
    rho[4 * i + 2] = h1[s].x;
    rho[4 * i + 3] = h1[s].y;
    if(drho) {
      drho[4 * i + 0] = h0[n + s];
      drho[4 * i + 1] = 0.0;
      drho[4 * i + 2] = h1[n + s].x;
      drho[4 * i + 3] = h1[n + s].y;
    }
  }
  return QUEST_OK;
}

int quest_pair_delays_range(quest_ctx *c, int64_t first, int64_t count, int32_t *delays)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  const int64_t total = static_cast<int64_t>(c->n) * (c->n - 1) / 2;
  if(first < 0 || count < 0 || first > total || count > total - first)
    return fail(c, QUEST_ERR_ARG, "pair range [" + std::to_string(first) + ", +" + std::to_string(count) +
                                      ") outside [0, " + std::to_string(total) + ")");
  if(count == 0) return QUEST_OK;
  if(!delays) return fail(c, QUEST_ERR_ARG, "delays is required");
  HIPC(c, hipSetDevice(c->P.device));
  const int64_t chunk = std::min<int64_t>(count, int64_t(1) << 26);
  int32_t *buf = nullptr;
  HIPC(c, hipMalloc(&buf, sizeof(int32_t) * chunk));
  for(int64_t done = 0; done < count; done += chunk) {
    const int64_t cnt = std::min(chunk, count - done);
    hipError_t e = launch_pair_delays(c->k, c->d, c->P.precision, buf, first + done, cnt, c->stream);
    if(e == hipSuccess)
      e = hipMemcpyAsync(delays + done, buf, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, c->stream);
    if(e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if(e != hipSuccess) {
      (void)hipFree(buf);
      return fail(c, QUEST_ERR_HIP, std::string("pair delays: ") + hipGetErrorString(e));
    }
  }
  HIPC(c, hipFree(buf));
  return QUEST_OK;
}

int quest_pair_delays(quest_ctx *c, int32_t *delays)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  return quest_pair_delays_range(c, 0, static_cast<int64_t>(c->n) * (c->n - 1) / 2, delays);
}

int quest_set_timing(quest_ctx *c, int enable)
{
  if(!c) return QUEST_ERR_ARG;
  c->timing = enable != 0;
  return QUEST_OK;
}

int quest_pair_kernel_time(quest_ctx *c, double *ms, int64_t *launches, int reset)
{
  if(!c) return QUEST_ERR_ARG;
  if(ms) *ms = c->pair_ms;
  if(launches) *launches = c->pair_launches;
  if(reset) {
    c->pair_ms = 0;
    c->pair_launches = 0;
  }
  return QUEST_OK;
}

int quest_short_kernel_time(quest_ctx *c, double *ms, int64_t *launches, int reset)
{
  if(!c) return QUEST_ERR_ARG;
  if(ms) *ms = c->short_ms;
  if(launches) *launches = c->short_launches;
  if(reset) {
    c->short_ms = 0;
    c->short_launches = 0;
  }
  return QUEST_OK;
}

int quest_ctx_info(const quest_ctx *c, int64_t *info)
{
  if(!c || !info) return QUEST_ERR_ARG;
  info[0] = c->n;
  info[1] = c->k.ring_slots;
  info[2] = c->dmax_bound;
  info[3] = c->near_total;
  info[4] = c->k.n_work;
  info[5] = c->k.grid;
  info[6] = c->k.state_slots;
  info[7] = c->P.precision;
  info[8] = c->n_sys;
  info[9] = c->chunk_len;
  info[10] = c->k.max_chunks;
  info[11] = c->graph_exec != nullptr;
  info[12] = c->step_block;
  info[13] = c->max_group;
  info[14] = c->k.dyn_items;
  info[15] = c->k.pipelined;
  info[16] = c->pair_method;
  info[17] = c->k.pair_blocked ? c->k.block_steps : 0;
  info[18] = c->short_total;
  info[19] = c->k.pair_blocked ? c->k.hist_slots : 0;
  return QUEST_OK;
}

int quest_sync(quest_ctx *c)
{
  if(!c) return QUEST_ERR_ARG;
  HIPC(c, hipSetDevice(c->P.device));
  HIPC(c, hipStreamSynchronize(c->stream));
  return QUEST_OK;
}

}  // extern "C"
