// quest_capi.cpp -- the C ABI (include/quest_amd.h) over the gfx950 kernels.
//
// A context owns the device copy of one dot system in Morton order, the
// interpolant ring of the pair sum, the predictor-corrector state ring and one
// HIP stream.  Two ways to drive it, one per context:
//   per-call  quest_hist_evaluate   -- HistoryInteraction::evaluate replacement
//   whole-run quest_run             -- Integrator::solve replacement, all steps
//                                      on the device
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "quest_internal.hpp"

using namespace quest;

struct quest_ctx {
  quest_params P{};
  int n = 0;
  std::vector<int> perm;  // perm[sorted] = caller index
  std::vector<int> inv;   // inv[caller] = sorted index
  Consts k{};
  DeviceState d{};
  hipStream_t stream = nullptr;
  size_t ring_bytes = 0;
  int dmax_bound = 0;
  int near_total = 0;
  int phase_len = 0;
  int mode = 0;           // 0 unset, 1 per-call, 2 whole-run
  bool started = false;
  int cur_step = 0;       // whole-run: next step; per-call: step of the current values (-1 = none)
  int bad = -1;
  bool pc_uploaded = false;
  std::vector<double> pcw;  // ps[2W], cs[2W]
  double fc = 0;
  PulseParams pulse{};
  std::string err;
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double pair_ms = 0;
  int64_t pair_launches = 0;
  double2 *h_stage = nullptr;  // pinned [n]
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

// Morton (Z-order) permutation of the dots: 21 bits per axis over the
// bounding box, stable in the caller's order for equal keys.
std::vector<int> morton_order(int n, const double *pos)
{
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for(int i = 0; i < n; ++i)
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
  std::vector<uint64_t> key(n);
  for(int i = 0; i < n; ++i) {
    uint64_t q[3];
    for(int a = 0; a < 3; ++a) {
      const double span = hi[a] - lo[a];
      const double t = span > 0 ? (pos[3 * i + a] - lo[a]) / span : 0.0;
      q[a] = static_cast<uint64_t>(std::min(std::max(t, 0.0), 1.0) * 2097151.0);
    }
    key[i] = spread(q[0]) | spread(q[1]) << 1 | spread(q[2]) << 2;
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
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

int default_splits(int n, int tiles)
{
  int s = (2048 + tiles - 1) / tiles;
  s = std::min(s, std::max(1, n / 64));
  if(s >= 8) s = (s + 7) / 8 * 8;
  return std::max(1, s);
}

int ensure_phase(quest_ctx *c, int upto)
{
  if(upto <= c->phase_len) return QUEST_OK;
  int len = std::max(1024, next_pow2(upto));
  std::vector<double2> h(len);
  const long double theta = static_cast<long double>(c->k.omega) * static_cast<long double>(c->k.dt);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for(int s = 0; s < len; ++s) {
    const long double a = fmodl(theta * s, two_pi);
    h[s] = make_double2(static_cast<double>(cosl(a)), static_cast<double>(sinl(a)));
  }
  if(c->d.phase) HIPC(c, hipFree(c->d.phase));
  c->d.phase = nullptr;
  HIPC(c, hipMalloc(&c->d.phase, sizeof(double2) * len));
  HIPC(c, hipMemcpyAsync(c->d.phase, h.data(), sizeof(double2) * len, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->phase_len = len;
  return QUEST_OK;
}

int start_state(quest_ctx *c)
{
  HIPC(c, launch_fill_ground_state(c->k, c->d, c->stream));
  HIPC(c, launch_clear_ring(c->k, c->d, c->ring_bytes, c->stream));
  c->started = true;
  c->bad = -1;
  return QUEST_OK;
}

int ensure_pc(quest_ctx *c)
{
  if(c->pc_uploaded) return QUEST_OK;
  const int W = c->P.history_window;
  if(c->pcw.empty()) {
    c->pcw.assign(4 * W, 0.0);
    if(pc_weights(18, W, 3.15, c->pcw.data(), c->pcw.data() + 2 * W, &c->fc) != 0)
      return fail(c, QUEST_ERR_ARG, "pc weight solve failed");
  }
  HIPC(c, hipMemcpyAsync(c->d.pcw, c->pcw.data(), sizeof(double) * 4 * W, hipMemcpyHostToDevice,
                         c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->pc_uploaded = true;
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
    c->pair_ms += ms;
    c->pair_launches += 1;
  }
  c->ev_used = 0;
  return QUEST_OK;
}

int pair_sum(quest_ctx *c, int step)
{
  if(c->n < 2) return QUEST_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if(c->timing) {
    e0 = pool_event(c);
    e1 = pool_event(c);
    if(!e0 || !e1) return fail(c, QUEST_ERR_HIP, "hipEventCreate failed");
    HIPC(c, hipEventRecord(e0, c->stream));
  }
  HIPC(c, launch_pair_past(c->k, c->d, c->P.precision, step, c->stream));
  if(c->timing) HIPC(c, hipEventRecord(e1, c->stream));
  return QUEST_OK;
}

void free_device(quest_ctx *c)
{
  DeviceState &d = c->d;
  void *ptrs[] = {d.pos64, d.dip64, d.pos32, d.dip32, d.freq, d.t1, d.t2, d.ring, d.y0, d.y1,
                  d.d0, d.d1, d.cur0, d.cur1, d.yc0, d.yc1, d.field, d.partial, d.phase, d.out,
                  d.near_start, d.near_src, d.near_tgt, d.near_n, d.near_coef, d.pcw,
                  d.bad_step, d.counts, d.perm_dev};
  for(void *p : ptrs)
    if(p) (void)hipFree(p);
  d = DeviceState{};
  if(c->h_stage) (void)hipHostFree(c->h_stage);
  c->h_stage = nullptr;
  for(hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  c->ev_pool.clear();
  if(c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
}

int create(quest_ctx *c, const quest_params *p, int n, const double *pos, const double *dip,
           const double *freq, const double *t1, const double *t2)
{
  c->P = *p;
  c->n = n;
  HIPC(c, hipSetDevice(p->device));
  HIPC(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

  c->perm = morton_order(n, pos);
  c->inv.assign(n, 0);
  for(int s = 0; s < n; ++s) c->inv[c->perm[s]] = s;

  // bounding-box diagonal bounds every pair's delay
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for(int i = 0; i < n; ++i)
    for(int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], pos[3 * i + a]);
      hi[a] = std::max(hi[a], pos[3 * i + a]);
    }
  const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const double cdt = p->c0 * p->dt;
  if(!(diag / cdt < 1e6)) return fail(c, QUEST_ERR_ARG, "system extent / (c0 dt) too large");
  c->dmax_bound = static_cast<int>(std::floor(diag / cdt)) + 2;

  Consts &k = c->k;
  k.n = n;
  k.order = p->interp_order;
  k.kstride = p->interp_order + 1;  // slot_stride<M>()
  k.ring_slots = next_pow2(c->dmax_bound + 1);
  k.ring_mask = k.ring_slots - 1;
  k.tiles = (n + kBlock - 1) / kBlock;
  k.splits = p->source_split > 0 ? p->source_split : default_splits(n, k.tiles);
  k.src_per_split = (n + k.splits - 1) / k.splits;
  k.splits = (n + k.src_per_split - 1) / k.src_per_split;
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
  k.cc = p->c0 * p->c0;
  k.oo = p->laser_omega * p->laser_omega;
  k.oc = p->laser_omega * p->c0;
  k.m_cdt = -p->c0 / p->dt;
  k.m_o2dt = -2 * p->laser_omega / p->dt;
  k.two_idt2 = 2 / (p->dt * p->dt);
  lagrange_monomials(p->interp_order, k.vinv);
  // FP32 positions are stored relative to the box centre; each coordinate is
  // off by at most half an ulp of the largest centred coordinate, so a pair
  // distance by at most sqrt(3) ulp.  The arithmetic (squares, sum, rsqrt,
  // two products, the rounded 1/(c0 dt)) adds at most ~3.5e-7 relative.
  double ctr[3], half = 0;
  for(int a = 0; a < 3; ++a) {
    ctr[a] = 0.5 * (lo[a] + hi[a]);
    half = std::max(half, 0.5 * (hi[a] - lo[a]));
  }
  const double ulp = half > 0 ? std::ldexp(1.0, std::ilogb(half) - 23) : 0x1p-126;
  k.eps_rel = 8e-7;
  k.eps_abs = 2.0 * std::sqrt(3.0) * ulp / cdt + 2e-6;
  k.dyadic = p->dyadic;

  // sorted host copies
  std::vector<double4> p64(n), d64(n);
  std::vector<float4> p32(n), d32(n);
  std::vector<double> f(n, 0.0), a1(n, 1.0), a2(n, 1.0);
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    p64[s] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], 0.0);
    d64[s] = make_double4(dip[3 * i], dip[3 * i + 1], dip[3 * i + 2], 0.0);
    p32[s] = make_float4(static_cast<float>(p64[s].x - ctr[0]), static_cast<float>(p64[s].y - ctr[1]),
                         static_cast<float>(p64[s].z - ctr[2]), 0.0f);
    d32[s] = make_float4(static_cast<float>(d64[s].x), static_cast<float>(d64[s].y),
                         static_cast<float>(d64[s].z), 0.0f);
    if(freq) f[s] = freq[i];
    if(t1) a1[s] = t1[i];
    if(t2) a2[s] = t2[i];
  }

  DeviceState &d = c->d;
  const size_t nn = static_cast<size_t>(n);
  const size_t r2size = p->precision == QUEST_PREC_FP64 ? sizeof(double2) : sizeof(float2);
  c->ring_bytes = nn * k.ring_slots * k.kstride * r2size;
  const size_t ss = static_cast<size_t>(k.state_slots) * nn;
  auto alloc = [&](void **ptr, size_t bytes) -> hipError_t { return hipMalloc(ptr, std::max<size_t>(bytes, 16)); };
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pos64), sizeof(double4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.dip64), sizeof(double4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pos32), sizeof(float4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.dip32), sizeof(float4) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.freq), sizeof(double) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.t1), sizeof(double) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.t2), sizeof(double) * nn));
  {
    hipError_t e = alloc(&d.ring, c->ring_bytes);
    if(e != hipSuccess)
      return fail(c, QUEST_ERR_NOMEM, "ring allocation of " + std::to_string(c->ring_bytes) + " bytes failed");
  }
  HIPC(c, alloc(reinterpret_cast<void **>(&d.y0), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.y1), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.d0), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.d1), sizeof(double2) * ss));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.cur0), sizeof(double2) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.cur1), sizeof(double2) * 2 * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.yc0), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.yc1), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.field), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.partial), sizeof(double2) * nn * k.splits));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.out), sizeof(double2) * nn));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.near_start), sizeof(int) * (nn + 1)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.counts), sizeof(int) * (nn + 1)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.pcw), sizeof(double) * 4 * kMaxWindow));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.bad_step), sizeof(int)));
  HIPC(c, alloc(reinterpret_cast<void **>(&d.perm_dev), sizeof(int) * nn));
  HIPC(c, hipHostMalloc(reinterpret_cast<void **>(&c->h_stage), sizeof(double2) * nn * 2));

  HIPC(c, hipMemcpyAsync(d.pos64, p64.data(), sizeof(double4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.dip64, d64.data(), sizeof(double4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.pos32, p32.data(), sizeof(float4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.dip32, d32.data(), sizeof(float4) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.freq, f.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.t1, a1.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.t2, a2.data(), sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
  HIPC(c, hipMemcpyAsync(d.perm_dev, c->inv.data(), sizeof(int) * nn, hipMemcpyHostToDevice, c->stream));

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
  HIPC(c, hipStreamSynchronize(c->stream));
  int rc = ensure_phase(c, 1024);
  if(rc != QUEST_OK) return rc;
  return start_state(c);
}

int check_params(const quest_params *p, int n)
{
  if(!p) return fail(nullptr, QUEST_ERR_ARG, "params is NULL");
  if(n < 1) return fail(nullptr, QUEST_ERR_ARG, "n_dots must be >= 1");
  if(p->interp_order < 1 || p->interp_order > kMaxOrder)
    return fail(nullptr, QUEST_ERR_ARG, "interp_order must be in 1..6");
  if(p->history_window < 1 || p->history_window > kMaxWindow)
    return fail(nullptr, QUEST_ERR_ARG, "history_window must be in 1..64");
  if(p->precision != QUEST_PREC_FP64 && p->precision != QUEST_PREC_MIXED)
    return fail(nullptr, QUEST_ERR_ARG, "unknown precision");
  if(p->dyadic != QUEST_DYADIC_ROTATING && p->dyadic != QUEST_DYADIC_LAB)
    return fail(nullptr, QUEST_ERR_ARG, "unknown dyadic");
  if(!(p->dt > 0) || !(p->c0 > 0) || !(p->hbar > 0) || !std::isfinite(p->mu0) ||
     !std::isfinite(p->laser_omega))
    return fail(nullptr, QUEST_ERR_ARG, "constants must be finite, dt, c0, hbar > 0");
  if(p->source_split < 0) return fail(nullptr, QUEST_ERR_ARG, "source_split must be >= 0");
  return QUEST_OK;
}

}  // namespace

extern "C" {

int quest_ctx_create(const quest_params *params, int n_dots, const double *pos,
                     const double *dip, const double *freq, const double *t1,
                     const double *t2, quest_ctx **out)
{
  if(!out) return fail(nullptr, QUEST_ERR_ARG, "out is NULL");
  *out = nullptr;
  int rc = check_params(params, n_dots);
  if(rc != QUEST_OK) return rc;
  if(!pos || !dip) return fail(nullptr, QUEST_ERR_ARG, "pos and dip are required");
  for(int i = 0; i < 3 * n_dots; ++i)
    if(!std::isfinite(pos[i])) return fail(nullptr, QUEST_ERR_ARG, "non-finite position");
  quest_ctx *c = new(std::nothrow) quest_ctx;
  if(!c) return fail(nullptr, QUEST_ERR_NOMEM, "context allocation failed");
  rc = create(c, params, n_dots, pos, dip, freq, t1, t2);
  if(rc != QUEST_OK) {
    g_create_error = c->err;
    free_device(c);
    delete c;
    return rc;
  }
  *out = c;
  return QUEST_OK;
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
  if(c->cur_step >= 0 && step < c->cur_step)
    return fail(c, QUEST_ERR_STATE, "step went backwards");
  if(c->cur_step >= 0 && step > c->cur_step + 1)
    return fail(c, QUEST_ERR_STATE, "step skipped: every step must be evaluated");
  int rc = ensure_phase(c, step + 2);
  if(rc != QUEST_OK) return rc;
  if(step != c->cur_step) {
    if(c->cur_step >= 0) HIPC(c, launch_hist_commit(c->k, c->d, c->P.precision, c->cur_step, c->stream));
    rc = pair_sum(c, step);
    if(rc != QUEST_OK) return rc;
    HIPC(c, launch_hist_reduce(c->k, c->d, step, c->stream));
    c->cur_step = step;
  }
  const int n = c->n;
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    c->h_stage[s] = make_double2(rho01[2 * i], rho01[2 * i + 1]);
  }
  HIPC(c, hipMemcpyAsync(c->d.cur1, c->h_stage, sizeof(double2) * n, hipMemcpyHostToDevice, c->stream));
  HIPC(c, launch_hist_near(c->k, c->d, step, c->stream));
  HIPC(c, hipMemcpyAsync(c->h_stage + n, c->d.out, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    out[2 * i] = c->h_stage[n + s].x;
    out[2 * i + 1] = c->h_stage[n + s].y;
  }
  if(c->timing) return collect_timing(c);
  return QUEST_OK;
}

int quest_set_pulse(quest_ctx *c, const quest_pulse *pl)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!pl) {
    c->pulse = PulseParams{};
    return QUEST_OK;
  }
  PulseParams &q = c->pulse;
  q.on = 1;
  q.amplitude = pl->amplitude;
  q.delay = pl->delay;
  q.width = pl->width;
  q.freq = pl->freq;
  for(int a = 0; a < 3; ++a) {
    q.k[a] = pl->wavevector[a];
    q.pol[a] = pl->polarization[a];
  }
  return QUEST_OK;
}

int quest_pulse_evaluate(quest_ctx *c, int step, double *out)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!out) return fail(c, QUEST_ERR_ARG, "out is required");
  HIPC(c, hipSetDevice(c->P.device));
  HIPC(c, launch_pulse(c->k, c->d, c->pulse, step, c->stream));
  const int n = c->n;
  HIPC(c, hipMemcpyAsync(c->h_stage + n, c->d.out, sizeof(double2) * n, hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    out[2 * i] = c->h_stage[n + s].x;
    out[2 * i + 1] = c->h_stage[n + s].y;
  }
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
  c->pc_uploaded = false;
  return QUEST_OK;
}

int quest_run(quest_ctx *c, int n_steps)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(n_steps < 0) return fail(c, QUEST_ERR_ARG, "n_steps must be >= 0");
  if(c->mode == 1) return fail(c, QUEST_ERR_STATE, "context is driven by quest_hist_evaluate");
  HIPC(c, hipSetDevice(c->P.device));
  if(c->mode == 0) {
    c->mode = 2;
    c->cur_step = 0;
  }
  if(c->bad >= 0) return fail(c, QUEST_ERR_DOMAIN, "unbounded history values at or before step " + std::to_string(c->bad));
  int rc = ensure_pc(c);
  if(rc != QUEST_OK) return rc;
  rc = ensure_phase(c, c->cur_step + n_steps + 1);
  if(rc != QUEST_OK) return rc;
  const int W = c->P.history_window;
  for(int step = c->cur_step; step < c->cur_step + n_steps; ++step) {
    rc = pair_sum(c, step);
    if(rc != QUEST_OK) return rc;
    HIPC(c, launch_pc_begin(c->k, c->d, c->pulse, W, step, c->stream));
    for(int it = 0; it <= kNumCorrectorSteps; ++it)
      HIPC(c, launch_pc_iter(c->k, c->d, c->P.precision, W, c->fc, step, it, c->stream));
  }
  int bad = INT_MAX;
  HIPC(c, hipMemcpyAsync(&bad, c->d.bad_step, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  c->cur_step += n_steps;
  if(c->timing) {
    rc = collect_timing(c);
    if(rc != QUEST_OK) return rc;
  }
  if(bad != INT_MAX) {
    c->bad = bad;
    return fail(c, QUEST_ERR_DOMAIN, "unbounded history values at or before step " + std::to_string(bad));
  }
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
  std::vector<double2> h(4 * static_cast<size_t>(n));
  double2 *srcs[4] = {c->d.y0, c->d.y1, c->d.d0, c->d.d1};
  for(int q = 0; q < 4; ++q)
    HIPC(c, hipMemcpyAsync(h.data() + static_cast<size_t>(q) * n, srcs[q] + off, sizeof(double2) * n,
                           hipMemcpyDeviceToHost, c->stream));
  HIPC(c, hipStreamSynchronize(c->stream));
  for(int s = 0; s < n; ++s) {
    const int i = c->perm[s];
    rho[4 * i + 0] = h[s].x;
    rho[4 * i + 1] = h[s].y;
    rho[4 * i + 2] = h[n + s].x;
    rho[4 * i + 3] = h[n + s].y;
    if(drho) {
      drho[4 * i + 0] = h[2 * n + s].x;
      drho[4 * i + 1] = h[2 * n + s].y;
      drho[4 * i + 2] = h[3 * n + s].x;
      drho[4 * i + 3] = h[3 * n + s].y;
    }
  }
  return QUEST_OK;
}

int quest_pair_delays(quest_ctx *c, int32_t *delays)
{
  if(!c) return fail(nullptr, QUEST_ERR_ARG, "ctx is NULL");
  if(!delays) return fail(c, QUEST_ERR_ARG, "delays is required");
  HIPC(c, hipSetDevice(c->P.device));
  const int64_t total = static_cast<int64_t>(c->n) * (c->n - 1) / 2;
  if(total == 0) return QUEST_OK;
  const int64_t chunk = std::min<int64_t>(total, int64_t(1) << 26);
  int32_t *buf = nullptr;
  HIPC(c, hipMalloc(&buf, sizeof(int32_t) * chunk));
  for(int64_t first = 0; first < total; first += chunk) {
    const int64_t cnt = std::min(chunk, total - first);
    hipError_t e = launch_pair_delays(c->k, c->d, buf, first, cnt, c->stream);
    if(e == hipSuccess)
      e = hipMemcpyAsync(delays + first, buf, sizeof(int32_t) * cnt, hipMemcpyDeviceToHost, c->stream);
    if(e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if(e != hipSuccess) {
      (void)hipFree(buf);
      return fail(c, QUEST_ERR_HIP, std::string("pair delays: ") + hipGetErrorString(e));
    }
  }
  HIPC(c, hipFree(buf));
  return QUEST_OK;
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

int quest_ctx_info(const quest_ctx *c, int64_t *info)
{
  if(!c || !info) return QUEST_ERR_ARG;
  info[0] = c->n;
  info[1] = c->k.ring_slots;
  info[2] = c->dmax_bound;
  info[3] = c->near_total;
  info[4] = c->k.splits;
  info[5] = c->k.tiles;
  info[6] = c->k.state_slots;
  info[7] = c->P.precision;
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
