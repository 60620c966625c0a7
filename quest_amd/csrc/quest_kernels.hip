// quest_kernels.hip -- gfx950 kernels of the retarded-interaction step.
//
// Data layout in HBM (all in the library's Morton-sorted dot order):
//   pos/dip        (x, y, z, 0) per dot in um (FP64), and interleaved
//                  (position relative to the centre in units of c0 dt,
//                  dipole) copies in FP64 and FP32 for the pair sums
//   ring           for every source dot and every finished step s (slot
//                  s & ring_mask) the monomial coefficients c_k (in x - 1/2)
//                  of its Lagrange interpolant of rho01 through steps s, s-1,
//                  ..., s-M, multiplied by e^{i theta s} e^{-i theta / 2}
//                  (rotating frame), in FP64 or FP32, in coefficient planes
//                  (see ring_index)
//   state ring     double2 [state_slots][n] rho00, rho01 and derivatives, for
//                  the predictor-corrector window
//
//   sample ring    (time-blocked method) rho01 itself per source and finished
//                  step, rows with a mirrored tail, so that a window of K + M
//                  steps is contiguous
//
// The pair sum is a gather: one thread per target dot, sources walked in
// order.  Source data are wave-uniform (scalar loads); the ring read is a
// per-lane gather of M + 1 complex values (per step), or of K + M samples
// (time-blocked, per 8 steps).  No MFMA: the work is a gather-and-accumulate,
// ~78 VALU operations per ordered pair and step (per step), ~37 (time-blocked).
#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <type_traits>

#include "quest_internal.hpp"

namespace quest {

// Partial sums of the pair kernel for quest_run_noalloc: one static array per
// GPU, loaded with the code object, so that entry point allocates nothing.
// Host code never passes `scratch` itself to a kernel (on the host the name
// is only a stand-in for the symbol); it uses the device address
// scratch_address() looked up per device.  Nothing in it has to start at
// zero: every step's pair kernel writes every partial row it reads.
static __device__ __align__(256) double2 scratch[kScratchBytes / sizeof(double2)];

hipError_t scratch_address(int device, double2 **ptr)
{
  static std::mutex mu;
  static double2 *addr[kMaxDevices] = {};
  *ptr = nullptr;
  if(device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if(!addr[device]) {
    // the symbol's address on the current device (the caller made `device` current)
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if(e != hipSuccess) return e;
    addr[device] = static_cast<double2 *>(p);
  }
  *ptr = addr[device];
  return hipSuccess;
}

namespace {

// ------------------------------------------------------------------ helpers

template <typename R> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };
template <typename R> struct V4;
template <> struct V4<double> { using type = double4; };
template <> struct V4<float> { using type = float4; };

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b)
{
  return make_double2(a.x + b.x, a.y + b.y);
}

// r^2 (units of (c0 dt)^2) and 1/r of the FP32 pair arithmetic (positions
// stored in units of c0 dt, so r is the delay tau).  Written with explicit
// fused operations and contraction off so that every kernel that classifies a
// pair (the pair sum and the zero-delay list builder) gets bit-identical
// results.  r^2 is
// clamped to 1e-20 (pairs that close have delay 0), so the geometry of a dot
// with itself or a coincident one stays finite (1/r <= 1e10) and its term is
// an exact 0 against the zeroed slot it reads.  (A max, not a seed added in
// the first FMA: FP32 FMAs issue at half the rate of muls and maxes here.)
__device__ __forceinline__ void r2_f32(float dx, float dy, float dz, float &r2, float &ir)
{
#pragma clang fp contract(off)
  r2 = __builtin_fmaxf(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), 1e-20f);
  ir = __builtin_amdgcn_rsqf(r2);
}

// FP64 delay split, bit-identical to the reference's
//   n = trunc(sqrt((dx^2 + dy^2) + dz^2) / (c0 dt))   (no fused operations)
// (history_interaction.cpp:31-32, split_double :88-97).  A fast estimate from
// the hardware reciprocal square root (two Newton steps, relative error a few
// ulp) decides the integer part unless the estimate lies within 1e-9 of an
// integer; then the correctly rounded sqrt and division are evaluated.
// Returns n, the fraction x and 1/r.
__device__ __forceinline__ int split_f64(double dx, double dy, double dz, double cdt,
                                         double icdt, double &x, double &ir, double &r2)
{
#pragma clang fp contract(off)
  r2 = (dx * dx + dy * dy) + dz * dz;
  // estimate from r^2 >= 1e-20 um^2: pairs closer than 1e-10 um (delay 0;
  // the self pair) get a finite 1/r <= 2.25e10 and a finite fraction, so their
  // pair-kernel terms are an exact 0 against the zeroed slot they read
  double y = __builtin_amdgcn_rsq(__builtin_fmax(r2, 1e-20));
  {
    const double h = r2 * y;
    const double e = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
  }
  {
    const double h = r2 * y;
    const double e = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
  }
  ir = y;
  const double tau = (r2 * y) * icdt;
  double fl = __builtin_floor(tau);
  double fr = tau - fl;
  if(__builtin_expect(fr < 1e-9 || fr > 1.0 - 1e-9, 0)) {
    const double r = __builtin_sqrt(r2);
    const double te = r / cdt;
    fl = __builtin_floor(te);
    fr = te - fl;
  }
  x = fr;
  // NaN (coincident dots) -> 0; keeps the conversion defined
  return static_cast<int>(__builtin_fmax(fl, 0.0));
}

// Delay split of the FP64 pair sum, whose positions are in units of L = c0 dt
// (dx, dy, dz: a pair's separation in those units, so its length is the
// delay tau).  The estimate tau = r2 rsq(r2) (hardware reciprocal square root,
// two Newton steps; within ~1e-13 of r / (c0 dt) including the rounding of
// the scaled positions) decides the integer part unless it lies within 1e-9
// of an integer; then the reference's own split of the um positions
// (split_f64's exact branch: correctly rounded sqrt of the unfused sum, then
// the division) decides.  So n is the reference's trunc(r / (c0 dt))
// (history_interaction.cpp:31-32).  Returns n, the centred fraction
// xc = x - 1/2 and 1/r in units of 1/L.  r2 >= 0, so the floor is never
// negative; r2 is clamped to 1e-20 (pairs that close have delay 0: the self
// pair) so that their geometry stays finite.
__device__ __forceinline__ int split_l64(double dx, double dy, double dz, const double4 &pa64,
                                         const double4 *__restrict__ pos64, int b, double cdt, double &xc,
                                         double &ir)
{
  const double r2 = __builtin_fmax(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx)), 1e-20);
  double y = __builtin_amdgcn_rsq(r2);
  {
    const double e = __builtin_fma(-(r2 * y), y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
  }
  {
    const double e = __builtin_fma(-(r2 * y), y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
  }
  ir = y;
  const double tau = r2 * y;
  double fl = __builtin_floor(tau);
  double fr = tau - fl;
  if(__builtin_expect(fr < 1e-9 || fr > 1.0 - 1e-9, 0)) {
#pragma clang fp contract(off)
    const double4 pb = pos64[b];
    const double ex = pa64.x - pb.x, ey = pa64.y - pb.y, ez = pa64.z - pb.z;
    const double te = __builtin_sqrt((ex * ex + ey * ey) + ez * ez) / cdt;
    fl = __builtin_floor(te);
    fr = te - fl;
  }
  xc = fr - 0.5;
  return static_cast<int>(fl);
}

// Delay split of the mixed-precision path: the FP32 estimate decides the
// integer part unless its fraction lies within the estimate's error bound
// eps = eps_rel tau + eps_abs (position rounding plus five roundings, with a
// factor two of margin) of an integer -- tested as |fr - 1/2| > 1/2 - eps --
// and then the exact FP64 split decides.  The integer delays are therefore the
// reference's in both precisions, and the pair kernel and the zero-delay list
// builder classify every pair alike.  The estimate's integer part is
// n = rint(tau - 1/2), which is floor(tau) whenever the test passes, taken
// with a rounding constant ms = kRoundMagic + sm (0 <= sm < 2^22): no
// float-to-int conversion, and the bit pattern of ms - n, returned, is
// kRoundMagicBits + sm - n, so with sm = step & mask its low bits are the
// ring slot (step - n) & mask.  lim0 = 1/2 - eps_abs - eps_rel / 2.  Also
// returns the fraction centred on 1/2, xc = x - 1/2 (the FP32 interpolant's
// variable).
//
// 1.5 * 2^23: for 0 <= sm < 2^22 and |v| < 2^22 the float
// (kRoundMagic + sm) - v is exactly kRoundMagic + sm - rint(v) (a float in
// [2^23, 2^24), unit in the last place 1), whose bit pattern is
// kRoundMagicBits + sm - rint(v).
constexpr float kRoundMagic = 12582912.0f;
constexpr int kRoundMagicBits = 0x4B400000;

__device__ __forceinline__ int split_mixed(float dx, float dy, float dz, float eps_rel, float lim0, float ms,
                                           const double4 &pa64, const double4 *__restrict__ pos64, int b,
                                           double cdt, double icdt64, float &xc, float &ir)
{
  float r2;
  r2_f32(dx, dy, dz, r2, ir);
  const float u = __builtin_fmaf(r2, ir, -0.5f);  // tau - 1/2, tau = r / (c0 dt)
  const float t = ms - u;                         // ms - n, n = rint(tau - 1/2)
  xc = u - (ms - t);                              // exact: tau - n - 1/2 in [-1/2, 1/2]
  const float lim = __builtin_fmaf(-eps_rel, u, lim0);  // 1/2 - eps_rel tau - eps_abs
  int code = __builtin_bit_cast(int, t);
  if(__builtin_expect(__builtin_fabsf(xc) > lim, 0)) {
    const double4 pb = pos64[b];
    double xd, ird, r2d;
    const int n = split_f64(pa64.x - pb.x, pa64.y - pb.y, pa64.z - pb.z, cdt, icdt64, xd, ird, r2d);
    xc = static_cast<float>(xd - 0.5);
    code = __builtin_bit_cast(int, ms) - n;
  }
  return code;  // bit pattern of ms - n
}

// sin and cos of theta x for x in [-1/2, 1/2), used while theta <= 0.28
// (|theta x| <= 0.14): series in x^2 with theta folded into the coefficients;
// sn[] = the sine's two Chebyshev-economized terms (error <= (theta/2)^5/1920,
// 1e-8 at theta = 0.228), cs[] = -theta^2/2, theta^4/24 (error <= (theta/2)^6/720,
// 3e-9 at theta = 0.228): below FP32 rounding of the term.
__device__ __forceinline__ void sincos_theta(const float *sn, const float *cs, float x, float &s,
                                             float &c)
{
  const float x2 = x * x;
  s = x * __builtin_fmaf(x2, sn[1], sn[0]);
  c = __builtin_fmaf(x2, __builtin_fmaf(x2, cs[1], cs[0]), 1.0f);
}
// sin and cos of theta x for |x| <= 1/2, used while theta <= 0.33: Taylor
// series in x^2 with theta folded into the coefficients (Consts::trig64,
// kernel arguments, so each FMA takes its coefficient from a scalar register),
// truncated where the next term is below FP64 rounding (|theta x| <= 0.165:
// sine to u^9, cosine to u^10).
__device__ __forceinline__ void sincos_centred64(const double *tg, double x, double &s, double &c)
{
  const double x2 = x * x;
  double ps = tg[0];
#pragma unroll
  for(int k = 1; k < 5; ++k) ps = __builtin_fma(ps, x2, tg[k]);
  s = x * ps;
  double pc = tg[5];
#pragma unroll
  for(int k = 6; k < 10; ++k) pc = __builtin_fma(pc, x2, tg[k]);
  c = __builtin_fma(pc, x2, 1.0);
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Interpolant ring.  Entry e = a * ring_slots + slot holds the M + 1
// coefficients (monomials in xc = x - 1/2) of source a for one slot;
// E = n * ring_slots entries, in coefficient planes:
//   FP64: double2 [M + 1][E]
//   FP32: (M + 1) / 2 planes of two coefficients float4 [(M + 1) / 2][E],
//         then, for even M, one plane of c_M float2 [E]
// A wave's lanes read one source at the ~30 consecutive slots its targets'
// delays span, so each load instruction touches a short contiguous run of
// 16-B entries (~4 cache lines) instead of ~20 lines of whole records.
// Measured against 80-B records: FP64 -18 % at N = 131072 and within 1 % at
// N = 65536; FP32 (40-B records) -3.5 % at N = 65536.
// Index of coefficient kk of entry e, in units of the complex type.
__device__ __forceinline__ size_t ring_index(const double2 *, int /*order*/, size_t E, size_t e, int kk)
{
  return static_cast<size_t>(kk) * E + e;
}
__device__ __forceinline__ size_t ring_index(const float2 *, int order, size_t E, size_t e, int kk)
{
  const int pairs = (order + 1) / 2;
  if(kk < 2 * pairs) return (static_cast<size_t>(kk >> 1) * E + e) * 2 + (kk & 1);
  return static_cast<size_t>(2 * pairs) * E + e;
}

// Load at a wave-uniform base plus a 32-bit per-lane byte offset (the
// base + zero-extended offset form of the global loads: no per-lane 64-bit
// address arithmetic).
template <typename T>
__device__ __forceinline__ T load_off(const T *__restrict__ base, unsigned byte_off)
{
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Value q0 = p(x), slope q1 = p'(x) and half curvature q2 = p''(x)/2 of the
// degree-M polynomial sum_k c_k x^k (simultaneous Horner).
template <typename R, int M>
__device__ __forceinline__ void horner3(const typename V2<R>::type *c, R x, R &q0r, R &q0i,
                                        R &q1r, R &q1i, R &q2r, R &q2i)
{
  q0r = c[M].x;
  q0i = c[M].y;
  q1r = 0;
  q1i = 0;
  q2r = 0;
  q2i = 0;
  if constexpr(M >= 1) {
    q1r = q0r;
    q1i = q0i;
    q0r = fmaT(q0r, x, c[M - 1].x);
    q0i = fmaT(q0i, x, c[M - 1].y);
  }
  if constexpr(M >= 2) {
    q2r = q1r;
    q2i = q1i;
    q1r = fmaT(q1r, x, q0r);
    q1i = fmaT(q1i, x, q0i);
    q0r = fmaT(q0r, x, c[M - 2].x);
    q0i = fmaT(q0i, x, c[M - 2].y);
  }
#pragma unroll
  for(int k = M - 3; k >= 0; --k) {
    q2r = fmaT(q2r, x, q1r);
    q2i = fmaT(q2i, x, q1i);
    q1r = fmaT(q1r, x, q0r);
    q1i = fmaT(q1i, x, q0i);
    q0r = fmaT(q0r, x, c[k].x);
    q0i = fmaT(q0i, x, c[k].y);
  }
}

// Pair constants converted to the pair arithmetic's type.  Both pair sums
// work in units of L = c0 dt (1/r, g3 and g1 are L, L^2 and L times their um
// values) and divide every term by unit_scale = two_idt2 / L, so that the
// constants of the pair term reduce to theta and theta^2 / 2 (pair_term).
template <typename R> struct PairK {
  R oo, theta, eps_rel, lim0;
  R sn[2], cs[2];  // sincos_theta coefficients (FP32)
  double tg[10];   // sincos_centred64 coefficients (FP64)
  double cdt, icdt64;
  __host__ __device__ explicit PairK(const Consts &k)
      : oo((R)k.oo_l), theta((R)k.theta), eps_rel((R)k.eps_rel), lim0((R)(0.5 - k.eps_abs - 0.5 * k.eps_rel)),
        sn{(R)k.sin_c[0], (R)k.sin_c[1]}, cs{(R)k.cos_c[0], (R)k.cos_c[1]}, cdt(k.cdt), icdt64(k.icdt)
  {
    for(int i = 0; i < 10; ++i) tg[i] = k.trig64[i];
  }
};

// The field at a target from one source, given the geometry and the source's
// interpolant at the retarded time (before the global factor -K e^{-i theta step}):
//   G . = e^{-i theta x} [ T0 p + T1 (p' + i w p) + T2 (p'' + 2 i w p' - w^2 p) ]
// with time derivatives p' = -q1/dt, p'' = 2 q2/dt^2, T0 = c^2 g3/r,
// T1 = c g3, T2 = g1, g3 = (d_a.d_b - 3 B)/r^2, g1 = (d_a.d_b - B)/r,
// B = (d_a.dr)(d_b.dr)/r^2 (rotating_green_function.cpp:13-20 projected on the
// dipoles, history_interaction.cpp:40-43).  LAB drops the rotation and the
// i w mixing (green_function.cpp:12-17).  Grouped by the two geometric
// factors, the bracket is g3 G3 + g1 G1 with
//   G3 = (c^2/r) q0 + i (w c) q0 - (c/dt) q1,
//   G1 = (2/dt^2) q2 - w^2 q0 - i (2w/dt) q1,
// and the rotated term is added to (accr, acci) in the same fused operations.
// In the pair sums' units (lengths in L = c0 dt, the whole term divided by
// unit_scale = 2 / (dt^2 L)) the constants of G3 and G1 reduce to
//   c^2/r -> (1/2) ir,  w c -> theta / 2,  c/dt -> 1/2,
//   w^2 -> theta^2 / 2,  2w/dt -> theta,  2/dt^2 -> 1      (theta = w dt),
// so with g3h = g3 / 2
//   g3 G3 = g3h [ir q0 - q1 + i theta q0],  G1 = q2 - (theta^2/2) q0 - i theta q1.
// g3h and g1 come from the halved target dipole dah = d_a / 2 (an exact
// power-of-two scale): aab/2 = dah.d_b, g3h = (aab/2 - (3/2) B) / r^2 and
// g1 = (2 (aab/2) - B) / r, the same value as (aab - B) / r.  x is the
// fraction centred on 1/2 (the ring's basis; e^{-i theta / 2} is folded into
// the ring).  GEN: e^{-i theta x} from the library sincos (theta beyond the
// short series' range); a separate kernel variant, so the common one carries
// none of it.
// Kahan's compensated addition: s + c carries the sum of the added terms
// with the rounding error of every addition kept in c (-c is the error).
template <typename R> __device__ __forceinline__ void kahan_add(R &s, R &c, R x)
{
  const R y = x - c;
  const R t = s + y;
  c = (t - s) - y;
  s = t;
}

// COMP: the term is formed on its own and added to (accr, acci) by
// kahan_add with compensations (cmpr, cmpi); else it is added in the same
// fused operations that rotate it.
template <typename R, bool LAB, bool GEN = false, bool COMP = false>
__device__ __forceinline__ void pair_term(const PairK<R> &K, R ir, R dx, R dy, R dz, const typename V4<R>::type &da,
                                          const typename V4<R>::type &dah, const typename V4<R>::type &db, R x,
                                          R q0r, R q0i, R q1r, R q1i, R q2r, R q2i, R &accr, R &acci, R &cmpr,
                                          R &cmpi)
{
  const R ir2 = ir * ir;
  const R ba = da.x * dx + da.y * dy + da.z * dz;
  const R bb = db.x * dx + db.y * dy + db.z * dz;
  const R aabh = dah.x * db.x + dah.y * db.y + dah.z * db.z;
  const R bq = ba * bb * ir2;
  const R g3h = fmaT(R(-1.5), bq, aabh) * ir2;
  const R g1 = fmaT(R(2), aabh, -bq) * ir;
  if constexpr(LAB) {
    const R g3r = fmaT(ir, q0r, -q1r);
    const R g3i = fmaT(ir, q0i, -q1i);
    if constexpr(COMP) {
      kahan_add(accr, cmpr, fmaT(g3h, g3r, g1 * q2r));
      kahan_add(acci, cmpi, fmaT(g3h, g3i, g1 * q2i));
    } else {
      accr = fmaT(g3h, g3r, fmaT(g1, q2r, accr));
      acci = fmaT(g3h, g3i, fmaT(g1, q2i, acci));
    }
  } else {
    const R th = K.theta;
    const R g3r = fmaT(ir, q0r, fmaT(-th, q0i, -q1r));
    const R g3i = fmaT(ir, q0i, fmaT(th, q0r, -q1i));
    const R g1r = fmaT(-K.oo, q0r, fmaT(th, q1i, q2r));
    const R g1i = fmaT(-K.oo, q0i, fmaT(-th, q1r, q2i));
    const R cr = g3h * g3r + g1 * g1r;
    const R ci = g3h * g3i + g1 * g1i;
    R sn, cs;
    if constexpr(sizeof(R) == sizeof(double)) {
      if constexpr(GEN)
        sincos(K.theta * x, &sn, &cs);
      else
        sincos_centred64(K.tg, x, sn, cs);
    } else {
      if constexpr(GEN)
        __sincosf(K.theta * x, &sn, &cs);
      else
        sincos_theta(K.sn, K.cs, x, sn, cs);
    }
    if constexpr(COMP) {
      kahan_add(accr, cmpr, fmaT(cr, cs, ci * sn));
      kahan_add(acci, cmpi, fmaT(ci, cs, -(cr * sn)));
    } else {
      accr = fmaT(cr, cs, fmaT(ci, sn, accr));
      acci = fmaT(ci, cs, fmaT(-cr, sn, acci));
    }
  }
}

// ------------------------------------------------------------ pair kernels

// Past-only pair sum: every ordered pair (a <- b) of one system, b != a,
// whose integer delay n >= 1 reads finished steps only, so this sum is formed
// once per step; the zero-delay pairs are evaluated at every corrector
// iteration from the near list.
//
// Persistent kernel over a work list: item w = (target tile of one wave of
// dots, source range) of one system; wave j of block i takes items
// 4 i + j, 4 (i + grid) + j, ...  A block's four consecutive items share a
// source chunk (for systems of >= 256 dots: one 256-target tile), the chunks
// of a group of tiles are consecutive, and `grid` is a multiple of 8, so with
// 8 chunks per tile group each XCD (round-robin block dispatch) walks one
// source range and its L2 holds that range's ring rows.
__device__ __forceinline__ int first_item()
{
  return static_cast<int>(blockIdx.x) * (kBlock / kWave) +
         __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / kWave);
}
__device__ __forceinline__ int items_stride() { return static_cast<int>(gridDim.x) * (kBlock / kWave); }

// Work items of the persistent pair kernels, one wave per item.  Static
// (DYN false): wave j of block i takes items 4i + j, 4i + j + 4 grid, ...
// (XCD-aware with the list's chunk grouping).  Dynamic: every wave starts
// with its static first item, then claims item W + c (W the grid's waves, c
// from a device counter), so uneven items (the systems of a batch) balance
// without the whole grid's first claims queueing on the counter at once.
// Each wave that takes an item makes exactly one failing claim, so the
// claims number n_work in all: the one that draws n_work - 1 is the last and
// resets the counter for the next launch.
// ctr: [0] claims drawn (0 between launches).
template <bool DYN>
struct ItemCursor {
  static constexpr bool dyn = DYN;
  int *ctr;
  int n_work;
  int w;
  __device__ ItemCursor(int *c, int n) : ctr(c), n_work(n), w(0) {}
  __device__ int claim()
  {
    int t = 0;
    if((threadIdx.x & (kWave - 1)) == 0) {
      t = atomicAdd(ctr, 1);
      if(t == n_work - 1) atomicExch(ctr, 0);
    }
    return items_stride() + __builtin_amdgcn_readfirstlane(t);
  }
  __device__ int first()
  {
    w = first_item();
    return w;
  }
  __device__ int next()
  {
    w = dyn ? claim() : w + items_stride();
    return w;
  }
};

constexpr int kChunk = 64;  // sources summed in the pair type before the FP64 fold

template <int M, bool LAB, bool DYN, bool GEN>
__global__ __launch_bounds__(kBlock) void pair_past_f64(
    const double4 *__restrict__ pd, const double4 *__restrict__ pos64, const double2 *__restrict__ ring,
    const int4 *__restrict__ work, const int *__restrict__ work_row, int *step_ptr, double2 *__restrict__ partial,
    Consts k)
{
  const PairK<double> K(k);
  const int step = *step_ptr;
  const size_t E = static_cast<size_t>(k.n) * k.ring_slots;  // entries per coefficient plane
  ItemCursor<DYN> items(step_ptr + 1, k.n_work);
  for(int w = items.first(); w < k.n_work; w = items.next()) {
    const int4 wi = work[w];
    const int tgt = wi.x + static_cast<int>(threadIdx.x % kWave);
    const int a = tgt < wi.y ? tgt : wi.y - 1;
    const double4 pa = pd[2 * a], da = pd[2 * a + 1];
    const double4 dah = make_double4(0.5 * da.x, 0.5 * da.y, 0.5 * da.z, 0.0);
    const double4 pa64 = pos64[a];
    double accr = 0, acci = 0;
    for(int b = wi.z; b < wi.w; ++b) {
      const double4 pb = pd[2 * b], db = pd[2 * b + 1];
      const double dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
      double xc, ir;
      // delay 0 (the self pair, zero-delay partners) reads the current step's
      // slot, which holds zeros until the step commits: its term is 0
      const int n = split_l64(dx, dy, dz, pa64, pos64, b, K.cdt, xc, ir);
      const unsigned off = (static_cast<unsigned>(step - n) & k.ring_mask) << 4;  // slot * 16 B
      const double2 *row = ring + static_cast<size_t>(b) * k.ring_slots;       // source b in plane 0
      double2 cv[M + 1];
#pragma unroll
      for(int i = 0; i <= M; ++i) cv[i] = load_off(row + static_cast<size_t>(i) * E, off);
      double q0r, q0i, q1r, q1i, q2r, q2i;
      horner3<double, M>(cv, xc, q0r, q0i, q1r, q1i, q2r, q2i);
      pair_term<double, LAB, GEN>(K, ir, dx, dy, dz, da, dah, db, xc, q0r, q0i, q1r, q1i, q2r, q2i, accr, acci,
                                  accr, acci);
    }
    if(tgt < wi.y)
      partial[static_cast<size_t>(work_row[w]) * k.n + tgt] = make_double2(accr * k.unit_scale, acci * k.unit_scale);
  }
}

// SGPR budget: at 82 SGPRs (what the compiler picks for this loop nest) the
// kernel ran 1.47x slower on MI355X than at 62, with identical instruction
// counts and cache traffic but ~40% fewer wave-cycles in flight, although the
// compiler reports 8 waves/SIMD for both: fewer waves were resident than
// that.  Capping at 64 costs a handful of SGPR spills to VGPR lanes outside
// the inner loop and keeps all 8 waves/SIMD resident.
// One source of the FP32 pair loop, staged ahead of its arithmetic: dipole,
// separation, centred fraction x, 1/r and the interpolant coefficients.
template <int M> struct StagedF32 {
  float4 db;
  float dx, dy, dz, x, ir;
  float2 cv[M + 1];
};

// Base of each coefficient plane of the FP32 ring, passed as separate kernel
// arguments so that each stays a scalar base of its loads (computed in the
// kernel, ring + E would be folded into the per-lane 64-bit address).
struct RingPlanesF32 {
  const float4 *pair[3];  // planes of (c_2i, c_2i+1)
  const float2 *last;     // plane of c_M (even M)
};

template <int M, bool LAB, bool DYN, bool GEN, bool PF, bool COMP>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_sgpr(64))) void pair_past_f32(
    const float4 *__restrict__ pd, const double4 *__restrict__ pos64, const RingPlanesF32 planes,
    const int4 *__restrict__ work, const int *__restrict__ work_row, int *step_ptr,
    double2 *__restrict__ partial, Consts k)
{
  const PairK<float> K(k);
  const int step = *step_ptr;
  constexpr int kPairs = (M + 1) / 2;
  const unsigned mask = static_cast<unsigned>(k.ring_mask);
  const float ms = kRoundMagic + static_cast<float>(step & k.ring_mask);  // slot = code & mask
  // a lane's 32-bit byte offset into a 16-B plane is (b ring_slots + slot) * 16
  // (the host keeps a plane below 4 GiB), formed as (slot * 16) | (b ring_slots * 16)
  // since slot < ring_slots, a power of two; the 8-B plane's offset is half of it
  const unsigned row_step16 = static_cast<unsigned>(k.ring_slots) << 4;
  ItemCursor<DYN> items(step_ptr + 1, k.n_work);
  for(int w = items.first(); w < k.n_work; w = items.next()) {
    const int4 wi = work[w];
    const int tgt = wi.x + static_cast<int>(threadIdx.x % kWave);
    const int a = tgt < wi.y ? tgt : wi.y - 1;
    const float4 pa = pd[2 * a], da = pd[2 * a + 1];
    const float4 dah = make_float4(0.5f * da.x, 0.5f * da.y, 0.5f * da.z, 0.0f);
    const double4 pa64 = pos64[a];
    double accr = 0, acci = 0;
    for(int bt = wi.z; bt < wi.w; bt += kChunk) {
      const int be = min(wi.w, bt + kChunk);
      float sr = 0, si = 0;
      float cr = 0, ci = 0;  // Kahan compensations (COMP)
      // A source is staged (position and dipole: one scalar load of 32 B;
      // delay split; coefficient gathers) and then evaluated.  PF: software-
      // pipelined by one source -- source b + 1 is staged before source b's
      // arithmetic, so the gathers' cache latency overlaps it (the last step
      // re-stages b, in bounds and unused); the same arithmetic in the same
      // order as the plain loop.  Byte offsets of the slot (step - n) & mask
      // in the 16-B and 8-B planes; delay 0 (the self pair, zero-delay
      // partners) reads the current step's slot, which holds zeros until the
      // step commits.
      auto stage = [&](int b, StagedF32<M> &st) {
        const float4 pb = pd[2 * b];
        st.db = pd[2 * b + 1];
        st.dx = pa.x - pb.x;
        st.dy = pa.y - pb.y;
        st.dz = pa.z - pb.z;
        const int code =
            split_mixed(st.dx, st.dy, st.dz, K.eps_rel, K.lim0, ms, pa64, pos64, b, K.cdt, K.icdt64, st.x, st.ir);
        const unsigned off16 = ((static_cast<unsigned>(code) & mask) << 4) | (static_cast<unsigned>(b) * row_step16);
#pragma unroll
        for(int i = 0; i < kPairs; ++i) {
          const float4 v = load_off(planes.pair[i], off16);
          st.cv[2 * i] = make_float2(v.x, v.y);
          st.cv[2 * i + 1] = make_float2(v.z, v.w);
        }
        if constexpr(2 * kPairs == M) st.cv[M] = load_off(planes.last, off16 >> 1);
      };
      auto evaluate = [&](const StagedF32<M> &st) {
        float q0r, q0i, q1r, q1i, q2r, q2i;
        horner3<float, M>(st.cv, st.x, q0r, q0i, q1r, q1i, q2r, q2i);
        pair_term<float, LAB, GEN, COMP>(K, st.ir, st.dx, st.dy, st.dz, da, dah, st.db, st.x, q0r, q0i, q1r, q1i,
                                         q2r, q2i, sr, si, cr, ci);
      };
      StagedF32<M> cur;
      if constexpr(PF) {
        stage(bt, cur);
        for(int b = bt; b < be; ++b) {
          StagedF32<M> nxt;
          stage(min(b + 1, be - 1), nxt);
          evaluate(cur);
          cur = nxt;
        }
      } else {
        for(int b = bt; b < be; ++b) {
          stage(b, cur);
          evaluate(cur);
        }
      }
      if constexpr(COMP) {  // the compensated block sum is s - c
        accr += static_cast<double>(sr) - static_cast<double>(cr);
        acci += static_cast<double>(si) - static_cast<double>(ci);
      } else {
        accr += static_cast<double>(sr);
        acci += static_cast<double>(si);
      }
    }
    if(tgt < wi.y)
      partial[static_cast<size_t>(work_row[w]) * k.n + tgt] = make_double2(accr * k.unit_scale, acci * k.unit_scale);
  }
}

// ------------------------------------------------- time-blocked pair sums
//
// The reference's sum for target a at step t (history_interaction.cpp:51-67),
//   Omega_a(t) = sum_b sum_m coef_ab[m] rho01_b(t - n_ab - m),
// reads, for a pair of integer delay n, steps up to t - n only.  At the start
// of step t0 (steps < t0 finished) every pair with n >= K therefore holds all
// it needs for the K steps t0 .. t0 + K - 1: its weights (geometry, Green's
// function and Lagrange weights, the expensive part) are formed once and
// applied to K consecutive windows of the source's rho01 samples (K = 8 or
// 16).  The pairs with 1 <= n < 4 (the short list, ~4 per dot at 512 um^-3)
// are summed step by step, those with 4 <= n < 8 (~24 per dot) in 4-step
// windows at every multiple of 4, with K = 16 those with 8 <= n < 16 (~207
// per dot) in 8-step windows at every multiple of 8 (pair_list: a pair of
// delay n >= w has its samples for w steps), and the zero-delay pairs stay on
// the near list.
//
// Sample ring: rho01 of source b at finished step u in row b, entry
// u & hist_mask, and again at hist_slots + (u & hist_mask) when that is below
// K + M - 1, so that a window of K + M consecutive steps is one contiguous
// run of the row whatever its start (2 Re rho01 in the lab frame, the
// reference's polarization_prefactor, green_function.h:23-27).
//
// Per pair, with x the delay's fraction and p(x) = sum_m L_m(x) rho01(s - m)
// the reference's interpolant (lagrange_set.cpp:17-41), the pair term of
// pair_term is  A p + B p' + C p''/2  with complex A, B, C fixed by the
// geometry (rotation e^{-i theta (n + x)} included).  So its weight on
// sample s - m is W_m = A L_m(x) + B L_m'(x) + C L_m''(x)/2, formed from the
// monomial coefficients of the L_m in z = x - M/2 (nodes centred on 0) as
// W_m = sum_j l[j][m] Gamma_j with
// Gamma_j = A z^j + j B z^(j-1) + j(j-1)/2 C z^(j-2).

// Monomial coefficients of the Lagrange basis on the centred nodes
// m - M/2, m = 0..M: L_m(z) = sum_j c[j][m] z^j with z = x - M/2, exact
// dyadic products over one integer denominator, rounded once.  The nodes are
// symmetric about 0, so c[j][M - m] = (-1)^j c[j][m]: a weight and its mirror
// share their even and odd parts (block_weights).
template <int M> struct LagrangeMonomials {
  double c[M + 1][M + 1];
  constexpr LagrangeMonomials() : c{}
  {
    for(int m = 0; m <= M; ++m) {
      double p[M + 1] = {};
      p[0] = 1.0;
      double den = 1.0;
      int deg = 0;
      for(int j = 0; j <= M; ++j) {
        if(j == m) continue;
        const double node = j - 0.5 * M;
        for(int e = deg + 1; e >= 1; --e) p[e] = p[e - 1] - node * p[e];  // times (z - node)
        p[0] = -node * p[0];
        ++deg;
        den *= m - j;
      }
      for(int j = 0; j <= M; ++j) c[j][m] = p[j] / den;
    }
  }
};

// Weights W_m = wr[m] + i wi[m] of one pair on samples s - m, m = 0..M,
// before the global factor -K (rotating) or K (lab) and the unit scale.
// live false: all weights 0 (the pair belongs to another list).  rot_n =
// e^{-i theta (n + 1/2)} (rotating frame), x = xc + 1/2 on nodes 0..M.
template <typename R, int M, bool LAB, bool GEN>
__device__ __forceinline__ void block_weights(const PairK<R> &K, R ir, R dx, R dy, R dz,
                                              const typename V4<R>::type &da, const typename V4<R>::type &dah,
                                              const typename V4<R>::type &db, R xc,
                                              const typename V2<R>::type &rot_n, bool live, R (&wr)[M + 1],
                                              R (&wi)[M + 1])
{
  const R ir2 = ir * ir;
  const R ba = da.x * dx + da.y * dy + da.z * dz;
  const R bb = db.x * dx + db.y * dy + db.z * dz;
  const R aabh = dah.x * db.x + dah.y * db.y + dah.z * db.z;
  const R bq = ba * bb * ir2;
  const R g3h = live ? fmaT(R(-1.5), bq, aabh) * ir2 : R(0);
  const R g1 = live ? fmaT(R(2), aabh, -bq) * ir : R(0);
  // the pair term's q0, q1, q2 coefficients (pair_term):
  //   A = g3h (ir + i theta) - g1 theta^2/2,  B = -g3h - i theta g1,  C = g1
  R ar, ai, br, bi, cr, ci;
  if constexpr(LAB) {
    ar = g3h * ir;
    ai = 0;
    br = -g3h;
    bi = 0;
    cr = g1;
    ci = 0;
  } else {
    const R a_r = fmaT(g3h, ir, -(g1 * K.oo)), a_i = g3h * K.theta;
    const R b_i = -(K.theta * g1);
    R sn, cs;
    if constexpr(sizeof(R) == sizeof(double)) {
      if constexpr(GEN)
        sincos(K.theta * xc, &sn, &cs);
      else
        sincos_centred64(K.tg, xc, sn, cs);
    } else {
      if constexpr(GEN)
        __sincosf(K.theta * xc, &sn, &cs);
      else
        sincos_theta(K.sn, K.cs, xc, sn, cs);
    }
    // rotation e^{-i theta (n + 1/2)} e^{-i theta xc}
    const R rr = fmaT(rot_n.x, cs, rot_n.y * sn);
    const R ri = fmaT(rot_n.y, cs, -(rot_n.x * sn));
    ar = fmaT(rr, a_r, -(ri * a_i));
    ai = fmaT(rr, a_i, ri * a_r);
    br = fmaT(-rr, g3h, -(ri * b_i));
    bi = fmaT(rr, b_i, -(ri * g3h));
    cr = rr * g1;
    ci = ri * g1;
  }
  // z = x - M/2 with x = xc + 1/2 the fraction on nodes 0..M
  const R z = xc + R(0.5 - 0.5 * M);
  R pw[M + 1];
  pw[0] = R(1);
#pragma unroll
  for(int j = 1; j <= M; ++j) pw[j] = j == 1 ? z : pw[j - 1] * z;
  R gr[M + 1], gi[M + 1];
  gr[0] = ar;
  gi[0] = ai;
#pragma unroll
  for(int j = 1; j <= M; ++j) {
    const R q = j == 1 ? R(1) : R(j) * pw[j - 1];
    gr[j] = fmaT(ar, pw[j], br * q);
    gi[j] = fmaT(ai, pw[j], bi * q);
    if(j >= 2) {
      const R s = j == 2 ? R(1) : R(j * (j - 1) / 2) * pw[j - 2];
      gr[j] = fmaT(cr, s, gr[j]);
      gi[j] = fmaT(ci, s, gi[j]);
    }
  }
  // W_m = E_m + O_m and W_{M-m} = E_m - O_m, E and O the sums over even and
  // odd j of c[j][m] Gamma_j: 11 terms at M = 4 instead of the 21 nonzero
  // c[j][m] Gamma_j of all five weights
  constexpr LagrangeMonomials<M> L{};
  auto part = [&](int m, int parity, R &sr, R &si) {
    bool first = true;
#pragma unroll
    for(int j = parity; j <= M; j += 2) {
      const double c = L.c[j][m];
      if(c == 0.0) continue;
      if(first) {
        sr = R(c) * gr[j];
        si = R(c) * gi[j];
        first = false;
      } else {
        sr = fmaT(R(c), gr[j], sr);
        si = fmaT(R(c), gi[j], si);
      }
    }
    if(first) sr = si = R(0);
  };
#pragma unroll
  for(int m = 0; 2 * m < M; ++m) {
    R er, ei, orr, oi;
    part(m, 0, er, ei);
    part(m, 1, orr, oi);
    wr[m] = er + orr;
    wi[m] = ei + oi;
    wr[M - m] = er - orr;
    wi[M - m] = ei - oi;
  }
  if constexpr(M % 2 == 0) part(M / 2, 0, wr[M / 2], wi[M / 2]);  // the odd c[j][M/2] vanish
}

// The time-blocked pair sum: for the K steps t0 .. t0 + K - 1 (t0 = the step
// counter) and every ordered pair of delay n >= K, the weighted windows of
// the source's samples, into partial slice j = 0..K-1 (slice j, row r, target
// a at (j (max_chunks + 1) + r) n + a).  Same work list, items and claiming
// as pair_past_*.  Each step keeps three sums of real products (the complex
// product with three multiplications), combined into re and im at the end
// (FP64) or when the FP32 (mixed) sums are folded into FP64, every 64
// sources.  The order-4 FP64 16-step pass with the series sincos is held to
// three waves per SIMD (168 VGPRs, a few of them spilled): 2 % faster than
// two waves at 193 VGPRs.  The FP32 one loses at three (45 spilled) and keeps
// two, and so does the FP64 one with the library sincos (GEN), which would
// spill hundreds of bytes per lane; other orders are left to the compiler.
template <typename R, int M, int K, bool LAB, bool DYN, bool GEN>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == sizeof(double) && K == 16 && M == 4 && !GEN ? 3 : 1, 8))) void
pair_block(const typename V4<R>::type *__restrict__ pd,
                                                     const double4 *__restrict__ pos64,
                                                     const typename V2<R>::type *__restrict__ hist,
                                                     const typename V2<R>::type *__restrict__ rot,
                                                     const int4 *__restrict__ work, const int *__restrict__ work_row,
                                                     int *step_ptr, double2 *__restrict__ partial, Consts k)
{
  using R2 = typename V2<R>::type;
  using R4 = typename V4<R>::type;
  constexpr bool F64 = sizeof(R) == sizeof(double);
  const PairK<R> Kc(k);
  const int t0 = *step_ptr;
  const int sm = (t0 - M) & k.hist_mask;  // window start of a delay-0 pair
  const unsigned hmask = static_cast<unsigned>(k.hist_mask);
  const unsigned row_len = static_cast<unsigned>(k.hist_row);
  const size_t slice = static_cast<size_t>(k.max_chunks + 1) * k.n;
  const float ms = kRoundMagic + static_cast<float>(sm);
  const int code0 = kRoundMagicBits + sm;  // n = code0 - code (FP32 split)
  ItemCursor<DYN> items(step_ptr + 1, k.n_work);
  for(int w = items.first(); w < k.n_work; w = items.next()) {
    const int4 wi = work[w];
    const int tgt = wi.x + static_cast<int>(threadIdx.x % kWave);
    const int a = tgt < wi.y ? tgt : wi.y - 1;
    const R4 pa = pd[2 * a], da = pd[2 * a + 1];
    R4 dah = da;
    dah.x = R(0.5) * da.x;
    dah.y = R(0.5) * da.y;
    dah.z = R(0.5) * da.z;
    const double4 pa64 = pos64[a];
    // one source into the three step sums p1..p3 (see below)
    auto source = [&](int b, R(&p1)[K], R(&p2)[K], R(&p3)[K]) {
      const R4 pb = pd[2 * b], db = pd[2 * b + 1];
      const R dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
      R xc, ir;
      int n;
      if constexpr(F64) {
        n = split_l64(dx, dy, dz, pa64, pos64, b, Kc.cdt, xc, ir);
      } else {
        const int code = split_mixed(dx, dy, dz, Kc.eps_rel, Kc.lim0, ms, pa64, pos64, b, Kc.cdt, Kc.icdt64, xc, ir);
        n = code0 - code;
      }
      // window of steps t0 - n - M .. t0 - n + K - 1 in the source's row
      const unsigned first = (static_cast<unsigned>(sm - n) & hmask) + static_cast<unsigned>(b) * row_len;
      const R2 *win = reinterpret_cast<const R2 *>(reinterpret_cast<const char *>(hist) +
                                                   first * static_cast<unsigned>(sizeof(R2)));
      R2 smp[K + M];
#pragma unroll
      for(int i = 0; i < K + M; ++i) smp[i] = win[i];
      const R2 rn = rot[n];
      R wr[M + 1], wim[M + 1];
      block_weights<R, M, LAB, GEN>(Kc, ir, dx, dy, dz, da, dah, db, xc, rn, n >= K, wr, wim);
      // step t0 + j reads sample t0 + j - n - m = window entry j - m + M
      if constexpr(LAB) {
        // real weights on real samples (2 Re rho01; the imaginary parts are 0)
#pragma unroll
        for(int i = 0; i < K + M; ++i)
#pragma unroll
          for(int m = 0; m <= M; ++m) {
            const int j = i + m - M;
            if(j < 0 || j >= K) continue;
            p1[j] = fmaT(wr[m], smp[i].x, p1[j]);
          }
      } else {
        // (wr + i wi)(x + i y) with three products (Gauss): the sums of wr x,
        // wi y and (wr + wi)(x + y) give re = p1 - p2, im = p3 - p1 - p2,
        // formed when the sums are folded: 3 FMAs per tap instead of 4
        R ws[M + 1];
#pragma unroll
        for(int m = 0; m <= M; ++m) ws[m] = wr[m] + wim[m];
#pragma unroll
        for(int i = 0; i < K + M; ++i) {
          const R ss = smp[i].x + smp[i].y;
#pragma unroll
          for(int m = 0; m <= M; ++m) {
            const int j = i + m - M;
            if(j < 0 || j >= K) continue;
            p1[j] = fmaT(wr[m], smp[i].x, p1[j]);
            p2[j] = fmaT(wim[m], smp[i].y, p2[j]);
            p3[j] = fmaT(ws[m], ss, p3[j]);
          }
        }
      }
    };
    double accr[K], acci[K];
#pragma unroll
    for(int j = 0; j < K; ++j) accr[j] = acci[j] = 0.0;
    if constexpr(F64) {
      double acc3[K];
#pragma unroll
      for(int j = 0; j < K; ++j) acc3[j] = 0.0;
      for(int b = wi.z; b < wi.w; ++b) source(b, accr, acci, acc3);
      if constexpr(!LAB) {
#pragma unroll
        for(int j = 0; j < K; ++j) {
          const double q1 = accr[j], q2 = acci[j];
          accr[j] = q1 - q2;
          acci[j] = acc3[j] - q1 - q2;
        }
      }
    } else {
      // FP32 sums of up to 64 sources, folded into FP64
      for(int bt = wi.z; bt < wi.w; bt += kChunk) {
        const int be = min(wi.w, bt + kChunk);
        R q1[K], q2[K], q3[K];
#pragma unroll
        for(int j = 0; j < K; ++j) q1[j] = q2[j] = q3[j] = R(0);
        for(int b = bt; b < be; ++b) source(b, q1, q2, q3);
#pragma unroll
        for(int j = 0; j < K; ++j) {
          const double d1 = q1[j];
          if constexpr(LAB) {
            accr[j] += d1;
          } else {
            const double d2 = q2[j];
            accr[j] += d1 - d2;
            acci[j] += static_cast<double>(q3[j]) - d1 - d2;
          }
        }
      }
    }
    if(tgt < wi.y) {
      double2 *out = partial + static_cast<size_t>(work_row[w]) * k.n + tgt;
#pragma unroll
      for(int j = 0; j < K; ++j)
        out[j * slice] = make_double2(accr[j] * k.unit_scale, acci[j] * k.unit_scale);
    }
  }
}

// Listed pairs of every target (CSR by target) over KS steps from t0 (the
// step counter): the short list (delays 1 .. 3) at every step (KS = 1), the
// delays 4 .. 7 at every multiple of 4 (KS = 4) and with 16-step passes the
// delays 8 .. 15 at every multiple of 8 (KS = 8).  Per pair the weights of
// block_weights once, then KS windows of
// the source's samples (all finished: t0 + KS - 1 - n <= t0 - 1), summed in
// FP64 into out[j n + a], j < KS.  kListLanes consecutive lanes share a
// target and take every kListLanes-th pair of its list (a chain of dependent
// loads per pair: list entry, source, samples; U of them in flight per lane),
// then add their sums by a butterfly over the lanes.
constexpr int kListLanes = 8;

template <typename R, int M, int KS, bool LAB, bool GEN>
__global__ __launch_bounds__(kBlock) void pair_list(const typename V4<R>::type *__restrict__ pd,
                                                    const double4 *__restrict__ pos64,
                                                    const typename V2<R>::type *__restrict__ hist,
                                                    const typename V2<R>::type *__restrict__ rot,
                                                    const int *__restrict__ list_start,
                                                    const int *__restrict__ list_src, const int *step_ptr,
                                                    double2 *__restrict__ out, Consts k)
{
  using R2 = typename V2<R>::type;
  using R4 = typename V4<R>::type;
  constexpr bool F64 = sizeof(R) == sizeof(double);
  static_assert(kWave % kListLanes == 0, "a target's lanes lie in one wave");
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int sub = gid % kListLanes;
  const bool valid = gid / kListLanes < k.n;
  const int a = valid ? gid / kListLanes : k.n - 1;
  const PairK<R> Kc(k);
  const int t0 = *step_ptr;
  const int sm = (t0 - M) & k.hist_mask;  // window start of a delay-0 pair
  const R4 pa = pd[2 * a], da = pd[2 * a + 1];
  R4 dah = da;
  dah.x = R(0.5) * da.x;
  dah.y = R(0.5) * da.y;
  dah.z = R(0.5) * da.z;
  const double4 pa64 = pos64[a];
  double accr[KS], acci[KS];
#pragma unroll
  for(int j = 0; j < KS; ++j) accr[j] = acci[j] = 0.0;
  // U pairs of the lane's share in flight at once: their list entries, then
  // their sources, then their sample windows are loaded together, so a lane
  // waits for U chains of dependent loads at a time instead of one.  The
  // window lists (~3 and ~26 pairs per lane at 512 um^-3) take 2; the short
  // list 1 (with 4 the ensemble's list of delays 1 .. 7, ~3 pairs per lane,
  // ran 58 -> 68 us per step: half the occupancy, a quarter of the slots
  // empty)
  constexpr int U = KS == 1 ? 1 : 2;
  const int e1 = valid ? list_start[a + 1] : 0;
  for(int e = list_start[a] + sub; e < e1; e += kListLanes * U) {
    int b[U];
    bool ok[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      ok[u] = e + u * kListLanes < e1;
      b[u] = list_src[ok[u] ? e + u * kListLanes : e];  // a repeated entry where the share ends, not summed
    }
    R4 pb[U], db[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      pb[u] = pd[2 * b[u]];
      db[u] = pd[2 * b[u] + 1];
    }
    R dx[U], dy[U], dz[U], xc[U], ir[U];
    int n[U];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      dx[u] = pa.x - pb[u].x;
      dy[u] = pa.y - pb[u].y;
      dz[u] = pa.z - pb[u].z;
      if constexpr(F64) {
        n[u] = split_l64(dx[u], dy[u], dz[u], pa64, pos64, b[u], Kc.cdt, xc[u], ir[u]);
      } else {
        const int code = split_mixed(dx[u], dy[u], dz[u], Kc.eps_rel, Kc.lim0, kRoundMagic, pa64, pos64, b[u],
                                     Kc.cdt, Kc.icdt64, xc[u], ir[u]);
        n[u] = kRoundMagicBits - code;
      }
    }
    // windows of steps t0 - n - M .. t0 - n + KS - 1 (contiguous in the row)
    R2 rn[U], smp[U][KS + M];
#pragma unroll
    for(int u = 0; u < U; ++u) {
      rn[u] = rot[n[u]];
      const R2 *win = hist + static_cast<size_t>(b[u]) * k.hist_row + ((sm - n[u]) & k.hist_mask);
#pragma unroll
      for(int i = 0; i < KS + M; ++i) smp[u][i] = win[i];
    }
#pragma unroll
    for(int u = 0; u < U; ++u) {
      R wr[M + 1], wim[M + 1];
      block_weights<R, M, LAB, GEN>(Kc, ir[u], dx[u], dy[u], dz[u], da, dah, db[u], xc[u], rn[u], true, wr, wim);
      // the three real products of pair_block's complex multiply-adds (one
      // in the lab frame, where weights and samples are real)
      R q1[KS], q2[KS], q3[KS];
#pragma unroll
      for(int j = 0; j < KS; ++j) q1[j] = q2[j] = q3[j] = R(0);
      R ws[M + 1];
#pragma unroll
      for(int m = 0; m <= M; ++m) ws[m] = wr[m] + wim[m];
#pragma unroll
      for(int i = 0; i < KS + M; ++i) {
        const R ss = smp[u][i].x + smp[u][i].y;
#pragma unroll
        for(int m = 0; m <= M; ++m) {
          const int j = i + m - M;
          if(j < 0 || j >= KS) continue;
          q1[j] = fmaT(wr[m], smp[u][i].x, q1[j]);
          if constexpr(!LAB) {
            q2[j] = fmaT(wim[m], smp[u][i].y, q2[j]);
            q3[j] = fmaT(ws[m], ss, q3[j]);
          }
        }
      }
      if(ok[u])
#pragma unroll
        for(int j = 0; j < KS; ++j) {
          const double d1 = q1[j];
          if constexpr(LAB) {
            accr[j] += d1;
          } else {
            const double d2 = q2[j];
            accr[j] += d1 - d2;
            acci[j] += static_cast<double>(q3[j]) - d1 - d2;
          }
        }
    }
  }
#pragma unroll
  for(int j = 0; j < KS; ++j)
#pragma unroll
    for(int off = 1; off < kListLanes; off <<= 1) {
      accr[j] += __shfl_xor(accr[j], off, kListLanes);
      acci[j] += __shfl_xor(acci[j], off, kListLanes);
    }
  if(valid)
#pragma unroll
    for(int j = 0; j < KS; ++j)
      if(j % kListLanes == sub)
        out[static_cast<size_t>(j) * k.n + a] = make_double2(accr[j] * k.unit_scale, acci[j] * k.unit_scale);
}

// Past-only pair sum from the coefficient table (the reference's evaluate,
// history_interaction.cpp:47-70, as a gather: per target, sum over sources of
// coef[m] rho01_b(step - n - m)).  Same work list, grid and partial layout as
// the on-the-fly kernels; the coefficients are complete, so pair_field applies
// no further scale.  Bound by HBM: per ordered pair the kernel streams
// (M + 1) x 16 B of coefficients and a 2-B delay; the sample gather (M + 1)
// x 16 B per lane hits L1/L2 (a source's few hundred samples are shared by
// all targets).
template <int M, bool LAB>
__global__ __launch_bounds__(kBlock) void pair_table_f64(
    const double2 *__restrict__ table, const int16_t *__restrict__ table_n,
    const double2 *__restrict__ samp, const int4 *__restrict__ work, const int *__restrict__ work_row,
    int *step_ptr, const int *__restrict__ sys_of, const int2 *__restrict__ sys_range,
    const long long *__restrict__ sys_toff, double2 *__restrict__ partial, Consts k)
{
  const int step = *step_ptr;
  const size_t P = static_cast<size_t>(k.table_plane);
  ItemCursor<false> items(step_ptr + 1, k.n_work);  // table items are uniform: static stride
  for(int w = items.first(); w < k.n_work; w = items.next()) {
    const int4 wi = work[w];
    const int s = sys_of[wi.x];
    const int2 rg = sys_range[s];
    const size_t ns = static_cast<size_t>(rg.y - rg.x);
    const int tgt = wi.x + static_cast<int>(threadIdx.x % kWave);
    const int a = tgt < wi.y ? tgt : wi.y - 1;
    const size_t base = static_cast<size_t>(sys_toff[s]) + static_cast<size_t>(a - rg.x);
    double accr = 0, acci = 0;
    for(int b = wi.z; b < wi.w; ++b) {
      const size_t e = base + static_cast<size_t>(b - rg.x) * ns;
      const int n = table_n[e];
      const double2 *sb = samp + static_cast<size_t>(b) * k.samp_slots;
#pragma unroll
      for(int m = 0; m <= M; ++m) {
        const double2 c = table[m * P + e];
        double2 r = sb[(step - n - m) & k.samp_mask];
        if constexpr(LAB) r = make_double2(2.0 * r.x, 0.0);
        accr = fma(c.x, r.x, fma(-c.y, r.y, accr));
        acci = fma(c.x, r.y, fma(c.y, r.x, acci));
      }
    }
    if(tgt < wi.y) partial[static_cast<size_t>(work_row[w]) * k.n + tgt] = make_double2(accr, acci);
  }
}

// ------------------------------------------------------- per-dot kernels

// e^{i theta s} for 0 <= s < 2^31 from the three factor tables of
// DeviceState::phase (s = 2^20 h + 2^10 m + l): e^{i theta 2^20 h} e^{i theta 2^10 m}
// e^{i theta l}, each entry rounded once from a long-double argument reduction
// on the host, so the product is within a few FP64 units of e^{i theta s}.
__device__ __forceinline__ double2 phase_at(const double2 *phase, int s)
{
  const double2 hi = phase[2 * kPhaseLevel + (s >> 20)];
  const double2 mid = phase[kPhaseLevel + ((s >> 10) & (kPhaseLevel - 1))];
  const double2 lo = phase[s & (kPhaseLevel - 1)];
  return cmul(cmul(hi, mid), lo);
}

// Field scale of the pair sum at `step`: -K e^{-i theta step} (rotating) or K (lab).
__device__ __forceinline__ double2 hist_scale(const Consts &k, const double2 *phase, int step)
{
  if(k.dyadic == QUEST_DYADIC_LAB) return make_double2(k.k_green, 0.0);
  const double2 ph = phase_at(phase, step);
  return make_double2(-k.k_green * ph.x, k.k_green * ph.y);
}

// Sum of the pair-kernel partials of target a over its system's source chunks
// (rows of `partial`, n entries each).
__device__ __forceinline__ double2 reduce_partials(const Consts &k, const DeviceState &d,
                                                   const double2 *partial, int a)
{
  double2 s = make_double2(0.0, 0.0);
  const int rows = d.sys_chunks[d.sys_of[a]];
  // unrolled so that several rows' loads are in flight; the sum stays in row order
#pragma unroll 8
  for(int j = 0; j < rows; ++j) s = cadd(s, partial[static_cast<size_t>(j) * k.n + a]);
  return s;
}

// Field of the past pair sum at target a: the partials times -K e^{-i theta
// step} (on-the-fly), as they are (the table holds complete coefficients), or
// (time-blocked) the slice of this step plus the sums of the short list and
// of the window lists for this step, times -K (rotating; the rotation is in
// the weights) or K (lab).
__device__ __forceinline__ double2 pair_field(const Consts &k, const DeviceState &d, int a, int step)
{
  if(k.pair_blocked) {
    const double2 *slice = d.partial + static_cast<size_t>(step % k.block_steps) * (k.max_chunks + 1) * k.n;
    double2 sum = cadd(reduce_partials(k, d, slice, a), d.short_sum[a]);
#pragma unroll
    for(int l = 0; l < kWindowLists; ++l)
      if(k.win_steps[l]) sum = cadd(sum, d.win_sum[l][static_cast<size_t>(step % k.win_steps[l]) * k.n + a]);
    const double g = k.dyadic == QUEST_DYADIC_LAB ? k.k_green : -k.k_green;
    return make_double2(g * sum.x, g * sum.y);
  }
  const double2 sum = reduce_partials(k, d, d.partial, a);
  return k.pair_table ? sum : cmul(sum, hist_scale(k, d.phase, step));
}

// Pulse::operator() projected on the dipole (pulse.cpp:15-20,
// pulse_interaction.cpp:11-16, gaussian math_utils.cpp:27-30).
__device__ __forceinline__ double2 pulse_field(const PulseParams &pl, double hbar, double4 r,
                                               double4 d, double t)
{
  const double kr = (pl.k[0] * r.x + pl.k[1] * r.y) + pl.k[2] * r.z;
  const double arg = kr - pl.freq * (t - pl.delay);
  const double u = arg / pl.width;
  const double g = exp(-(u * u) / 2);
  const double e0 = (pl.amplitude / 2 * pl.pol[0]) * g;
  const double e1 = (pl.amplitude / 2 * pl.pol[1]) * g;
  const double e2 = (pl.amplitude / 2 * pl.pol[2]) * g;
  return make_double2(((e0 * d.x + e1 * d.y) + e2 * d.z) / hbar, 0.0);
}

// The pulses of dot a's system, summed in list order from 0 (the reference
// integrator accumulates its interactions from zero, integrator.cpp:173-179).
__device__ __forceinline__ double2 pulse_sum(const Consts &k, const DeviceState &d, int a, double t)
{
  const PulseParams *pl = d.pulses + static_cast<size_t>(d.sys_of[a]) * QUEST_MAX_PULSES;
  double2 f = make_double2(0.0, 0.0);
  for(int p = 0; p < QUEST_MAX_PULSES && pl[p].on; ++p)
    f = cadd(f, pulse_field(pl[p], k.hbar, d.pos64[a], d.dip64[a], t));
  return f;
}

// QuantumDot::liouville_rhs (quantum_dot.cpp:10-19).
__device__ __forceinline__ void bloch_rhs(double r0, double2 r1, double2 om, double laser,
                                          double freq, double t1, double t2, double &m0,
                                          double2 &m1)
{
  // z = om conj(r1) - conj(om) r1 = i 2 Im(om conj(r1)): its real part is the
  // difference of two identically rounded products, an exact 0
  const double zy = (om.y * r1.x - om.x * r1.y) - (om.x * r1.y - om.y * r1.x);
  // -i z - (r0 - 1) / t1, real (its imaginary part -Re z - Im r0 / t1 is 0)
  m0 = zy - (r0 - 1.0) / t1;
  // w = om (1 - 2 r0) + r1 (laser - freq);  -i w - r1 / t2
  const double f = 1.0 - 2.0 * r0;
  const double det = laser - freq;
  const double2 w = make_double2(om.x * f + r1.x * det, om.y * f + r1.y * det);
  m1 = make_double2(w.y - r1.x / t2, -w.x - r1.y / t2);
}

__device__ __forceinline__ bool finite2(double2 v) { return isfinite(v.x) && isfinite(v.y); }

// rho01 of dot b at time t, seen from the evaluation of step `step`: the
// current iterate for t == step, the state ring for finished steps (the ring
// holds the ground state, rho01 = 0, for t < 0).
__device__ __forceinline__ double2 rho01_at(const Consts &k, const double2 *cur1,
                                            const double2 *y1, int b, int t, int step)
{
  if(t == step) return cur1[b];
  return y1[static_cast<size_t>(t & k.state_mask) * k.n + b];
}

// Zero-delay (near) pairs of target a, evaluated with the reference's
// per-pair coefficients (history_interaction.cpp:56-66, one direction).
__device__ __forceinline__ double2 near_sum(const Consts &k, const DeviceState &d,
                                            const double2 *cur1, int a, int step)
{
  double2 acc = make_double2(0.0, 0.0);
  const int e0 = d.near_start[a], e1 = d.near_start[a + 1];
  for(int e = e0; e < e1; ++e) {
    const int b = d.near_src[e];
    const int nn = d.near_n[e];
    const double2 *cf = d.near_coef + static_cast<size_t>(e) * (k.order + 1);
    for(int m = 0; m <= k.order; ++m) {
      double2 rho = rho01_at(k, cur1, d.y1, b, step - nn - m, step);
      if(k.dyadic == QUEST_DYADIC_LAB) rho = make_double2(2.0 * rho.x, 0.0);
      acc = cadd(acc, cmul(rho, cf[m]));
    }
  }
  return acc;
}

// The same sum split for the one-launch step, where every near pair has delay
// 0 (the list holds the pairs with n < 1): the terms m >= 1 read finished
// steps and are formed once per step (near_fixed); only the m = 0 term reads
// the current iterate of the partner (near_current, every iteration).
__device__ __forceinline__ double2 near_fixed(const Consts &k, const DeviceState &d, int a, int step)
{
  double2 acc = make_double2(0.0, 0.0);
  const int e0 = d.near_start[a], e1 = d.near_start[a + 1];
  for(int e = e0; e < e1; ++e) {
    const int b = d.near_src[e];
    const double2 *cf = d.near_coef + static_cast<size_t>(e) * (k.order + 1);
    for(int m = 1; m <= k.order; ++m) {
      double2 rho = d.y1[static_cast<size_t>((step - m) & k.state_mask) * k.n + b];
      if(k.dyadic == QUEST_DYADIC_LAB) rho = make_double2(2.0 * rho.x, 0.0);
      acc = cadd(acc, cmul(rho, cf[m]));
    }
  }
  return acc;
}
__device__ __forceinline__ double2 near_current(const Consts &k, const DeviceState &d, const double2 *xchg,
                                                int a)
{
  double2 acc = make_double2(0.0, 0.0);
  const int e0 = d.near_start[a], e1 = d.near_start[a + 1];
  for(int e = e0; e < e1; ++e) {
    double2 rho = xchg[d.near_lane[e]];  // the partner's iterate, from its thread's LDS entry
    if(k.dyadic == QUEST_DYADIC_LAB) rho = make_double2(2.0 * rho.x, 0.0);
    acc = cadd(acc, cmul(rho, d.near_coef[static_cast<size_t>(e) * (k.order + 1)]));
  }
  return acc;
}

// Writes the interpolant coefficients of slot s for source a:
// c_k = e^{i theta s} sum_m vinv[k][m] rho01(s - m)  (rotating), or of
// 2 Re rho01 without the phase (lab), and zeros into slot s + 1.  The pair
// kernels of step s + 1 read slot s + 1 for exactly the pairs of delay 0 (a
// dot with itself, zero-delay partners, which the near list evaluates), so
// those terms vanish without a per-pair test; the slot's previous content
// (step s + 1 - ring_slots) is older than any delay.
//   Time-blocked method: `ring` is the sample ring instead, and slot s of
// source a gets rho01(s) (2 Re rho01 in the lab frame) at entry s & hist_mask
// of row a, and again hist_slots entries further while that is below
// K + M - 1 (the window's contiguous tail).
template <typename R2>
__device__ __forceinline__ void build_slot(const Consts &k, const double2 *phase, R2 *ring,
                                           double2 *samp, const double2 *y1, double2 rho_s, int a,
                                           int s)
{
  if(k.pair_blocked) {
    R2 v;
    v.x = k.dyadic == QUEST_DYADIC_LAB ? 2.0 * rho_s.x : rho_s.x;
    v.y = k.dyadic == QUEST_DYADIC_LAB ? 0.0 : rho_s.y;
    R2 *row = ring + static_cast<size_t>(a) * k.hist_row;
    const int e = s & k.hist_mask;
    row[e] = v;
    if(e < k.block_steps + k.order - 1) row[k.hist_slots + e] = v;
    return;
  }
  if(samp) samp[static_cast<size_t>(a) * k.samp_slots + (s & k.samp_mask)] = rho_s;
  // loops unrolled to the largest order (guards on k.order) keep smp in registers
  double2 smp[kMaxOrder + 1];
  smp[0] = rho_s;
#pragma unroll
  for(int m = 1; m <= kMaxOrder; ++m)
    smp[m] = m <= k.order ? y1[static_cast<size_t>((s - m) & k.state_mask) * k.n + a] : make_double2(0.0, 0.0);
  const bool lab = k.dyadic == QUEST_DYADIC_LAB;
  if(lab)
#pragma unroll
    for(int m = 0; m <= kMaxOrder; ++m) smp[m] = make_double2(2.0 * smp[m].x, 0.0);
  // coefficients about x = 1/2, so the rotation e^{-i theta x} of the pair
  // kernels splits into e^{-i theta / 2} (here) and e^{-i theta xc}
  const double *basis = k.vinv_c;
  const double2 ph =
      lab ? make_double2(1.0, 0.0) : cmul(phase_at(phase, s), make_double2(k.half_phase_re, k.half_phase_im));
  const size_t E = static_cast<size_t>(k.n) * k.ring_slots;
  const size_t e = static_cast<size_t>(a) * k.ring_slots + (s & k.ring_mask);
  for(int kk = 0; kk <= k.order; ++kk) {
    double2 c = make_double2(0.0, 0.0);
#pragma unroll
    for(int m = 0; m <= kMaxOrder; ++m) {
      if(m > k.order) break;
      const double v = basis[kk * (kMaxOrder + 1) + m];
      c.x += v * smp[m].x;
      c.y += v * smp[m].y;
    }
    c = cmul(c, ph);
    R2 &dst = ring[ring_index(ring, k.order, E, e, kk)];
    dst.x = c.x;
    dst.y = c.y;
  }
  const size_t e_next = static_cast<size_t>(a) * k.ring_slots + ((s + 1) & k.ring_mask);
  for(int kk = 0; kk <= k.order; ++kk) {
    R2 &dst = ring[ring_index(ring, k.order, E, e_next, kk)];
    dst.x = 0;
    dst.y = 0;
  }
}

// Ground state in every state-ring slot (main.cpp:32-36), zero derivatives.
__global__ void fill_ground_state_kernel(Consts k, double *y0, double2 *y1, double *d0,
                                         double2 *d1, int *bad_step)
{
  const size_t total = static_cast<size_t>(k.state_slots) * k.n;
  for(size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
      i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    y0[i] = 1.0;
    y1[i] = make_double2(0.0, 0.0);
    d0[i] = 0.0;
    d1[i] = make_double2(0.0, 0.0);
  }
  if(blockIdx.x == 0 && threadIdx.x == 0) *bad_step = INT_MAX;
}

__global__ void clear_words_kernel(uint4 *p, size_t n16)
{
  for(size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n16;
      i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = make_uint4(0, 0, 0, 0);
}

// Start of a step (whole-run mode): fixed field (pulse + past pair sum), the
// predictor (integrator.cpp:142-154) and the step-invariant part of the
// corrector (integrator.cpp:156-169).
__device__ __forceinline__ void pc_predict(const Consts &k, const DeviceState &d, int window, int a, int step,
                                           double2 &field, double &p0, double2 &p1, double &c0, double2 &c1)
{
  const double dt = k.dt;
  double2 hist = make_double2(0.0, 0.0);
  if(k.n_work > 0) hist = pair_field(k, d, a, step);
  const double2 pulse = pulse_sum(k, d, a, step * dt);
  field = cadd(pulse, hist);

  const double *ps0 = d.pcw, *ps1 = d.pcw + window;
  const double *cs0 = d.pcw + 2 * window, *cs1 = d.pcw + 3 * window;
  p0 = 0.0;
  c0 = 0.0;
  p1 = make_double2(0.0, 0.0);
  c1 = p1;
  const int start = step - window;
#pragma unroll 4
  for(int h = 0; h < window; ++h) {
    const size_t o = static_cast<size_t>((start + h) & k.state_mask) * k.n + a;
    const double y0 = d.y0[o], g0 = d.d0[o];
    const double2 y1 = d.y1[o], g1 = d.d1[o];
    const double pw = ps0[h], pd = ps1[h] * dt, cw = cs0[h], cd = cs1[h] * dt;
    p0 += y0 * pw + g0 * pd;
    p1.x += y1.x * pw + g1.x * pd;
    p1.y += y1.y * pw + g1.y * pd;
    c0 += y0 * cw + g0 * cd;
    c1.x += y1.x * cw + g1.x * cd;
    c1.y += y1.y * cw + g1.y * cd;
  }
}

__device__ __forceinline__ void pc_begin_dot(const Consts &k, const DeviceState &d, int window, int a)
{
  double2 field, p1, c1;
  double p0, c0;
  pc_predict(k, d, window, a, *d.step, field, p0, p1, c0, c1);
  d.field[a] = field;
  d.cur0[a] = p0;
  d.cur1[a] = p1;
  d.yc0[a] = c0;
  d.yc1[a] = c1;
}

__global__ __launch_bounds__(kBlock) void pc_begin_kernel(Consts k, DeviceState d, int window)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a < k.n) pc_begin_dot(k, d, window, a);
}

// One evaluation of the step (integrator.cpp:171-185) and the next corrector
// update (integrator.cpp:156-169) for iterations 0..9; iteration 10 commits
// the step to the state ring, checks finiteness (integrator.cpp:187-207) and
// writes the source's interpolant slot for the next steps' pair sums.
// The step's final iterate and derivative into the state ring; finiteness
// check (integrator.cpp:187-207); the source's interpolant slot.
template <typename R2>
__device__ __forceinline__ void pc_commit(const Consts &k, const DeviceState &d, R2 *ring, int a, int step,
                                          double y0, double2 y1, double m0, double2 m1)
{
  const size_t o = static_cast<size_t>(step & k.state_mask) * k.n + a;
  d.y0[o] = y0;
  d.y1[o] = y1;
  d.d0[o] = m0;
  d.d1[o] = m1;
  if(!(isfinite(y0) && finite2(y1) && isfinite(m0) && finite2(m1))) atomicMin(d.bad_step, step);
  build_slot(k, d.phase, ring, d.samp, d.y1, y1, a, step);
}

template <typename R2>
__device__ __forceinline__ void pc_iter_dot(const Consts &k, const DeviceState &d, R2 *ring,
                                            double future_coef, int iter, int a)
{
  const int step = *d.step;
  const int rd = iter & 1, wr = rd ^ 1;
  const double *cur0 = d.cur0 + static_cast<size_t>(rd) * k.n;
  const double2 *cur1 = d.cur1 + static_cast<size_t>(rd) * k.n;
  const double y0 = cur0[a];
  const double2 y1 = cur1[a];
  double2 om = d.field[a];
  if(d.near_start[a + 1] > d.near_start[a]) om = cadd(om, near_sum(k, d, cur1, a, step));
  double m0;
  double2 m1;
  bloch_rhs(y0, y1, om, k.laser_omega, d.freq[a], d.t1[a], d.t2[a], m0, m1);
  if(iter < kNumCorrectorSteps) {
    const double f = future_coef * k.dt;
    d.cur0[static_cast<size_t>(wr) * k.n + a] = f * m0 + d.yc0[a];
    d.cur1[static_cast<size_t>(wr) * k.n + a] = make_double2(f * m1.x + d.yc1[a].x, f * m1.y + d.yc1[a].y);
    return;
  }
  pc_commit(k, d, ring, a, step, y0, y1, m0, m1);
}

template <typename R2>
__global__ __launch_bounds__(kBlock) void pc_iter_kernel(Consts k, DeviceState d, R2 *ring,
                                                         double future_coef, int iter)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a < k.n) pc_iter_dot(k, d, ring, future_coef, iter, a);
}

// The whole predictor-corrector step of every dot in one launch: predictor,
// then 11 x (evaluate, correct or commit), the same arithmetic as
// pc_begin_kernel + pc_iter_kernel.  A dot's iterate, corrector base, fixed
// field and parameters stay in registers.  The only coupling inside a step is
// a zero-delay partner's current iterate: the host packs dots (step_map:
// thread -> dot, -1 idle) so that every connected group of zero-delay
// partners sits in one of the first `sync_blocks` blocks; there a dot with
// partners publishes each iterate to an LDS double buffer (near_lane: each
// near pair's source as a thread of the block), and a block barrier
// separates the iterations.
// Blocks of dots without partners run barrier-free.  The last block to finish
// advances the step counter (what advance_step_kernel does after the
// per-iteration kernels): every block has read it by then.
template <typename R2>
__global__ __launch_bounds__(kBlock) void pc_step_kernel(Consts k, DeviceState d, R2 *ring,
                                                         double future_coef, int window, int sync_blocks)
{
  const int a = d.step_map[blockIdx.x * blockDim.x + threadIdx.x];
  const bool sync = static_cast<int>(blockIdx.x) < sync_blocks;
  const int step = *d.step;
  // iterates of zero-delay partners, exchanged through LDS (double buffer)
  __shared__ double2 xchg[2][kBlock];
  double2 field, y1, c1;
  double y0, c0;
  double freq = 0, t1 = 1, t2 = 1;
  bool partners = false;
  if(a >= 0) {
    pc_predict(k, d, window, a, step, field, y0, y1, c0, c1);
    freq = d.freq[a];
    t1 = d.t1[a];
    t2 = d.t2[a];
    partners = sync && d.near_start[a + 1] > d.near_start[a];
    if(partners) {
      xchg[0][threadIdx.x] = y1;
      field = cadd(field, near_fixed(k, d, a, step));
    }
  }
  if(sync) __syncthreads();
  const double f = future_coef * k.dt;
  double m0;
  double2 m1;
  for(int iter = 0; iter <= kNumCorrectorSteps; ++iter) {
    if(a >= 0) {
      double2 om = field;
      if(partners) om = cadd(om, near_current(k, d, xchg[iter & 1], a));
      bloch_rhs(y0, y1, om, k.laser_omega, freq, t1, t2, m0, m1);
      if(iter < kNumCorrectorSteps) {
        y0 = f * m0 + c0;
        y1 = make_double2(f * m1.x + c1.x, f * m1.y + c1.y);
        if(partners) xchg[(iter & 1) ^ 1][threadIdx.x] = y1;
      }
    }
    if(sync && iter < kNumCorrectorSteps) __syncthreads();
  }
  if(a >= 0) pc_commit(k, d, ring, a, step, y0, y1, m0, m1);
  // No fence: the counters order nothing but themselves (every thread's read
  // of the step completed before the barrier; data reaches the next kernel at
  // the kernel boundary), and a device-scope fence writes back the L2 here.
  // Block b counts itself in group b % kStepGroups; the group's last block
  // counts the group, and the last group advances the step (one counter for
  // all blocks queued ~3500 atomics on one line: 14 us of the ensemble's
  // step kernel).
  __syncthreads();
  if(threadIdx.x == 0) {
    const int groups = min(static_cast<int>(gridDim.x), kStepGroups);
    const int g = static_cast<int>(blockIdx.x) % kStepGroups;
    const int members = (static_cast<int>(gridDim.x) - g + kStepGroups - 1) / kStepGroups;
    int *group = d.step + kCounterStride * (1 + g);
    if(atomicAdd(group, 1) == members - 1) {
      atomicExch(group, 0);
      if(atomicAdd(d.step + 3, 1) == groups - 1) {
        atomicExch(d.step + 3, 0);
        atomicAdd(d.step, 1);
      }
    }
  }
}

// Per-call mode: the values passed for `step` become final -- state ring
// slot and interpolant slot of every source.
template <typename R2>
__global__ __launch_bounds__(kBlock) void hist_commit_kernel(Consts k, DeviceState d, R2 *ring,
                                                             int step)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a >= k.n) return;
  const double2 y1 = d.cur1[a];
  d.y1[static_cast<size_t>(step & k.state_mask) * k.n + a] = y1;
  build_slot(k, d.phase, ring, d.samp, d.y1, y1, a, step);
}

// Per-call mode: the pair sum of the finished history, scaled.
__global__ __launch_bounds__(kBlock) void hist_reduce_kernel(Consts k, DeviceState d)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a >= k.n) return;
  const int step = *d.step;
  double2 hist = make_double2(0.0, 0.0);
  if(k.n_work > 0) hist = pair_field(k, d, a, step);
  d.field[a] = hist;
}

// Per-call mode: the caller's rho01 (caller order, io[0..n)) into the sorted
// iterate; the results back into caller order (io[n..2n)).  inv: caller ->
// sorted index.
__global__ __launch_bounds__(kBlock) void permute_in_kernel(int n, const double2 *__restrict__ io,
                                                            const int *__restrict__ inv, double2 *__restrict__ cur1)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) cur1[inv[i]] = io[i];
}
__global__ __launch_bounds__(kBlock) void permute_out_kernel(int n, const double2 *__restrict__ out,
                                                             const int *__restrict__ inv, double2 *__restrict__ io)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if(i < n) io[n + i] = out[inv[i]];
}

// Per-call mode: out = past pair sum + zero-delay pairs with the current rho01.
__global__ __launch_bounds__(kBlock) void hist_near_kernel(Consts k, DeviceState d)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a >= k.n) return;
  const int step = *d.step;
  double2 om = d.field[a];
  if(d.near_start[a + 1] > d.near_start[a]) om = cadd(om, near_sum(k, d, d.cur1, a, step));
  d.out[a] = om;
}

__global__ __launch_bounds__(kBlock) void pulse_kernel(Consts k, DeviceState d)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a >= k.n) return;
  d.out[a] = pulse_sum(k, d, a, *d.step * k.dt);
}

__global__ void set_step_kernel(int *step, int value) { *step = value; }
__global__ void advance_step_kernel(int *step) { *step += 1; }

// ------------------------------------------------------- zero-delay list

// Number (FILL = false) or list (FILL = true, in source order) of the sources
// b != a of each target whose pair the pair kernels of this precision
// classify with an integer delay in [lo, hi): the zero-delay list ([0, 1))
// and the time-blocked method's lists ([1, 4), [4, 8), [8, 16)).
template <bool F64, bool FILL>
__global__ __launch_bounds__(kBlock) void near_scan_kernel(Consts k, DeviceState d, int lo, int hi,
                                                           const int *__restrict__ start, int *__restrict__ src,
                                                           int *__restrict__ tgt, int *__restrict__ counts)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if(a >= k.n) return;
  int cnt = 0;
  int base = FILL ? start[a] : 0;
  const int2 range = d.sys_range[d.sys_of[a]];
  // FP32 prefilter: a pair whose FP32 distance (units of c0 dt, error at most
  // eps_abs + eps_rel r) is at least hi + 2 (eps_abs + eps_rel hi) + 1e-3 has
  // delay >= hi for certain and skips the exact split; the lists are unchanged.
  const float4 pa32 = d.pd32[2 * a];
  const double cut = hi + 2.0 * (k.eps_abs + k.eps_rel * hi) + 1e-3;
  const float cut2 = static_cast<float>(cut * cut);
  const double4 pa64 = d.pos64[a];
  const PairK<float> K(k);
  // the exact split of the precision's pair kernels for a candidate b
  auto consider = [&](int b) {
    int n;
    if constexpr(F64) {
      const double4 pb = d.pos64[b];
      double x, ir, r2;
      n = split_f64(pa64.x - pb.x, pa64.y - pb.y, pa64.z - pb.z, k.cdt, k.icdt, x, ir, r2);
    } else {
      const float4 pb = d.pd32[2 * b];
      float x, ir;
      n = kRoundMagicBits - split_mixed(pa32.x - pb.x, pa32.y - pb.y, pa32.z - pb.z, K.eps_rel, K.lim0,
                                        kRoundMagic, pa64, d.pos64, b, K.cdt, K.icdt64, x, ir);
    }
    if(b != a && n >= lo && n < hi) {
      if(FILL) src[base + cnt] = b;
      ++cnt;
    }
  };
  auto far = [&](const float4 &pb) {
    const float ex = pa32.x - pb.x, ey = pa32.y - pb.y, ez = pa32.z - pb.z;
    return ex * ex + ey * ey + ez * ez >= cut2;
  };
  // eight source positions in flight per step (the loop is bound by their
  // load latency, one dot per thread), tested in source order
  constexpr int kAhead = 8;
  int b = range.x;
  for(; b + kAhead <= range.y; b += kAhead) {
    float4 q[kAhead];
#pragma unroll
    for(int j = 0; j < kAhead; ++j) q[j] = d.pd32[2 * (b + j)];
#pragma unroll
    for(int j = 0; j < kAhead; ++j)
      if(!far(q[j])) consider(b + j);
  }
  for(; b < range.y; ++b)
    if(!far(d.pd32[2 * b])) consider(b);
  if(!FILL) counts[a] = cnt;
  else if(tgt)
    for(int e = 0; e < cnt; ++e) tgt[base + e] = a;
}

// Lagrange weights at x on nodes 0..M and their first and second x-derivatives,
// as products (no division by x - m, so x on a node is regular).
__device__ void lagrange_products(double x, int M, double *l0, double *l1, double *l2)
{
  for(int b = 0; b <= M; ++b) {
    double den = 1;
    for(int m = 0; m <= M; ++m)
      if(m != b) den *= static_cast<double>(b - m);
    double v = 1, d1 = 0, d2 = 0;
    for(int m = 0; m <= M; ++m) {
      if(m == b) continue;
      const double f = x - m;
      d2 = d2 * f + 2 * d1;
      d1 = d1 * f + v;
      v = v * f;
    }
    l0[b] = v / den;
    l1[b] = d1 / den;
    l2[b] = d2 / den;
  }
}

// The reference's FP64 coefficients[pair][m] of target a <- source b
// (history_interaction.cpp:29-43 with lagrange_set.cpp:17-41 and
// rotating_green_function.cpp:3-24 / green_function.cpp:3-20), already
// multiplied by the prefactor -K e^{-i w r / c} (rotating) or K (lab); r > 0.
// Returns the integer delay.
__device__ int ref_pair_coefs(const Consts &k, double4 pa, double4 pb, double4 da, double4 db,
                              double2 *out)
{
  const double dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
  double x, ir, r2;
  const int n = split_f64(dx, dy, dz, k.cdt, k.icdt, x, ir, r2);
  const int M = k.order;
  double l0[kMaxOrder + 1], l1[kMaxOrder + 1], l2[kMaxOrder + 1];
  lagrange_products(x, M, l0, l1, l2);
  const double r = sqrt(r2);
  const double ba = da.x * dx + da.y * dy + da.z * dz;
  const double bb = db.x * dx + db.y * dy + db.z * dz;
  const double aab = da.x * db.x + da.y * db.y + da.z * db.z;
  const double bq = ba * bb / r2;
  const double T0 = k.c * k.c * (aab - 3 * bq) / (r2 * r);
  const double T1 = k.c * (aab - 3 * bq) / r2;
  const double T2 = (aab - bq) / r;
  const double w = k.omega;
  double2 ph = make_double2(1.0, 0.0);
  if(k.dyadic != QUEST_DYADIC_LAB) {
    double s, c;
    sincos(-w * r / k.c, &s, &c);
    ph = make_double2(-k.k_green * c, -k.k_green * s);
  } else {
    ph = make_double2(k.k_green, 0.0);
  }
  for(int m = 0; m <= M; ++m) {
    // time-derivative weights: w1 = -L'(x)/dt, w2 = L''(x)/dt^2
    const double w0 = l0[m], w1 = -l1[m] / k.dt, w2 = l2[m] / (k.dt * k.dt);
    double2 g;
    if(k.dyadic == QUEST_DYADIC_LAB) g = make_double2(T0 * w0 + T1 * w1 + T2 * w2, 0.0);
    else
      g = make_double2(T0 * w0 + T1 * w1 + T2 * (w2 - w * w * w0), T1 * w * w0 + T2 * 2.0 * w * w1);
    out[m] = cmul(ph, g);
  }
  return n;
}

// FP64 per-pair coefficients of the zero-delay list.
__global__ __launch_bounds__(kBlock) void near_coef_kernel(Consts k, DeviceState d, int total)
{
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if(e >= total) return;
  const int a = d.near_tgt[e], b = d.near_src[e];
  double2 cf[kMaxOrder + 1];
  d.near_n[e] = ref_pair_coefs(k, d.pos64[a], d.pos64[b], d.dip64[a], d.dip64[b], cf);
  double2 *out = d.near_coef + static_cast<size_t>(e) * (k.order + 1);
  for(int m = 0; m <= k.order; ++m) out[m] = cf[m];
}

// Coefficient table of one system (the reference's build_coefficient_table,
// history_interaction.cpp:21-45, for every ordered pair): pairs with integer
// delay >= 1 get their coefficients and delay; the self pair and the
// zero-delay pairs (evaluated from the near list) get zeros.
__global__ __launch_bounds__(kBlock) void table_build_kernel(Consts k, DeviceState d, int2 range,
                                                             long long toff)
{
  const long long ns = range.y - range.x;
  const long long total = ns * ns;
  const size_t P = static_cast<size_t>(k.table_plane);
  for(long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
      i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int b = range.x + static_cast<int>(i / ns), a = range.x + static_cast<int>(i % ns);
    const size_t e = static_cast<size_t>(toff + i);
    double2 cf[kMaxOrder + 1];
    int n = 0;
    if(a != b) {
      const double4 pa = d.pos64[a], pb = d.pos64[b];
      double x, ir, r2;
      n = split_f64(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z, k.cdt, k.icdt, x, ir, r2);
      if(n >= 1) n = ref_pair_coefs(k, pa, pb, d.dip64[a], d.dip64[b], cf);
    }
    if(n < 1) {
      n = 0;
      for(int m = 0; m <= k.order; ++m) cf[m] = make_double2(0.0, 0.0);
    }
    for(int m = 0; m <= k.order; ++m) d.table[m * P + e] = cf[m];
    d.table_n[e] = static_cast<int16_t>(n);
  }
}

// Diagnostic: integer delays of unordered pairs [first, first + count) in the
// reference's lower-triangle order (history_interaction.cpp:80-86; row = the
// pair's source, col < row its observer), from the delay split the pair
// kernel of the precision uses: split_l64 (F64) or split_mixed, the FP32
// estimate with the FP64 split where it lies near an integer.
template <bool F64>
__global__ void pair_delays_kernel(Consts k, DeviceState d, int32_t *out, int64_t first, int64_t count)
{
  const PairK<float> K(k);
  for(int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < count;
      i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t idx = first + i;
    int64_t row = static_cast<int64_t>(floor((sqrt(static_cast<double>(1 + 8 * idx)) + 1) / 2.0));
    while(row * (row - 1) / 2 > idx) --row;
    while((row + 1) * row / 2 <= idx) ++row;
    const int64_t col = idx - row * (row - 1) / 2;
    const int rs = d.perm_dev[row], cs = d.perm_dev[col];
    if constexpr(F64) {
      // observer cs as the target a, source rs as b, as in the pair kernel
      const double4 pa = d.pd64[2 * cs], pb = d.pd64[2 * rs];
      double xc, ir;
      out[i] = split_l64(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z, d.pos64[cs], d.pos64, rs, k.cdt, xc, ir);
    } else {
      // observer cs as the target a, source rs as b, as in the pair kernel
      const float4 pa = d.pd32[2 * cs], pb = d.pd32[2 * rs];
      float x, ir;
      out[i] = kRoundMagicBits - split_mixed(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z, K.eps_rel, K.lim0,
                                             kRoundMagic, d.pos64[cs], d.pos64, rs, K.cdt, K.icdt64, x, ir);
    }
  }
}

inline dim3 dot_grid(int n) { return dim3((n + kBlock - 1) / kBlock); }

template <int M, bool LAB, bool DYN, bool GEN>
void launch_pair_mld(const Consts &k, DeviceState &d, int precision, hipStream_t s)
{
  const dim3 grid(k.grid), block(kBlock);
  if(precision == QUEST_PREC_FP64)
    hipLaunchKernelGGL((pair_past_f64<M, LAB, DYN, GEN>), grid, block, 0, s, d.pd64, d.pos64,
                       static_cast<const double2 *>(d.ring), d.work, d.work_row, d.step, d.partial, k);
  else {
    // FP32 ring layout (see ring_index): (M + 1) / 2 float4 planes, then the c_M plane
    const size_t E = static_cast<size_t>(k.n) * k.ring_slots;
    const float4 *r4 = static_cast<const float4 *>(d.ring);
    RingPlanesF32 planes{};
    for(int i = 0; i < (M + 1) / 2; ++i) planes.pair[i] = r4 + i * E;
    planes.last = reinterpret_cast<const float2 *>(r4 + ((M + 1) / 2) * E);
    if(k.compensated) {
      if(k.pipelined)
        hipLaunchKernelGGL((pair_past_f32<M, LAB, DYN, GEN, true, true>), grid, block, 0, s, d.pd32, d.pos64,
                           planes, d.work, d.work_row, d.step, d.partial, k);
      else
        hipLaunchKernelGGL((pair_past_f32<M, LAB, DYN, GEN, false, true>), grid, block, 0, s, d.pd32, d.pos64,
                           planes, d.work, d.work_row, d.step, d.partial, k);
    } else if(k.pipelined)
      hipLaunchKernelGGL((pair_past_f32<M, LAB, DYN, GEN, true, false>), grid, block, 0, s, d.pd32, d.pos64,
                         planes, d.work, d.work_row, d.step, d.partial, k);
    else
      hipLaunchKernelGGL((pair_past_f32<M, LAB, DYN, GEN, false, false>), grid, block, 0, s, d.pd32, d.pos64,
                         planes, d.work, d.work_row, d.step, d.partial, k);
  }
}

template <typename R, int M, bool LAB, bool DYN, bool GEN>
void launch_block_mld(const Consts &k, DeviceState &d, hipStream_t s)
{
  using R2 = typename V2<R>::type;
  using R4 = typename V4<R>::type;
  const R4 *pd;
  if constexpr(sizeof(R) == sizeof(double))
    pd = d.pd64;
  else
    pd = d.pd32;
  const R2 *h = static_cast<const R2 *>(d.hist), *rt = static_cast<const R2 *>(d.rot);
  if(k.block_steps == 16)
    hipLaunchKernelGGL((pair_block<R, M, 16, LAB, DYN, GEN>), dim3(k.grid), dim3(kBlock), 0, s, pd, d.pos64, h, rt,
                       d.work, d.work_row, d.step, d.partial, k);
  else
    hipLaunchKernelGGL((pair_block<R, M, 8, LAB, DYN, GEN>), dim3(k.grid), dim3(kBlock), 0, s, pd, d.pos64, h, rt,
                       d.work, d.work_row, d.step, d.partial, k);
}

template <typename R, int M>
hipError_t launch_block_m(const Consts &k, DeviceState &d, hipStream_t s)
{
  if(k.dyadic == QUEST_DYADIC_LAB)
    k.dyn_items ? launch_block_mld<R, M, true, true, false>(k, d, s) : launch_block_mld<R, M, true, false, false>(k, d, s);
  else if(k.trig_poly)
    k.dyn_items ? launch_block_mld<R, M, false, true, false>(k, d, s)
                : launch_block_mld<R, M, false, false, false>(k, d, s);
  else
    k.dyn_items ? launch_block_mld<R, M, false, true, true>(k, d, s) : launch_block_mld<R, M, false, false, true>(k, d, s);
  return hipGetLastError();
}

template <typename R, int M, int KS>
hipError_t launch_list_mk(const Consts &k, DeviceState &d, const int *start, const int *src, double2 *out,
                          hipStream_t s)
{
  using R2 = typename V2<R>::type;
  using R4 = typename V4<R>::type;
  const R4 *pd;
  if constexpr(sizeof(R) == sizeof(double))
    pd = d.pd64;
  else
    pd = d.pd32;
  const R2 *h = static_cast<const R2 *>(d.hist), *rt = static_cast<const R2 *>(d.rot);
  const dim3 grid(static_cast<unsigned>((static_cast<size_t>(k.n) * kListLanes + kBlock - 1) / kBlock));
  if(k.dyadic == QUEST_DYADIC_LAB)
    hipLaunchKernelGGL((pair_list<R, M, KS, true, false>), grid, dim3(kBlock), 0, s, pd, d.pos64, h, rt, start, src,
                       d.step, out, k);
  else if(k.trig_poly)
    hipLaunchKernelGGL((pair_list<R, M, KS, false, false>), grid, dim3(kBlock), 0, s, pd, d.pos64, h, rt, start, src,
                       d.step, out, k);
  else
    hipLaunchKernelGGL((pair_list<R, M, KS, false, true>), grid, dim3(kBlock), 0, s, pd, d.pos64, h, rt, start, src,
                       d.step, out, k);
  return hipGetLastError();
}

template <typename R, int M>
hipError_t launch_list_m(const Consts &k, DeviceState &d, int list, hipStream_t s)
{
  if(list < 0) return launch_list_mk<R, M, 1>(k, d, d.short_start, d.short_src, d.short_sum, s);
  static_assert(kShortListEnd == 4 && kBlockStepsShort == 8, "window lists of 4 and 8 steps");
  return k.win_steps[list] == kBlockStepsShort
             ? launch_list_mk<R, M, kBlockStepsShort>(k, d, d.win_start[list], d.win_src[list], d.win_sum[list], s)
             : launch_list_mk<R, M, kShortListEnd>(k, d, d.win_start[list], d.win_src[list], d.win_sum[list], s);
}

template <int M>
hipError_t launch_pair_m(const Consts &k, DeviceState &d, int precision, hipStream_t s)
{
  const bool lab = k.dyadic == QUEST_DYADIC_LAB;
  if(lab)  // no rotation in the lab frame
    k.dyn_items ? launch_pair_mld<M, true, true, false>(k, d, precision, s)
                : launch_pair_mld<M, true, false, false>(k, d, precision, s);
  else if(k.trig_poly)
    k.dyn_items ? launch_pair_mld<M, false, true, false>(k, d, precision, s)
                : launch_pair_mld<M, false, false, false>(k, d, precision, s);
  else
    k.dyn_items ? launch_pair_mld<M, false, true, true>(k, d, precision, s)
                : launch_pair_mld<M, false, false, true>(k, d, precision, s);
  return hipGetLastError();
}

template <int M>
hipError_t launch_table_m(const Consts &k, DeviceState &d, hipStream_t s)
{
  const dim3 grid(k.grid), block(kBlock);
  if(k.dyadic == QUEST_DYADIC_LAB)
    hipLaunchKernelGGL((pair_table_f64<M, true>), grid, block, 0, s, d.table, d.table_n, d.samp, d.work,
                       d.work_row, d.step, d.sys_of, d.sys_range, d.sys_toff, d.partial, k);
  else
    hipLaunchKernelGGL((pair_table_f64<M, false>), grid, block, 0, s, d.table, d.table_n, d.samp, d.work,
                       d.work_row, d.step, d.sys_of, d.sys_range, d.sys_toff, d.partial, k);
  return hipGetLastError();
}

template <int M, int K>
hipError_t block_occupancy(int *nb, int precision, bool gen)
{
  if(precision == QUEST_PREC_FP64)
    return gen ? hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, pair_block<double, M, K, false, false, true>, kBlock, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, pair_block<double, M, K, false, false, false>, kBlock, 0);
  return gen ? hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, pair_block<float, M, K, false, false, true>, kBlock, 0)
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(nb, pair_block<float, M, K, false, false, false>, kBlock, 0);
}

template <int M>
int blocks_per_cu_m(int precision, bool table, bool gen, bool pf, bool comp, int blocked)
{
  int nb = 0;
  hipError_t e;
  if(blocked) {
    e = blocked == 16 ? block_occupancy<M, 16>(&nb, precision, gen) : block_occupancy<M, 8>(&nb, precision, gen);
    return e == hipSuccess && nb > 0 ? nb : 1;
  }
  // the FP32 variant of the context's loop (plain / pipelined, block /
  // compensated sums, short series / library sincos)
  auto f32 = [&](auto pf_c, auto comp_c) {
    constexpr bool P = decltype(pf_c)::value, C = decltype(comp_c)::value;
    return gen ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pair_past_f32<M, false, false, true, P, C>, kBlock, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pair_past_f32<M, false, false, false, P, C>, kBlock,
                                                              0);
  };
  using T = std::true_type;
  using F = std::false_type;
  if(table)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pair_table_f64<M, false>, kBlock, 0);
  else if(precision == QUEST_PREC_FP64)
    e = gen ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pair_past_f64<M, false, false, true>, kBlock, 0)
            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, pair_past_f64<M, false, false, false>, kBlock, 0);
  else if(pf)
    e = comp ? f32(T{}, T{}) : f32(T{}, F{});
  else
    e = comp ? f32(F{}, T{}) : f32(F{}, F{});
  return e == hipSuccess && nb > 0 ? nb : 1;
}

}  // namespace

// ---------------------------------------------------------------- launchers

hipError_t launch_fill_ground_state(const Consts &k, DeviceState &d, hipStream_t s)
{
  const size_t total = static_cast<size_t>(k.state_slots) * k.n;
  const int blocks = static_cast<int>(std::min<size_t>((total + kBlock - 1) / kBlock, 4096));
  hipLaunchKernelGGL(fill_ground_state_kernel, dim3(blocks), dim3(kBlock), 0, s, k, d.y0, d.y1,
                     d.d0, d.d1, d.bad_step);
  return hipGetLastError();
}

hipError_t launch_clear_ring(const Consts &k, DeviceState &d, size_t ring_bytes, hipStream_t s)
{
  (void)k;
  const size_t n16 = ring_bytes / 16;
  const int blocks = static_cast<int>(std::min<size_t>((n16 + kBlock - 1) / kBlock, 8192));
  hipLaunchKernelGGL(clear_words_kernel, dim3(blocks > 0 ? blocks : 1), dim3(kBlock), 0, s,
                     static_cast<uint4 *>(d.ring), n16);
  return hipGetLastError();
}

hipError_t launch_clear(void *p, size_t bytes, hipStream_t s)
{
  const size_t n16 = bytes / 16;
  const int blocks = static_cast<int>(std::min<size_t>((n16 + kBlock - 1) / kBlock, 8192));
  hipLaunchKernelGGL(clear_words_kernel, dim3(blocks > 0 ? blocks : 1), dim3(kBlock), 0, s,
                     static_cast<uint4 *>(p), n16);
  return hipGetLastError();
}

hipError_t launch_table_build(const Consts &k, DeviceState &d, int2 range, long long toff, hipStream_t s)
{
  const long long ns = range.y - range.x;
  const long long total = ns * ns;
  if(total == 0) return hipSuccess;
  const int blocks = static_cast<int>(std::min<long long>((total + kBlock - 1) / kBlock, 65536));
  hipLaunchKernelGGL(table_build_kernel, dim3(blocks), dim3(kBlock), 0, s, k, d, range, toff);
  return hipGetLastError();
}

hipError_t launch_set_step(DeviceState &d, int step, hipStream_t s)
{
  hipLaunchKernelGGL(set_step_kernel, dim3(1), dim3(1), 0, s, d.step, step);
  return hipGetLastError();
}

hipError_t launch_advance_step(DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(advance_step_kernel, dim3(1), dim3(1), 0, s, d.step);
  return hipGetLastError();
}

hipError_t launch_pair_past(const Consts &k, DeviceState &d, int precision, hipStream_t s)
{
  if(k.n_work == 0) return hipSuccess;
  if(k.pair_table) {
    switch(k.order) {
    case 1: return launch_table_m<1>(k, d, s);
    case 2: return launch_table_m<2>(k, d, s);
    case 3: return launch_table_m<3>(k, d, s);
    case 4: return launch_table_m<4>(k, d, s);
    case 5: return launch_table_m<5>(k, d, s);
    case 6: return launch_table_m<6>(k, d, s);
    default: return hipErrorInvalidValue;
    }
  }
  if(k.pair_blocked) {
    const bool f64 = precision == QUEST_PREC_FP64;
    switch(k.order) {
    case 1: return f64 ? launch_block_m<double, 1>(k, d, s) : launch_block_m<float, 1>(k, d, s);
    case 2: return f64 ? launch_block_m<double, 2>(k, d, s) : launch_block_m<float, 2>(k, d, s);
    case 3: return f64 ? launch_block_m<double, 3>(k, d, s) : launch_block_m<float, 3>(k, d, s);
    case 4: return f64 ? launch_block_m<double, 4>(k, d, s) : launch_block_m<float, 4>(k, d, s);
    case 5: return f64 ? launch_block_m<double, 5>(k, d, s) : launch_block_m<float, 5>(k, d, s);
    case 6: return f64 ? launch_block_m<double, 6>(k, d, s) : launch_block_m<float, 6>(k, d, s);
    default: return hipErrorInvalidValue;
    }
  }
  switch(k.order) {
  case 1: return launch_pair_m<1>(k, d, precision, s);
  case 2: return launch_pair_m<2>(k, d, precision, s);
  case 3: return launch_pair_m<3>(k, d, precision, s);
  case 4: return launch_pair_m<4>(k, d, precision, s);
  case 5: return launch_pair_m<5>(k, d, precision, s);
  case 6: return launch_pair_m<6>(k, d, precision, s);
  default: return hipErrorInvalidValue;
  }
}

int pair_blocks_per_cu(const Consts &k, int precision)
{
  const bool t = k.pair_table != 0;
  const bool g = k.dyadic != QUEST_DYADIC_LAB && !k.trig_poly;
  const bool pf = k.pipelined != 0, cp = k.compensated != 0;
  const int bl = k.pair_blocked ? k.block_steps : 0;
  switch(k.order) {
  case 1: return blocks_per_cu_m<1>(precision, t, g, pf, cp, bl);
  case 2: return blocks_per_cu_m<2>(precision, t, g, pf, cp, bl);
  case 3: return blocks_per_cu_m<3>(precision, t, g, pf, cp, bl);
  case 4: return blocks_per_cu_m<4>(precision, t, g, pf, cp, bl);
  case 5: return blocks_per_cu_m<5>(precision, t, g, pf, cp, bl);
  default: return blocks_per_cu_m<6>(precision, t, g, pf, cp, bl);
  }
}

hipError_t launch_pair_list(const Consts &k, DeviceState &d, int precision, int list, hipStream_t s)
{
  const bool f64 = precision == QUEST_PREC_FP64;
  switch(k.order) {
  case 1: return f64 ? launch_list_m<double, 1>(k, d, list, s) : launch_list_m<float, 1>(k, d, list, s);
  case 2: return f64 ? launch_list_m<double, 2>(k, d, list, s) : launch_list_m<float, 2>(k, d, list, s);
  case 3: return f64 ? launch_list_m<double, 3>(k, d, list, s) : launch_list_m<float, 3>(k, d, list, s);
  case 4: return f64 ? launch_list_m<double, 4>(k, d, list, s) : launch_list_m<float, 4>(k, d, list, s);
  case 5: return f64 ? launch_list_m<double, 5>(k, d, list, s) : launch_list_m<float, 5>(k, d, list, s);
  case 6: return f64 ? launch_list_m<double, 6>(k, d, list, s) : launch_list_m<float, 6>(k, d, list, s);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pc_step(const Consts &k, DeviceState &d, int precision, double future_coef, int window,
                          int block, int blocks, int sync_blocks, hipStream_t s)
{
  if(precision == QUEST_PREC_FP64)
    hipLaunchKernelGGL(pc_step_kernel<double2>, dim3(blocks), dim3(block), 0, s, k, d,
                       static_cast<double2 *>(k.pair_blocked ? d.hist : d.ring), future_coef, window, sync_blocks);
  else
    hipLaunchKernelGGL(pc_step_kernel<float2>, dim3(blocks), dim3(block), 0, s, k, d,
                       static_cast<float2 *>(k.pair_blocked ? d.hist : d.ring), future_coef, window, sync_blocks);
  return hipGetLastError();
}

hipError_t launch_pc_begin(const Consts &k, DeviceState &d, int window, hipStream_t s)
{
  hipLaunchKernelGGL(pc_begin_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k, d, window);
  return hipGetLastError();
}

hipError_t launch_pc_iter(const Consts &k, DeviceState &d, int precision, double future_coef,
                          int iter, hipStream_t s)
{
  if(precision == QUEST_PREC_FP64)
    hipLaunchKernelGGL(pc_iter_kernel<double2>, dot_grid(k.n), dim3(kBlock), 0, s, k, d,
                       static_cast<double2 *>(k.pair_blocked ? d.hist : d.ring), future_coef, iter);
  else
    hipLaunchKernelGGL(pc_iter_kernel<float2>, dot_grid(k.n), dim3(kBlock), 0, s, k, d,
                       static_cast<float2 *>(k.pair_blocked ? d.hist : d.ring), future_coef, iter);
  return hipGetLastError();
}

hipError_t launch_hist_commit(const Consts &k, DeviceState &d, int precision, int step, hipStream_t s)
{
  if(precision == QUEST_PREC_FP64)
    hipLaunchKernelGGL(hist_commit_kernel<double2>, dot_grid(k.n), dim3(kBlock), 0, s, k, d,
                       static_cast<double2 *>(k.pair_blocked ? d.hist : d.ring), step);
  else
    hipLaunchKernelGGL(hist_commit_kernel<float2>, dot_grid(k.n), dim3(kBlock), 0, s, k, d,
                       static_cast<float2 *>(k.pair_blocked ? d.hist : d.ring), step);
  return hipGetLastError();
}

hipError_t launch_hist_reduce(const Consts &k, DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(hist_reduce_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k, d);
  return hipGetLastError();
}

hipError_t launch_hist_near(const Consts &k, DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(hist_near_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k, d);
  return hipGetLastError();
}

hipError_t launch_permute_in(const Consts &k, DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(permute_in_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k.n, d.io, d.perm_dev, d.cur1);
  return hipGetLastError();
}

hipError_t launch_permute_out(const Consts &k, DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(permute_out_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k.n, d.out, d.perm_dev, d.io);
  return hipGetLastError();
}

hipError_t launch_pulse(const Consts &k, DeviceState &d, hipStream_t s)
{
  hipLaunchKernelGGL(pulse_kernel, dot_grid(k.n), dim3(kBlock), 0, s, k, d);
  return hipGetLastError();
}

hipError_t launch_delay_list(const Consts &k, DeviceState &d, int precision, int lo, int hi, bool fill,
                             const int *start, int *src, int *tgt, int *counts, hipStream_t s)
{
  const bool f64 = precision == QUEST_PREC_FP64;
  if(f64 && fill)
    hipLaunchKernelGGL((near_scan_kernel<true, true>), dot_grid(k.n), dim3(kBlock), 0, s, k, d, lo, hi, start,
                       src, tgt, counts);
  else if(f64)
    hipLaunchKernelGGL((near_scan_kernel<true, false>), dot_grid(k.n), dim3(kBlock), 0, s, k, d, lo, hi, start,
                       src, tgt, counts);
  else if(fill)
    hipLaunchKernelGGL((near_scan_kernel<false, true>), dot_grid(k.n), dim3(kBlock), 0, s, k, d, lo, hi, start,
                       src, tgt, counts);
  else
    hipLaunchKernelGGL((near_scan_kernel<false, false>), dot_grid(k.n), dim3(kBlock), 0, s, k, d, lo, hi, start,
                       src, tgt, counts);
  return hipGetLastError();
}

hipError_t launch_near_count(const Consts &k, DeviceState &d, int precision, hipStream_t s)
{
  return launch_delay_list(k, d, precision, 0, 1, false, nullptr, nullptr, nullptr, d.counts, s);
}

hipError_t launch_near_fill(const Consts &k, DeviceState &d, int precision, hipStream_t s)
{
  return launch_delay_list(k, d, precision, 0, 1, true, d.near_start, d.near_src, d.near_tgt, d.counts, s);
}

hipError_t launch_near_coef(const Consts &k, DeviceState &d, int near_total, hipStream_t s)
{
  if(near_total <= 0) return hipSuccess;
  hipLaunchKernelGGL(near_coef_kernel, dim3((near_total + kBlock - 1) / kBlock), dim3(kBlock), 0,
                     s, k, d, near_total);
  return hipGetLastError();
}

hipError_t launch_pair_delays(const Consts &k, DeviceState &d, int precision, int32_t *out, int64_t first,
                              int64_t count, hipStream_t s)
{
  const int blocks = static_cast<int>(std::min<int64_t>((count + kBlock - 1) / kBlock, 16384));
  if(precision == QUEST_PREC_FP64)
    hipLaunchKernelGGL(pair_delays_kernel<true>, dim3(blocks > 0 ? blocks : 1), dim3(kBlock), 0, s, k, d, out,
                       first, count);
  else
    hipLaunchKernelGGL(pair_delays_kernel<false>, dim3(blocks > 0 ? blocks : 1), dim3(kBlock), 0, s, k, d, out,
                       first, count);
  return hipGetLastError();
}

}  // namespace quest
