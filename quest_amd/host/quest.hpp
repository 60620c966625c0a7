// quest.hpp -- C++17 mirror of QuEST's host interface over the C ABI of
// libquest_amd.so (include/quest_amd.h).
//
// Same class names, constructor shapes, argument meaning and exceptions as the
// reference, without Eigen or Boost:
//   QuantumDot, DotVector, import_dots            src/quantum_dot.{h,cpp}
//   Pulse, read_pulse_config                       src/pulse.{h,cpp}
//   History::{soltype, HistoryArray,
//             make_shared_history, isfinite,
//             write_history}                       src/history.{h,cpp}
//   GreenFunction::{Dyadic, RotatingDyadic}        src/interactions/*green_function.h
//   Interaction, PulseInteraction,
//   HistoryInteraction                             src/interactions/*.h
//   PredictorCorrector::{Weights, Integrator}      src/integrator.{h,cpp}
//   Configuration config                           src/configuration.h
//
// HistoryInteraction and PulseInteraction evaluate on the GPU (one quest_ctx
// per interaction); PredictorCorrector::Integrator is the reference's host
// loop, so swapping these classes in leaves the integrator untouched.
// DeviceIntegrator runs the whole predictor-corrector loop on the device.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <complex>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/quest_amd.h"

using cmplx = std::complex<double>;
constexpr cmplx iu(0, 1);

// src/configuration.h:12-19 (values only; file parsing lives in quest_main.cpp)
struct Configuration {
  int num_particles = 0;
  int num_timesteps = 0;
  double c0 = 0, hbar = 0, mu0 = 0, omega = 0;
  double dt = 0, total_time = 0;
  int interpolation_order = 0;
  std::string qd_path = "dots.cfg", pulse_path = "pulse.cfg";
  int precision = QUEST_PREC_FP64;  // arithmetic of the pair sum on the device
  // QUEST_PAIR_TIME_BLOCKED (sums of 8 or 16 steps per pass; the fastest), QUEST_PAIR_ON_THE_FLY
  // (every pair every step) or QUEST_PAIR_TABLE (the reference's coefficient table, FP64)
  int pair_method = QUEST_PAIR_TIME_BLOCKED;
  int accumulation = QUEST_ACC_BLOCK;       // mixed precision's 64-source block sums (or compensated)
  bool autotune = false;                    // DeviceIntegrator: quest_autotune before the first step
  int device = 0;
};
inline Configuration config;

struct Vec3 {
  double x = 0, y = 0, z = 0;
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double dot(const Vec3 &o) const { return (x * o.x + y * o.y) + z * o.z; }
  double norm() const { return std::sqrt(dot(*this)); }
};

namespace quest_detail {
inline void check(int rc, const quest_ctx *ctx)
{
  if(rc == QUEST_OK) return;
  const std::string msg = quest_last_error(ctx);
  if(rc == QUEST_ERR_DOMAIN) throw std::domain_error(msg);
  throw std::runtime_error(msg);
}
}  // namespace quest_detail

// ------------------------------------------------------------ quantum_dot.h
using matrix_elements = std::array<cmplx, 2>;

class QuantumDot {
 public:
  QuantumDot() = default;
  QuantumDot(const Vec3 &pos, double freq, const std::pair<double, double> &damping, const Vec3 &dip)
      : pos(pos), freq(freq), damping(damping), dip(dip) {}

  // quantum_dot.cpp:10-19
  matrix_elements liouville_rhs(const matrix_elements &rho, const cmplx rabi,
                                const double laser_freq) const
  {
    const cmplx m0 = -iu * (rabi * std::conj(rho[1]) - std::conj(rabi) * rho[1]) -
                     (rho[0] - 1.0) / damping.first;
    const cmplx m1 = -iu * (rabi * (1.0 - 2.0 * rho[0]) + rho[1] * (laser_freq - freq)) -
                     rho[1] / damping.second;
    return {m0, m1};
  }

  const Vec3 &position() const { return pos; }
  const Vec3 &dipole() const { return dip; }
  double frequency() const { return freq; }
  const std::pair<double, double> &decay() const { return damping; }

  friend std::ostream &operator<<(std::ostream &os, const QuantumDot &qd)
  {
    os << qd.pos.x << " " << qd.pos.y << " " << qd.pos.z << " " << qd.freq << " "
       << qd.damping.first << " " << qd.damping.second << " " << qd.dip.x << " " << qd.dip.y
       << " " << qd.dip.z;
    return os;
  }
  // quantum_dot.cpp:33-38: x y z omega0 T1 T2 dx dy dz
  friend std::istream &operator>>(std::istream &is, QuantumDot &qd)
  {
    is >> qd.pos.x >> qd.pos.y >> qd.pos.z >> qd.freq >> qd.damping.first >>
        qd.damping.second >> qd.dip.x >> qd.dip.y >> qd.dip.z;
    return is;
  }

 private:
  Vec3 pos;
  double freq = 0;
  std::pair<double, double> damping{1, 1};
  Vec3 dip;
};

using DotVector = std::vector<QuantumDot>;

// quantum_dot.cpp:40-47
inline DotVector import_dots(const std::string &fname)
{
  std::ifstream ifs(fname);
  if(!ifs) throw std::runtime_error("Could not open " + fname);
  std::istream_iterator<QuantumDot> in_iter(ifs), eof;
  return DotVector(in_iter, eof);
}

// quantum_dot.cpp:21-24
inline Vec3 separation(const QuantumDot &d1, const QuantumDot &d2)
{
  return {d2.position().x - d1.position().x, d2.position().y - d1.position().y,
          d2.position().z - d1.position().z};
}

using rhs_func = std::function<matrix_elements(const matrix_elements &, const cmplx)>;

// quantum_dot.cpp:52-64
inline std::vector<rhs_func> rhs_functions(const DotVector &dots, const double laser_frequency)
{
  std::vector<rhs_func> funcs;
  funcs.reserve(dots.size());
  for(const QuantumDot &d : dots)
    funcs.push_back([d, laser_frequency](const matrix_elements &rho, const cmplx rabi) {
      return d.liouville_rhs(rho, rabi, laser_frequency);
    });
  return funcs;
}

// ------------------------------------------------------------------ pulse.h
class Pulse {
 public:
  Pulse() = default;
  // pulse.cpp:3-13 (the constructor normalises the polarisation)
  Pulse(double amplitude, double delay, double width, double freq, const Vec3 &wavevector,
        const Vec3 &polarization)
      : amplitude(amplitude), delay(delay), width(width), freq(freq), wavevector(wavevector)
  {
    const double n = polarization.norm();
    this->polarization = {polarization.x / n, polarization.y / n, polarization.z / n};
  }

  // pulse.cpp:15-20
  Vec3 operator()(const Vec3 &r, const double t) const
  {
    const double arg = wavevector.dot(r) - freq * (t - delay);
    const double g = std::exp(-std::pow(arg / width, 2) / 2);
    return {(amplitude / 2 * polarization.x) * g, (amplitude / 2 * polarization.y) * g,
            (amplitude / 2 * polarization.z) * g};
  }

  quest_pulse to_c() const
  {
    return {amplitude, delay, width, freq, {wavevector.x, wavevector.y, wavevector.z},
            {polarization.x, polarization.y, polarization.z}};
  }

  friend std::ostream &operator<<(std::ostream &os, const Pulse &p)
  {
    os << p.amplitude << " " << p.delay << " " << p.width << " " << p.freq << " " << p.wavevector.x
       << " " << p.wavevector.y << " " << p.wavevector.z << " " << p.polarization.x << " "
       << p.polarization.y << " " << p.polarization.z;
    return os;
  }
  // pulse.cpp:29-35 (no normalisation on this path, as in the reference)
  friend std::istream &operator>>(std::istream &is, Pulse &p)
  {
    is >> p.amplitude >> p.delay >> p.width >> p.freq >> p.wavevector.x >> p.wavevector.y >>
        p.wavevector.z >> p.polarization.x >> p.polarization.y >> p.polarization.z;
    return is;
  }

 private:
  double amplitude = 0, delay = 0, width = 1, freq = 0;
  Vec3 wavevector, polarization;
};

// pulse.cpp:37-46
inline Pulse read_pulse_config(const std::string &fname)
{
  std::ifstream ifs(fname);
  if(!ifs) throw std::runtime_error("Could not open " + fname);
  Pulse p;
  ifs >> p;
  return p;
}

// Every row of a pulse file, in order (an extension: the reference reads the
// first row only, pulse.cpp:37-46, though README.md:64 speaks of "pulse(s)").
// At most QUEST_MAX_PULSES rows feed the device path.
inline std::vector<Pulse> read_pulse_configs(const std::string &fname)
{
  std::ifstream ifs(fname);
  if(!ifs) throw std::runtime_error("Could not open " + fname);
  std::vector<Pulse> out;
  std::string line;
  while(std::getline(ifs, line)) {
    if(line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream is(line);
    Pulse p;
    if(is >> p) out.push_back(p);
  }
  return out;
}

// ---------------------------------------------------------------- history.h
namespace History {
using soltype = std::array<cmplx, 2>;

// boost::multi_array<soltype, 3>[N][-window, T)[2] (history.h:13-14): index as
// h[dot][t][0 state | 1 derivative] with t from -window.
class HistoryArray {
 public:
  HistoryArray(int n, int window, int n_steps)
      : n_(n), window_(window), len_(window + n_steps),
        data_(static_cast<size_t>(n) * (window + n_steps) * 2) {}

  struct TimeView {
    soltype *p;
    soltype &operator[](int k) const { return p[k]; }
  };
  struct DotView {
    soltype *p;
    int window;
    TimeView operator[](int t) const { return {p + static_cast<ptrdiff_t>(t + window) * 2}; }
  };
  struct ConstTimeView {
    const soltype *p;
    const soltype &operator[](int k) const { return p[k]; }
  };
  struct ConstDotView {
    const soltype *p;
    int window;
    ConstTimeView operator[](int t) const { return {p + static_cast<ptrdiff_t>(t + window) * 2}; }
  };
  DotView operator[](int dot) { return {data_.data() + static_cast<size_t>(dot) * len_ * 2, window_}; }
  ConstDotView operator[](int dot) const
  {
    return {data_.data() + static_cast<size_t>(dot) * len_ * 2, window_};
  }
  int num_dots() const { return n_; }
  int time_base() const { return -window_; }   // index_bases()[1]
  int time_extent() const { return len_; }     // shape()[1]

 private:
  int n_, window_, len_;
  std::vector<soltype> data_;
};

// history.cpp:3-9
inline std::shared_ptr<HistoryArray> make_shared_history(int num_particles, int window, int num_timesteps)
{
  return std::make_shared<HistoryArray>(num_particles, window, num_timesteps);
}

// history.cpp:11-15
inline bool isfinite(const soltype &sol)
{
  return std::isfinite(sol[0].real()) && std::isfinite(sol[0].imag()) &&
         std::isfinite(sol[1].real()) && std::isfinite(sol[1].imag());
}

// history.cpp:17-34: one line per step, each dot's (rho00, rho01) row as
// "(re,im) (re,im) " at setprecision(12) scientific.
inline void write_history(const std::shared_ptr<const HistoryArray> &history,
                          const std::string &filename, const int n = 0)
{
  std::ofstream output(filename);
  output << std::setprecision(12) << std::scientific;
  const int max_t = (n != 0) ? n : history->time_extent() + history->time_base();
  for(int t = 0; t < max_t; ++t) {
    for(int sol_idx = 0; sol_idx < history->num_dots(); ++sol_idx) {
      const soltype &s = (*history)[sol_idx][t][0];
      output << s[0] << " " << s[1] << " ";
    }
    output << std::endl;
  }
}
// Output of long or large runs without the O(N T) text history (an addition;
// the reference writes every dot at every step, history.cpp:17-34): every
// `every`-th step's (rho00, rho01) of a set of dots, as text in output.dat's
// row format (one line per recorded step, the listed dots only) or as binary
// records.  Binary layout, little-endian:
//   char[8] "QUESTOUT"; uint32 version (1), n_dots (dots in the system),
//   n_sel (dots recorded), every, n_records, 0; double dt;
//   int32 dots[n_sel] (row indices in dots.cfg, ascending);
//   n_records x { int32 step, int32 0, double [n_sel][4] (re, im rho00;
//   re, im rho01) }
// step t is output.dat's line t: the state after step t (t = every - 1,
// 2 every - 1, ...).
struct SampledOutput {
  std::string path;
  bool binary = false;
  int every = 1;
  std::vector<int> dots;  // ascending; all dots when the subsample is not asked for
};

// `count` distinct dots of n chosen by a seeded partial Fisher-Yates shuffle
// (splitmix64 stream), returned in ascending order: the same set on every
// platform for a given (n, count, seed).
inline std::vector<int> subsample_dots(int n, int count, uint64_t seed)
{
  if(count < 1 || count > n) throw std::runtime_error("subsample count must be in 1..num_particles");
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  uint64_t x = seed;
  auto next = [&x]() {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  for(int i = 0; i < count; ++i) {
    const uint64_t span = static_cast<uint64_t>(n - i);
    const int j = i + static_cast<int>(next() % span);
    std::swap(idx[i], idx[j]);
  }
  std::vector<int> out(idx.begin(), idx.begin() + count);
  std::sort(out.begin(), out.end());
  return out;
}

class SampledWriter {
 public:
  SampledWriter(const SampledOutput &spec, int n_dots, int num_timesteps, double dt) : spec_(spec)
  {
    if(spec_.every < 1) throw std::runtime_error("--every must be >= 1");
    if(spec_.dots.empty()) {
      spec_.dots.resize(n_dots);
      std::iota(spec_.dots.begin(), spec_.dots.end(), 0);
    }
    for(int d : spec_.dots)
      if(d < 0 || d >= n_dots) throw std::runtime_error("sampled dot index out of range");
    out_.open(spec_.path, spec_.binary ? std::ios::binary : std::ios::out);
    if(!out_) throw std::runtime_error("Could not open " + spec_.path);
    records_ = num_timesteps / spec_.every;
    if(spec_.binary) {
      const uint32_t head[6] = {1u, static_cast<uint32_t>(n_dots), static_cast<uint32_t>(spec_.dots.size()),
                                static_cast<uint32_t>(spec_.every), static_cast<uint32_t>(records_), 0u};
      out_.write("QUESTOUT", 8);
      out_.write(reinterpret_cast<const char *>(head), sizeof(head));
      out_.write(reinterpret_cast<const char *>(&dt), sizeof(dt));
      out_.write(reinterpret_cast<const char *>(spec_.dots.data()),
                 static_cast<std::streamsize>(sizeof(int) * spec_.dots.size()));
    } else {
      out_ << std::setprecision(12) << std::scientific;
    }
  }
  int every() const { return spec_.every; }
  int records() const { return records_; }
  // The state after step `step` of every dot: rho [n_dots][4] (re, im rho00;
  // re, im rho01), as quest_get_state returns it.
  void write(int step, const double *rho)
  {
    if(spec_.binary) {
      const int32_t tag[2] = {step, 0};
      out_.write(reinterpret_cast<const char *>(tag), sizeof(tag));
      for(int d : spec_.dots) out_.write(reinterpret_cast<const char *>(rho + 4 * static_cast<size_t>(d)), 32);
    } else {
      for(int d : spec_.dots) {
        const double *r = rho + 4 * static_cast<size_t>(d);
        out_ << cmplx(r[0], r[1]) << " " << cmplx(r[2], r[3]) << " ";
      }
      out_ << std::endl;
    }
    if(!out_) throw std::runtime_error("write to " + spec_.path + " failed");
  }
  // From a history array (the host loop): the recorded steps of `history`.
  void write_history(const HistoryArray &history, int num_timesteps)
  {
    std::vector<double> rho(4 * static_cast<size_t>(history.num_dots()));
    for(int t = spec_.every - 1; t < num_timesteps; t += spec_.every) {
      for(int i = 0; i < history.num_dots(); ++i) {
        const soltype &s = history[i][t][0];
        rho[4 * i] = s[0].real();
        rho[4 * i + 1] = s[0].imag();
        rho[4 * i + 2] = s[1].real();
        rho[4 * i + 3] = s[1].imag();
      }
      write(t, rho.data());
    }
  }

 private:
  SampledOutput spec_;
  std::ofstream out_;
  int records_ = 0;
};
}  // namespace History

// -------------------------------------------------------- green functions
namespace GreenFunction {
// green_function.h:16-56: the lab-frame dyadic (prefactor 2 Re rho01)
class Dyadic {
 public:
  Dyadic(const double mu0, const double c, const double hbar) : c_(c), mu0_(mu0), hbar_(hbar) {}
  virtual ~Dyadic() = default;
  virtual int kind() const { return QUEST_DYADIC_LAB; }
  virtual double omega() const { return 0.0; }
  double c() const { return c_; }
  double mu0() const { return mu0_; }
  double hbar() const { return hbar_; }

 protected:
  double c_, mu0_, hbar_;
};

// rotating_green_function.h:10-25 (prefactor rho01)
class RotatingDyadic : public Dyadic {
 public:
  RotatingDyadic(const double mu0, const double c, const double hbar, const double omega)
      : Dyadic(mu0, c, hbar), omega_(omega) {}
  int kind() const override { return QUEST_DYADIC_ROTATING; }
  double omega() const override { return omega_; }

 private:
  double omega_;
};
}  // namespace GreenFunction

// ------------------------------------------------------------ interactions
// interaction.h:9-21
class Interaction {
 public:
  using ResultArray = std::vector<cmplx>;
  explicit Interaction(const std::shared_ptr<const DotVector> &dots)
      : dots(dots), results(dots->size()) {}
  virtual ~Interaction() = default;
  const cmplx &operator[](const int i) const { return results[i]; }
  virtual const ResultArray &evaluate(const int) = 0;

 protected:
  std::shared_ptr<const DotVector> dots;
  ResultArray results;
};

namespace quest_detail {
// One device context for a dot system.
class Context {
 public:
  Context(const DotVector &dots, const quest_params &p)
  {
    const size_t n = dots.size();
    std::vector<double> pos(3 * n), dip(3 * n), freq(n), t1(n), t2(n);
    for(size_t i = 0; i < n; ++i) {
      for(int a = 0; a < 3; ++a) {
        pos[3 * i + a] = dots[i].position()[a];
        dip[3 * i + a] = dots[i].dipole()[a];
      }
      freq[i] = dots[i].frequency();
      t1[i] = dots[i].decay().first;
      t2[i] = dots[i].decay().second;
    }
    quest_ctx *c = nullptr;
    check(quest_ctx_create(&p, static_cast<int>(n), pos.data(), dip.data(), freq.data(),
                           t1.data(), t2.data(), &c),
          nullptr);
    ctx_.reset(c);
  }
  quest_ctx *get() const { return ctx_.get(); }

 private:
  struct Del {
    void operator()(quest_ctx *c) const { quest_ctx_destroy(c); }
  };
  std::unique_ptr<quest_ctx, Del> ctx_;
};

inline quest_params params_from_config(int dyadic_kind, double omega, int interp_order)
{
  quest_params p{};
  p.c0 = config.c0;
  p.hbar = config.hbar;
  p.mu0 = config.mu0;
  p.laser_omega = omega;
  p.dt = config.dt;
  p.interp_order = interp_order;
  p.history_window = 22;
  p.precision = config.precision;
  p.dyadic = dyadic_kind;
  p.device = config.device;
  p.accumulation = config.accumulation;
  return p;
}
}  // namespace quest_detail

// pulse_interaction.h:8-16 -- evaluated on the device
class PulseInteraction : public Interaction {
 public:
  PulseInteraction(const std::shared_ptr<const DotVector> &dots, const std::shared_ptr<const Pulse> pulse)
      : Interaction(dots), pulse(pulse),
        ctx(*dots, quest_detail::params_from_config(QUEST_DYADIC_ROTATING, config.omega, 1))
  {
    const quest_pulse q = pulse->to_c();
    quest_detail::check(quest_set_pulse(ctx.get(), &q), ctx.get());
    buf.resize(2 * dots->size());
  }

  const ResultArray &evaluate(const int time_idx) override
  {
    quest_detail::check(quest_pulse_evaluate(ctx.get(), time_idx, buf.data()), ctx.get());
    for(size_t i = 0; i < results.size(); ++i) results[i] = cmplx(buf[2 * i], buf[2 * i + 1]);
    return results;
  }

 private:
  std::shared_ptr<const Pulse> pulse;
  quest_detail::Context ctx;
  std::vector<double> buf;
};

// history_interaction.h:14-35 -- the HIP drop-in.  Reads the polarisation
// from the shared history exactly where the reference does ((*history)[i][step][0][1]);
// evaluates the retarded pair sum on the GPU.
class HistoryInteraction : public Interaction {
 public:
  HistoryInteraction(const std::shared_ptr<const DotVector> &dots,
                     const std::shared_ptr<const History::HistoryArray> &history,
                     const std::shared_ptr<GreenFunction::Dyadic> &dyadic, const int interp_order)
      : Interaction(dots), history(history), dyadic(dyadic),
        ctx(*dots, quest_detail::params_from_config(dyadic->kind(), dyadic->omega(), interp_order)),
        rho(2 * dots->size()), out(2 * dots->size())
  {
    if(config.pair_method != QUEST_PAIR_ON_THE_FLY)
      quest_detail::check(quest_set_pair_method(ctx.get(), config.pair_method), ctx.get());
    else if(config.autotune)
      quest_detail::check(quest_autotune(ctx.get(), 3, nullptr), ctx.get());
  }

  const ResultArray &evaluate(const int time_idx) override
  {
    const int n = static_cast<int>(results.size());
    for(int i = 0; i < n; ++i) {
      const cmplx r = (*history)[i][time_idx][0][1];
      rho[2 * i] = r.real();
      rho[2 * i + 1] = r.imag();
    }
    quest_detail::check(quest_hist_evaluate(ctx.get(), time_idx, rho.data(), out.data()), ctx.get());
    for(int i = 0; i < n; ++i) results[i] = cmplx(out[2 * i], out[2 * i + 1]);
    return results;
  }

 private:
  std::shared_ptr<const History::HistoryArray> history;
  std::shared_ptr<GreenFunction::Dyadic> dyadic;
  quest_detail::Context ctx;
  std::vector<double> rho, out;
};

// ----------------------------------------------------------- integrator.h
namespace PredictorCorrector {
constexpr int NUM_CORRECTOR_STEPS = 10;  // integrator.cpp:3

// integrator.cpp:85-102 (solved by the library's host Jacobi SVD)
class Weights {
 public:
  Weights(const int n_lambda, const int n_time, const double radius) : n_time(n_time)
  {
    ps.assign(2, std::vector<double>(n_time));
    cs.assign(2, std::vector<double>(n_time));
    std::vector<double> p(2 * n_time), c(2 * n_time);
    quest_detail::check(quest_pc_weights(n_lambda, n_time, radius, p.data(), c.data(), &future_coef), nullptr);
    for(int j = 0; j < 2; ++j)
      for(int h = 0; h < n_time; ++h) {
        ps[j][h] = p[j * n_time + h];
        cs[j][h] = c[j * n_time + h];
      }
  }
  std::vector<std::vector<double>> ps, cs;
  double future_coef = 0;
  int width() const { return n_time; }

 private:
  int n_time;
};

// integrator.cpp:104-216: the host loop, unchanged in structure.
class Integrator {
 public:
  Integrator(const double dt, const int n_lambda, const int n_time, const double radius,
             const std::shared_ptr<History::HistoryArray> &history,
             const std::vector<rhs_func> &rhs_funcs,
             std::vector<std::shared_ptr<Interaction>> interactions)
      : num_solutions(static_cast<int>(rhs_funcs.size())),
        time_idx_ubound(history->time_base() + history->time_extent()), dt(dt),
        weights(n_lambda, n_time, radius), history(history), rhs_funcs(rhs_funcs),
        interactions(std::move(interactions)) {}

  void solve(bool log = true) const
  {
    for(int step = 0; step < time_idx_ubound; ++step) {
      solve_step(step);
      if(log) log_percentage_complete(step);
      throw_if_unbounded_solution(step);
    }
  }

 private:
  int num_solutions, time_idx_ubound;
  double dt;
  Weights weights;
  std::shared_ptr<History::HistoryArray> history;
  std::vector<rhs_func> rhs_funcs;
  std::vector<std::shared_ptr<Interaction>> interactions;

  void solve_step(const int step) const
  {
    predictor(step);
    evaluator(step);
    for(int m = 0; m < NUM_CORRECTOR_STEPS; ++m) {
      corrector(step);
      evaluator(step);
    }
  }
  void predictor(const int step) const
  {
    const int start = step - weights.width();
    for(int s = 0; s < num_solutions; ++s) {
      History::soltype y{0.0, 0.0};
      for(int h = 0; h < weights.width(); ++h)
        for(int k = 0; k < 2; ++k)
          y[k] += (*history)[s][start + h][0][k] * weights.ps[0][h] +
                  (*history)[s][start + h][1][k] * weights.ps[1][h] * dt;
      (*history)[s][step][0] = y;
    }
  }
  void corrector(const int step) const
  {
    const int start = step - weights.width();
    for(int s = 0; s < num_solutions; ++s) {
      History::soltype y;
      for(int k = 0; k < 2; ++k) y[k] = weights.future_coef * (*history)[s][step][1][k] * dt;
      for(int h = 0; h < weights.width(); ++h)
        for(int k = 0; k < 2; ++k)
          y[k] += (*history)[s][start + h][0][k] * weights.cs[0][h] +
                  (*history)[s][start + h][1][k] * weights.cs[1][h] * dt;
      (*history)[s][step][0] = y;
    }
  }
  void evaluator(const int step) const
  {
    Interaction::ResultArray fields(num_solutions, 0.0);
    for(const auto &interaction : interactions) {
      const Interaction::ResultArray &r = interaction->evaluate(step);
      for(int s = 0; s < num_solutions; ++s) fields[s] += r[s];
    }
    for(int s = 0; s < num_solutions; ++s)
      (*history)[s][step][1] = rhs_funcs[s]((*history)[s][step][0], fields[s]);
  }
  bool all_finite(const int step) const
  {
    for(int s = 0; s < num_solutions; ++s)
      if(!History::isfinite((*history)[s][step][0]) || !History::isfinite((*history)[s][step][1]))
        return false;
    return true;
  }
  void throw_if_unbounded_solution(const int step) const
  {
    if(!all_finite(step))
      throw std::domain_error("unbounded history values at or before step " + std::to_string(step));
  }
  void log_percentage_complete(const int step) const
  {
    if(time_idx_ubound >= 10 && step % (time_idx_ubound / 10) == 0)
      std::cout << "\t" << static_cast<double>(step) / time_idx_ubound << std::endl;
  }
};

// The whole predictor-corrector loop on the device (quest_run).  Same wiring
// as main.cpp (pulse first, then the history interaction; ground state).
// record_every > 0 copies every record_every-th step into the history.
class DeviceIntegrator {
 public:
  DeviceIntegrator(const std::shared_ptr<const DotVector> &dots, const std::shared_ptr<const Pulse> &pulse,
                   const std::shared_ptr<GreenFunction::Dyadic> &dyadic, const int interp_order,
                   const std::shared_ptr<History::HistoryArray> &history)
      : dots(dots), history(history),
        ctx(*dots, quest_detail::params_from_config(dyadic->kind(), dyadic->omega(), interp_order))
  {
    if(pulse) {
      const quest_pulse q = pulse->to_c();
      quest_detail::check(quest_set_pulse(ctx.get(), &q), ctx.get());
    }
    if(config.pair_method != QUEST_PAIR_ON_THE_FLY)
      quest_detail::check(quest_set_pair_method(ctx.get(), config.pair_method), ctx.get());
    else if(config.autotune)
      quest_detail::check(quest_autotune(ctx.get(), 3, nullptr), ctx.get());
  }
  // Several pulse interactions (their fields add, in list order, ahead of the
  // history interaction -- the reference integrator's sum over interactions,
  // integrator.cpp:173-179).
  DeviceIntegrator(const std::shared_ptr<const DotVector> &dots,
                   const std::vector<std::shared_ptr<const Pulse>> &pulses,
                   const std::shared_ptr<GreenFunction::Dyadic> &dyadic, const int interp_order,
                   const std::shared_ptr<History::HistoryArray> &history)
      : DeviceIntegrator(dots, std::shared_ptr<const Pulse>(), dyadic, interp_order, history)
  {
    std::vector<quest_pulse> q;
    for(const auto &p : pulses) q.push_back(p->to_c());
    quest_detail::check(quest_set_pulse_list(ctx.get(), 0, static_cast<int>(q.size()), q.data()),
                        ctx.get());
  }
  void set_weights(const Weights &w)
  {
    std::vector<double> p, c;
    for(int j = 0; j < 2; ++j) {
      p.insert(p.end(), w.ps[j].begin(), w.ps[j].end());
      c.insert(c.end(), w.cs[j].begin(), w.cs[j].end());
    }
    quest_detail::check(quest_set_pc_weights(ctx.get(), w.width(), p.data(), c.data(), w.future_coef), ctx.get());
  }
  // Runs every step of the history (or n_steps); throws std::domain_error
  // with the reference's message on a non-finite state.
  void solve(int record_every = 1, int n_steps = -1)
  {
    const int total = n_steps >= 0 ? n_steps : history->time_base() + history->time_extent();
    const size_t n = dots->size();
    std::vector<double> rho(4 * n), drho(4 * n);
    int done = 0;
    while(done < total) {
      const int chunk = record_every > 0 ? std::min(record_every, total - done) : total - done;
      quest_detail::check(quest_run(ctx.get(), chunk), ctx.get());
      done += chunk;
      if(record_every > 0) {
        quest_detail::check(quest_get_state(ctx.get(), rho.data(), drho.data()), ctx.get());
        for(size_t i = 0; i < n; ++i) {
          (*history)[static_cast<int>(i)][done - 1][0] = {cmplx(rho[4 * i], rho[4 * i + 1]),
                                                          cmplx(rho[4 * i + 2], rho[4 * i + 3])};
          (*history)[static_cast<int>(i)][done - 1][1] = {cmplx(drho[4 * i], drho[4 * i + 1]),
                                                          cmplx(drho[4 * i + 2], drho[4 * i + 3])};
        }
      }
    }
  }
  // Runs n_steps steps and streams every writer.every()-th step's state to
  // `out` (no host history is filled).
  void solve_sampled(History::SampledWriter &out, int n_steps)
  {
    const size_t n = dots->size();
    std::vector<double> rho(4 * n);
    int done = 0;
    while(done < n_steps) {
      const int chunk = std::min(out.every(), n_steps - done);
      quest_detail::check(quest_run(ctx.get(), chunk), ctx.get());
      done += chunk;
      if(chunk == out.every()) {
        quest_detail::check(quest_get_state(ctx.get(), rho.data(), nullptr), ctx.get());
        out.write(done - 1, rho.data());
      }
    }
  }
  quest_ctx *context() const { return ctx.get(); }

 private:
  std::shared_ptr<const DotVector> dots;
  std::shared_ptr<History::HistoryArray> history;
  quest_detail::Context ctx;
};
}  // namespace PredictorCorrector
