// mirror_test.cpp -- host-side tests of the C++ mirror (quest_amd/host/quest.hpp),
// restating the reference's unit tests (test/quantum_dot_test.cpp,
// test/history_test.cpp, test/lagrange_set_test.cpp shapes) plus file-format
// and weight checks.  Needs no GPU.  Exit status 0 = all passed.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>

#include "../../quest_amd/host/quest.hpp"

static int failures = 0;
#define CHECK(cond)                                                      \
  do {                                                                   \
    if(!(cond)) {                                                        \
      std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                        \
    }                                                                    \
  } while(0)

// BOOST_CHECK_CLOSE: relative tolerance in percent
static bool close_pct(double a, double b, double pct)
{
  if(a == b) return true;
  return std::fabs(a - b) <= pct / 100 * std::max(std::fabs(a), std::fabs(b));
}

// test/quantum_dot_test.cpp:13-43
static void resonant_rhs()
{
  const Vec3 pos{0, 0, 0}, dipole{0, 0, 1};
  const double omega = 2278.9013;
  const std::pair<double, double> ts(10, 20);
  const QuantumDot qd(pos, omega, ts, dipole);
  const cmplx rabi = 1.0;  // E = z, d = z
  const cmplx H[2][2] = {{0, rabi}, {std::conj(rabi), 0}};
  const cmplx rho[2][2] = {{1, cmplx(0.5, 0.5)}, {cmplx(0.5, -0.5), 0}};
  const cmplx D[2][2] = {{(rho[0][0] - 1.0) / ts.first, rho[0][1] / ts.second},
                         {rho[1][0] / ts.second, rho[1][1] / ts.first}};
  cmplx m[2][2];
  for(int i = 0; i < 2; ++i)
    for(int j = 0; j < 2; ++j) {
      cmplx c = 0;
      for(int k = 0; k < 2; ++k) c += H[i][k] * rho[k][j] - rho[i][k] * H[k][j];
      m[i][j] = -iu * c - D[i][j];
    }
  const matrix_elements rhs = qd.liouville_rhs({rho[0][0], rho[0][1]}, rabi, omega);
  CHECK(close_pct(m[0][0].real(), rhs[0].real(), 1e-9));
  CHECK(close_pct(m[0][0].imag(), rhs[0].imag(), 1e-9));
  CHECK(close_pct(m[0][1].real(), rhs[1].real(), 1e-9));
  CHECK(close_pct(m[0][1].imag(), rhs[1].imag(), 1e-9));
}

// test/history_test.cpp:13-53 (with a true imaginary infinity)
static void nonfinite_detection()
{
  const double inf = std::numeric_limits<double>::infinity(), nan = std::numeric_limits<double>::quiet_NaN();
  const cmplx zero(0, 0), real_inf(inf, 0), imag_inf(0, inf), real_nan(nan, 0), imag_nan(0, nan);
  CHECK(History::isfinite({zero, zero}));
  for(const cmplx bad : {real_inf, imag_inf, real_nan, imag_nan}) {
    CHECK(!History::isfinite({bad, zero}));
    CHECK(!History::isfinite({zero, bad}));
  }
}

// history.h / main.cpp:30-36: index base -window, ground-state fill
static void history_layout()
{
  auto h = History::make_shared_history(3, 22, 10);
  CHECK(h->time_base() == -22);
  CHECK(h->time_extent() == 32);
  (*h)[2][-22][1] = {cmplx(1, 2), cmplx(3, 4)};
  (*h)[2][9][0] = {cmplx(5, 6), cmplx(7, 8)};
  CHECK((*h)[2][-22][1][1] == cmplx(3, 4));
  CHECK((*h)[2][9][0][0] == cmplx(5, 6));
  CHECK((*h)[1][9][0][0] == cmplx(0, 0));
}

// quantum_dot.cpp:33-47, pulse.cpp:29-46: the reference's input formats
static void file_formats()
{
  {
    std::ofstream f("mirror_test_dots.cfg");
    f << "0.1 -0.2 0.3 2278.9 10 20 0 0 5.2917721e-4\n1 2 3 2000 5 6 1e-4 2e-4 3e-4\n";
  }
  const DotVector d = import_dots("mirror_test_dots.cfg");
  CHECK(d.size() == 2);
  CHECK(d[1].position().z == 3 && d[1].frequency() == 2000 && d[1].decay().second == 6);
  CHECK(d[0].dipole().z == 5.2917721e-4);
  bool threw = false;
  try {
    import_dots("no_such_file.cfg");
  } catch(std::runtime_error &e) {
    threw = std::string(e.what()) == "Could not open no_such_file.cfg";
  }
  CHECK(threw);
  {
    std::ofstream f("mirror_test_pulse.cfg");
    f << "15589.2260227 5 227.89013 2278.9013 0 0 7.6015964 1 1 0\n";
  }
  const Pulse p = read_pulse_config("mirror_test_pulse.cfg");
  const quest_pulse q = p.to_c();
  CHECK(q.polarization[0] == 1 && q.polarization[1] == 1);   // read path: not normalised
  const Pulse pn(1, 0, 1, 1, {0, 0, 1}, {3, 4, 0});            // constructor normalises
  CHECK(std::fabs(pn.to_c().polarization[0] - 0.6) < 1e-15);
  // pulse value: (E0/2) pol gaussian((k.r - w (t - t0)) / width)
  const Vec3 e = p({0, 0, 0.5}, 5.0);
  const double arg = 7.6015964 * 0.5;
  CHECK(std::fabs(e.x - 15589.2260227 / 2 * std::exp(-std::pow(arg / 227.89013, 2) / 2)) < 1e-9);
  {
    std::ofstream f("mirror_test_pulse.cfg", std::ios::app);
    f << "\n1000 0.02 0.5 2278.9 7.6 0 0 0 0.6 0.8\n";
  }
  // every row, in order (the reference's reader takes the first only)
  const std::vector<Pulse> all = read_pulse_configs("mirror_test_pulse.cfg");
  CHECK(all.size() == 2);
  CHECK(all.size() == 2 && all[0].to_c().amplitude == 15589.2260227 && all[1].to_c().amplitude == 1000);
  CHECK(all.size() == 2 && all[1].to_c().polarization[2] == 0.8);
  CHECK(read_pulse_config("mirror_test_pulse.cfg").to_c().amplitude == 15589.2260227);
  std::remove("mirror_test_dots.cfg");
  std::remove("mirror_test_pulse.cfg");
}

// integrator.cpp:85-102 via quest_pc_weights: shape, sums, step factor
static void weights()
{
  const PredictorCorrector::Weights w(18, 22, 3.15);
  CHECK(w.width() == 22);
  double s0 = 0, c0 = 0;
  for(int h = 0; h < 22; ++h) {
    s0 += w.ps[0][h];
    c0 += w.cs[0][h];
  }
  CHECK(std::fabs(s0 - 1) < 1e-12 && std::fabs(c0 - 1) < 1e-12);
  CHECK(std::fabs(w.future_coef - 0.2815) < 1e-3);
}

// write_history format: "(re,im) (re,im) " per dot, setprecision(12) scientific
static void output_format()
{
  auto h = History::make_shared_history(2, 22, 2);
  (*h)[0][0][0] = {cmplx(1, 0), cmplx(0.5, -0.25)};
  (*h)[1][0][0] = {cmplx(0.75, 0), cmplx(-1e-3, 2e-3)};
  History::write_history(h, "mirror_test_output.dat", 1);
  std::ifstream f("mirror_test_output.dat");
  std::string line;
  std::getline(f, line);
  CHECK(line == "(1.000000000000e+00,0.000000000000e+00) (5.000000000000e-01,-2.500000000000e-01) "
                "(7.500000000000e-01,0.000000000000e+00) (-1.000000000000e-03,2.000000000000e-03) ");
  std::remove("mirror_test_output.dat");
}

// History::subsample_dots: COUNT distinct ascending indices, the same set for
// the same seed, a different one for another seed; SampledWriter: a binary
// and a text file of known values (checked by tests/test_cpp_mirror.py with
// quest_amd.outputs).
static void sampled_output()
{
  const std::vector<int> a = History::subsample_dots(1000, 64, 11), b = History::subsample_dots(1000, 64, 11);
  const std::vector<int> c = History::subsample_dots(1000, 64, 12);
  CHECK(a == b);
  CHECK(a != c);
  CHECK(a.size() == 64);
  for(size_t i = 1; i < a.size(); ++i) CHECK(a[i - 1] < a[i]);
  CHECK(a.front() >= 0 && a.back() < 1000);
  const std::vector<int> all = History::subsample_dots(5, 5, 3);
  CHECK((all == std::vector<int>{0, 1, 2, 3, 4}));
  bool threw = false;
  try {
    History::subsample_dots(5, 6, 1);
  } catch(std::runtime_error &) {
    threw = true;
  }
  CHECK(threw);
  // 6 dots, 7 steps, every 3rd step of dots {1, 4}: records at steps 2 and 5;
  // the state of dot i after step t is (t + i/10, -t) and (i, t/2)
  for(int binary : {1, 0}) {
    History::SampledOutput spec;
    spec.path = binary ? "sampled_test.bin" : "sampled_test.txt";
    spec.binary = binary;
    spec.every = 3;
    spec.dots = {1, 4};
    History::SampledWriter w(spec, 6, 7, 1e-4);
    CHECK(w.records() == 2);
    std::vector<double> rho(4 * 6);
    for(int t = 0; t < 7; ++t) {
      for(int i = 0; i < 6; ++i) {
        rho[4 * i] = t + i / 10.0;
        rho[4 * i + 1] = -t;
        rho[4 * i + 2] = i;
        rho[4 * i + 3] = t / 2.0;
      }
      if(t % 3 == 2) w.write(t, rho.data());
    }
  }
}

int main()
{
  resonant_rhs();
  nonfinite_detection();
  history_layout();
  file_formats();
  weights();
  output_format();
  sampled_output();
  if(failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
  else std::printf("mirror_test: all checks passed\n");
  return failures ? 1 : 0;
}
