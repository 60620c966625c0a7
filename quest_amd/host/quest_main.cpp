// quest_main.cpp -- the `quest` driver (src/main.cpp:17-63) over the GPU
// library.  Reads the reference's three input files and writes output.dat
// in the reference's format.
//
//   quest_hip [--config input.cfg] [--precision fp64|mixed]
//             [--accumulation block|compensated] [--tune] [--host-loop]
//             [--all-pulses] [--pair-method time-blocked|on-the-fly|table]
//             [--output output.dat] [--format text|binary] [--every K]
//             [--subsample COUNT:SEED] [--device N] [--help] [--version]
//
// input.cfg is the reference's INI file (configuration.cpp:17-42):
//   [files]      qd_path (dots.cfg), pulse_path (pulse.cfg)
//   [constants]  c0, hbar, mu0, laser_frequency       (all required)
//   [parameters] num_particles, total_time, timestep,
//                interpolation_order                  (all required)
// num_timesteps = ceil(total_time / timestep) (configuration.cpp:37-40).
// By default the whole predictor-corrector loop runs on the device;
// --host-loop keeps the reference's host Integrator and evaluates the
// interactions on the device at every predictor/corrector evaluation.
// --all-pulses drives the dots with every row of pulse.cfg (one pulse
// interaction each, fields added in file order) instead of the first row
// only, which is what the reference reads (pulse.cpp:37-46).
// --pair-method: time-blocked (default) forms the pair sums of the pairs of
// delay >= 8 for 8 steps in one pass (QUEST_PAIR_TIME_BLOCKED); on-the-fly
// evaluates every pair at every step; table streams the reference's per-pair
// coefficient table (FP64 only).
// --accumulation compensated makes mixed precision's 64-source FP32 block
// sums Kahan-compensated; --tune times the device loop's work splits
// (quest_autotune) before the first step.
// --format binary, --every K and --subsample COUNT:SEED write only every K-th
// step and/or a seeded subsample of the dots, as text rows in output.dat's
// format or as binary records (History::SampledWriter); with the device loop
// the states stream from the device every K steps and no O(N T) host history
// is held (at N = 65536 x 2000 steps the full text output is about 5 GB).
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>

#include "quest.hpp"

namespace {

// --help, --version or a missing configuration file: nothing to run.
struct CommandLineException : public std::exception {
  const char *what() const noexcept override { return "command line handled"; }
};

std::string trim(const std::string &s)
{
  const size_t a = s.find_first_not_of(" \t\r\n");
  if(a == std::string::npos) return "";
  const size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// INI reader with boost::program_options' conventions: "[section]" headers,
// "key = value" lines, '#' starts a comment.
std::map<std::string, std::string> read_ini(std::istream &is)
{
  std::map<std::string, std::string> kv;
  std::string line, section;
  while(std::getline(is, line)) {
    const size_t hash = line.find('#');
    if(hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if(line.empty()) continue;
    if(line.front() == '[' && line.back() == ']') {
      section = trim(line.substr(1, line.size() - 2));
      continue;
    }
    const size_t eq = line.find('=');
    if(eq == std::string::npos) throw std::runtime_error("malformed configuration line: " + line);
    kv[section + "." + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return kv;
}

double req_double(const std::map<std::string, std::string> &kv, const std::string &key)
{
  auto it = kv.find(key);
  if(it == kv.end()) throw std::runtime_error("the option '--" + key + "' is required but missing");
  return std::stod(it->second);
}

int req_int(const std::map<std::string, std::string> &kv, const std::string &key)
{
  auto it = kv.find(key);
  if(it == kv.end()) throw std::runtime_error("the option '--" + key + "' is required but missing");
  return std::stoi(it->second);
}

struct Options {
  std::string config_path = "input.cfg";
  std::string output = "output.dat";
  bool host_loop = false;
  bool all_pulses = false;
  bool binary = false;
  int every = 1;
  int sub_count = 0;  // 0: every dot
  uint64_t sub_seed = 0;
  bool sampled() const { return binary || every != 1 || sub_count > 0; }
};

// configuration.cpp:7-81
Options parse_configs(int argc, char *argv[])
{
  Options o;
  for(int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&]() -> std::string {
      if(i + 1 >= argc) throw std::runtime_error("missing value for " + a);
      return argv[++i];
    };
    if(a == "--help") {
      std::cout << "QuEST options (MI355X build):\n"
                   "  --help                 print this help message\n"
                   "  -v [ --version ]       print version string\n"
                   "  -c [ --config ] arg    path to configuration file (default input.cfg)\n"
                   "  --precision arg        fp64 (default) or mixed\n"
                   "  --accumulation arg     block (default) or compensated (mixed precision's block sums)\n"
                   "  --tune                 time the pair kernel's work splits before the run (device loop)\n"
                   "  --host-loop            host predictor-corrector, device interactions\n"
                   "  --all-pulses           every row of pulse.cfg (default: the first)\n"
                   "  --pair-method arg      time-blocked (default), on-the-fly or table (FP64 only)\n"
                   "  --output arg           output file (default output.dat)\n"
                   "  --format arg           text (default, output.dat's format) or binary\n"
                   "  --every arg            record every arg-th step (default 1)\n"
                   "  --subsample arg        COUNT:SEED, record a seeded subsample of the dots\n"
                   "  --device arg           HIP device ordinal (default 0)\n";
      throw CommandLineException();
    } else if(a == "--version" || a == "-v") {
      std::cout << "Quantum Electromagnetics Simulation Toolbox (QuEST), MI355X build" << std::endl;
      throw CommandLineException();
    } else if(a == "--config" || a == "-c") {
      o.config_path = next();
    } else if(a == "--precision") {
      const std::string v = next();
      if(v == "fp64") config.precision = QUEST_PREC_FP64;
      else if(v == "mixed") config.precision = QUEST_PREC_MIXED;
      else throw std::runtime_error("--precision must be fp64 or mixed");
    } else if(a == "--accumulation") {
      const std::string v = next();
      if(v == "block") config.accumulation = QUEST_ACC_BLOCK;
      else if(v == "compensated") config.accumulation = QUEST_ACC_COMPENSATED;
      else throw std::runtime_error("--accumulation must be block or compensated");
    } else if(a == "--tune") {
      config.autotune = true;
    } else if(a == "--host-loop") {
      o.host_loop = true;
    } else if(a == "--all-pulses") {
      o.all_pulses = true;
    } else if(a == "--pair-method") {
      const std::string v = next();
      if(v == "time-blocked") config.pair_method = QUEST_PAIR_TIME_BLOCKED;
      else if(v == "on-the-fly") config.pair_method = QUEST_PAIR_ON_THE_FLY;
      else if(v == "table") config.pair_method = QUEST_PAIR_TABLE;
      else throw std::runtime_error("--pair-method must be time-blocked, on-the-fly or table");
    } else if(a == "--output") {
      o.output = next();
    } else if(a == "--format") {
      const std::string v = next();
      if(v == "text") o.binary = false;
      else if(v == "binary") o.binary = true;
      else throw std::runtime_error("--format must be text or binary");
    } else if(a == "--every") {
      o.every = std::stoi(next());
      if(o.every < 1) throw std::runtime_error("--every must be >= 1");
    } else if(a == "--subsample") {
      const std::string v = next();
      const size_t colon = v.find(':');
      if(colon == std::string::npos) throw std::runtime_error("--subsample takes COUNT:SEED");
      o.sub_count = std::stoi(v.substr(0, colon));
      o.sub_seed = std::stoull(v.substr(colon + 1));
      if(o.sub_count < 1) throw std::runtime_error("--subsample COUNT must be >= 1");
    } else if(a == "--device") {
      config.device = std::stoi(next());
    } else {
      throw std::runtime_error("unrecognised option '" + a + "'");
    }
  }
  std::ifstream ifs(o.config_path);
  if(!ifs) {
    std::cerr << "ERROR: " << o.config_path << " not found" << std::endl;
    throw CommandLineException();
  }
  const auto kv = read_ini(ifs);
  if(kv.count("files.qd_path")) config.qd_path = kv.at("files.qd_path");
  if(kv.count("files.pulse_path")) config.pulse_path = kv.at("files.pulse_path");
  config.c0 = req_double(kv, "constants.c0");
  config.hbar = req_double(kv, "constants.hbar");
  config.mu0 = req_double(kv, "constants.mu0");
  config.omega = req_double(kv, "constants.laser_frequency");
  config.num_particles = req_int(kv, "parameters.num_particles");
  config.dt = req_double(kv, "parameters.timestep");
  config.total_time = req_double(kv, "parameters.total_time");
  config.num_timesteps = static_cast<int>(std::ceil(config.total_time / config.dt));
  config.interpolation_order = req_int(kv, "parameters.interpolation_order");
  return o;
}

}  // namespace

// main.cpp:17-63
int main(int argc, char *argv[])
{
  try {
    std::cout << std::setprecision(12) << std::scientific;
    const Options opt = parse_configs(argc, argv);

    std::cout << "Initializing..." << std::endl;
    auto qds = std::make_shared<DotVector>(import_dots(config.qd_path));
    qds->resize(config.num_particles);
    auto rhs_funs = rhs_functions(*qds, config.omega);

    // the device loop with sampled output streams its records and needs no host history
    const bool stream = opt.sampled() && !opt.host_loop;
    auto history = History::make_shared_history(config.num_particles, 22, stream ? 1 : config.num_timesteps);
    for(int t = -22; t <= 0; ++t)
      for(int sol_idx = 0; sol_idx < config.num_particles; ++sol_idx)
        (*history)[sol_idx][t][0] = {cmplx(1, 0), cmplx(0, 0)};  // ground state
    std::unique_ptr<History::SampledWriter> sampled;
    if(opt.sampled()) {
      History::SampledOutput spec;
      spec.path = opt.output;
      spec.binary = opt.binary;
      spec.every = opt.every;
      if(opt.sub_count > 0) spec.dots = History::subsample_dots(config.num_particles, opt.sub_count, opt.sub_seed);
      sampled = std::make_unique<History::SampledWriter>(spec, config.num_particles, config.num_timesteps, config.dt);
    }

    std::vector<std::shared_ptr<const Pulse>> pulses;
    if(opt.all_pulses) {
      for(const Pulse &p : read_pulse_configs(config.pulse_path)) pulses.push_back(std::make_shared<Pulse>(p));
      if(pulses.size() > QUEST_MAX_PULSES)
        throw std::runtime_error("pulse file has more than " + std::to_string(QUEST_MAX_PULSES) + " rows");
    } else {
      pulses.push_back(std::make_shared<Pulse>(read_pulse_config(config.pulse_path)));
    }
    auto rotating_dyadic =
        std::make_shared<GreenFunction::RotatingDyadic>(config.mu0, config.c0, config.hbar, config.omega);

    std::cout << "Solving..." << std::endl;
    if(opt.host_loop) {
      std::vector<std::shared_ptr<Interaction>> interactions;
      for(const auto &p : pulses) interactions.push_back(std::make_shared<PulseInteraction>(qds, p));
      interactions.push_back(
          std::make_shared<HistoryInteraction>(qds, history, rotating_dyadic, config.interpolation_order));
      PredictorCorrector::Integrator integrator(config.dt, 18, 22, 3.15, history, rhs_funs,
                                                std::move(interactions));
      integrator.solve();
    } else {
      PredictorCorrector::DeviceIntegrator integrator(qds, pulses, rotating_dyadic,
                                                      config.interpolation_order, history);
      integrator.set_weights(PredictorCorrector::Weights(18, 22, 3.15));
      if(stream) integrator.solve_sampled(*sampled, config.num_timesteps);
      else integrator.solve(1);
    }

    std::cout << "Writing output..." << std::endl;
    if(!opt.sampled()) History::write_history(history, opt.output);
    else if(opt.host_loop) sampled->write_history(*history, config.num_timesteps);
  } catch(CommandLineException &e) {
    // --help / --version / missing configuration file: nothing more to do
  } catch(std::domain_error &e) {
    std::cerr << "domain_error: " << e.what() << std::endl;
    return 2;
  } catch(std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
