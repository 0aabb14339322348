// The settings of a reconstruction under the command line's option names: the method, the
// regularizer, the data term of every component or of each, the primal-dual method's
// options, and those the reference takes but this library only keeps (the subgradient
// method's, the limits of reading). The names, ranges and defaults are the reference's
// (rust/unround/src/cli.rs:64-136, 338-580; cpp/src/settings.cpp:36-410).
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <vector>

namespace unround {

enum class Method : int32_t { mmse, tv, tgv, subgradient };

// The arithmetic of the iterations on the GPU: binary32 iterates, binary64 iterates, or
// binary32 iterates with the dual fields p and r stored in bfloat16 (an option of this
// library: --precision single|double|mixed).
enum class Precision : int32_t { single, dual, mixed };

// The options of G of one component (cpp/include/unround/model.hpp:39-63).
struct DataTerm {
  std::optional<double> mu;  // none: the rule mu_scale * mean(Q)^mu_power
  double mu_scale = 9.0;
  double mu_power = 0.9;
  double mu_chroma = 0.3;
  double slack = 0.0;
  double dc_weight = 1.0;
  bool mmse = false;  // the centres: MMSE (true) or the middles of the intervals
  double power = 2.0;
  double slack_cost = 0.0;
  double flat_weight = 0.1;
  uint32_t flat_levels = 0;
  std::optional<uint32_t> flat_dc = 1u;
};

struct Settings {
  Method method = Method::tgv;
  double alpha = 1.0;
  double alpha1 = 1.0;
  double alpha0 = 2.0;
  std::optional<std::vector<double>> channel_weights;
  bool coupled = true;
  std::vector<DataTerm> data = {DataTerm{}};
  // The primal-dual method (cpp/include/unround/pdhg.hpp:59-75).
  std::optional<uint64_t> iterations;
  std::optional<double> tolerance;
  double relative_tolerance = 0.0;
  double partial_tolerance = 0.0;
  std::optional<double> partial_radius;
  std::optional<double> step_ratio;
  std::optional<double> relaxation;
  double step_product = 0.99;
  std::optional<double> norm_squared;
  bool weight_scaling = true;
  uint64_t record_every = 10;
  double free_radius = 255.0;
  uint64_t threads = 0;
  // Kept for the round trip of the options; not used here.
  uint64_t subgradient_iterations = 50;
  double subgradient_step = 0.5;
  double subgradient_decay = 0.5;
  bool momentum = true;
  uint64_t subgradient_record_every = 1;
  uint64_t max_pixels = 0;
  uint64_t max_scans = 0;
  bool warnings_are_errors = false;
  Precision precision = Precision::single;
};

// The settings of the options, or every refusal joined by "; ".
bool parse_settings(const std::vector<std::string>& arguments, Settings& settings, std::string& errors);

// The options that give the settings: "--name=value" or "--flag", a space between, each
// number the shortest decimal that reads back as the same double.
std::string options_text(const Settings& settings);

// The shortest decimal that reads back as the same double, as the reference writes it
// (Rust's Debug of f64: "1.0", "0.001", "1e-7", "1e16").
std::string decimal(double value);

}  // namespace unround
