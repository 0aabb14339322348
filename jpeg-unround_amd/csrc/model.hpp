// The model of a picture's components on one canvas, as the host sets it up before the
// GPU solves: each component's levels, intervals' constants, weights, flat blocks and
// centres (docs/math.md 1.2-1.3, 2, 4.1, 4.4), and the plan of the primal-dual method
// (docs/math.md 5, 6.5). The per-coefficient arrays (ends, centres) are not stored: the
// kernels compute them from the int16 levels and the per-frequency constants, and the
// result makes them on demand (model.cpp says from what).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "settings.hpp"

namespace unround {

// One component on the canvas (cpp/include/unround/frames.hpp:47-56;
// cpp/include/unround/model.hpp:78-164).
struct ComponentModel {
  uint32_t rows = 0;     // block rows kept
  uint32_t columns = 0;  // blocks across kept
  uint32_t down = 1;     // canvas rows per sample of the component
  uint32_t across = 1;   // canvas columns per sample
  uint32_t h = 1, v = 1;
  std::vector<int16_t> levels;  // rows x columns x 64, natural order
  std::vector<uint8_t> flat;    // rows x columns
  std::array<uint16_t, 64> table{};
  std::array<double, 64> steps{};
  std::array<double, 64> weights{};
  std::array<double, 64> flat_weights{};
  std::array<double, 64> shrink{};  // delta / Q of the MMSE centres, 0 for the middles
  std::array<double, 64> costs{};   // slack_cost / Q where the slack has a price
  double slack = 0.0;
  bool mmse = false;
  bool costed = false;

  size_t blocks() const { return static_cast<size_t>(rows) * columns; }
  uint32_t cells() const { return down * across; }
  // The ends, the centre and the file's own ends of coefficient k of block b, in
  // binary64, as the reference computes them (cpp/src/model.cpp:44-48, 242-262).
  void coefficient(size_t b, int k, double& lower, double& upper, double& centre, double& own_lower,
                   double& own_upper) const;
};

struct FrameModel {
  uint32_t height = 0, width = 0;  // the picture
  uint32_t rows = 0, columns = 0;  // the canvas
  int32_t color_space = 1;
  std::vector<ComponentModel> components;

  size_t channels() const { return components.size(); }
  // Whether component c has free samples: cells of several samples, or canvas beyond its
  // blocks (cpp/src/frames.cpp:421-425).
  bool free(size_t c) const;
};

struct ComponentInput {
  const int16_t* levels = nullptr;
  uint32_t rows = 0, columns = 0;
  int32_t h = 1, v = 1;
  std::array<uint16_t, 64> table{};
};

// The frame of the inputs (cpp/src/frames.cpp:343-419; rust/unround/src/decode.rs:233-261):
// false with the reason, and a status, where refused (3 unsupported, 4 options).
bool make_frame(uint32_t height, uint32_t width, int32_t color_space, const std::vector<ComponentInput>& inputs,
                const Settings& settings, FrameModel& frame, std::string& message, int32_t& status);

// The plan of the primal-dual method (cpp/src/pdhg.cpp:147-233).
struct Plan {
  double tau = 0.0, sigma = 0.0;
  uint64_t iterations = 0;
  double tolerance = 0.0;
  double relative_tolerance = 0.0;
  double partial_tolerance = 0.0;
  bool has_partial = false;
  double partial_radius = 0.0;
  double relaxation = 1.9;
  uint64_t record_every = 10;
  double free_radius = 255.0;
  std::vector<double> gamma;
};

bool make_plan(const Settings& settings, size_t channels, Plan& plan, std::string& message);

// A picture of rounded samples (docs/math.md 11; rust/unround/src/samples.rs:56-140):
// channels x height x width levels of `bits` bits (1, 2, 4, 8 or 16), channel by channel,
// each standing for the values that round to it on the scale of 0 to 255, with the data
// term's weight mu (0: the intervals alone) and a slack in levels. The ends and centres of
// every level 0..top, in binary64: without a slack the doubles x with
// max(0, (v - 1/2) s) <= x < (v + 1/2) s, or x <= 255 at the top level, s = 255 / top,
// each end exact; with a slack, ((v -+ 1/2) -+ slack) s in floating point, within 0 and
// 255; the centre the double nearest v s.
struct SamplesModel {
  uint32_t channels = 0, height = 0, width = 0, bits = 8;
  double mu = 0.0, slack = 0.0;
  std::vector<uint16_t> levels;  // channels x height x width
  std::vector<double> low, high, centre;  // per level, 0..top

  uint32_t top() const { return (1u << bits) - 1u; }
  size_t count() const { return static_cast<size_t>(channels) * height * width; }
};

// The data term of a picture of rounded samples from the settings: --mu (0 where it follows
// JPEG's rule) and --slack, one value each; the other options of the data term are JPEG's
// and refused where they are not their defaults (rust/unround/src/rounded.rs:44-70).
bool samples_data_term(const Settings& settings, double& mu, double& slack, std::string& message);

// The model of the levels under the settings, whose data term holds --mu and --slack alone
// (rust/unround/src/rounded.rs:44-70; samples.rs:74-110): false with the reason where
// refused (status 4, options).
bool make_samples(const uint16_t* levels, uint32_t channels, uint32_t height, uint32_t width, uint32_t bits,
                  const Settings& settings, SamplesModel& model, std::string& message, int32_t& status);

// The plan of TV or TGV on a picture of rounded samples: the defaults are a tolerance of
// 1e-4, 10000 iterations and a step ratio of 10 (TV) or 0.3 (TGV), the tolerance times the
// first-order weight and the ratio over its square unless weight scaling is off
// (rust/unround/src/samples.rs:25-52).
bool make_samples_plan(const Settings& settings, size_t channels, Plan& plan, std::string& message);

// The Laplace model (docs/math.md 2.1-2.2; cpp/src/laplace.cpp:86-164).
double laplace_scale(uint64_t zeros, uint64_t others, uint64_t sum, double step);
double laplace_shrinkage(double rho);

// The orthonormal DCT basis C[k][n] from the nine cosines (docs/math.md:98-106).
const std::array<double, 64>& dct_basis();

}  // namespace unround
