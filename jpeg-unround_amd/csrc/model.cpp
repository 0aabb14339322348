// The host's setup of the model: frames, problems, Laplace centres and the plan.
#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace unround {

namespace {

using u128 = unsigned __int128;

double power_of(double step, double power) {
  if (power == 2.0) return step * step;
  if (power == 1.0) return step;
  return std::pow(step, power);
}

bool at_least_zero(double v) { return v >= 0.0 && std::isfinite(v); }

// ((q + half) + widen) Q + shift, each step rounded apart (cpp/src/model.cpp:44-48).
#pragma clang fp contract(off)
double interval_end(double level, double half, double widen, double step, double shift) {
  const double widened = (level + half) + widen;
  const double scaled = widened * step;
  return scaled + shift;
}

size_t ceiling(size_t n, size_t d) { return d == 0 ? 0 : (n + d - 1) / d; }

// B_2k / (2k)!, k = 1..11, each the double nearest to the rational (docs/math.md:286-296).
constexpr double series[11] = {
    0x1.5555555555555p-4,  -0x1.6c16c16c16c17p-10, 0x1.1566abc011567p-15, -0x1.bbd779334ef0bp-21,
    0x1.66a8f2bf70ebep-26, -0x1.22805d644267fp-31, 0x1.d6db2c4e09162p-37, -0x1.7da4e1f79955cp-42,
    0x1.355871d652e9ep-47, -0x1.f57d968caacf1p-53, 0x1.967e1f09c376fp-58,
};

}  // namespace

double laplace_scale(uint64_t zeros, uint64_t others, uint64_t sum, double step) {
  if (sum == 0) return 0.0;
  const u128 n0 = zeros, n1 = others, s = sum;
  // A t^2 + n0 t - S = 0 with A = n0 + S + 2 n1, its root written so that nothing cancels.
  const u128 a = n0 + s + 2u * n1;
  const u128 discriminant = n0 * n0 + 4u * a * s;
  const double root = std::sqrt(static_cast<double>(discriminant));
  const double n0_double = static_cast<double>(n0);
  const double t = static_cast<double>(2u * s) / (n0_double + root);
  double log_inverse = 0.0;
  if (t >= 0.5) {
    const double top = static_cast<double>(8u * s * (n0 + n1));
    const double first = root + static_cast<double>(2u * s - n0);
    const double second = n0_double + root;
    log_inverse = -std::log1p(-(top / (first * second)));
  } else {
    log_inverse = -std::log(t);
  }
  return step / (2.0 * log_inverse);
}

double laplace_shrinkage(double rho) {
  if (!(rho > 0.0)) return 0.0;
  if (std::isinf(rho)) return 0.5;
  if (rho < 1.0) {
    const double square = rho * rho;
    double total = series[10];
    for (int k = 9; k >= 0; --k) total = total * square + series[k];
    return total * rho;
  }
  const double decay = std::exp(-rho);
  return (0.5 - 1.0 / rho) + decay / -std::expm1(-rho);
}

const std::array<double, 64>& dct_basis() {
  static const std::array<double, 64> basis = [] {
    // cos(pi j / 16), j = 0..8, each literal's double the nearest to the cosine.
    const double cosines[9] = {1.0,
                               9.8078528040323044912618223613424e-1,
                               9.2387953251128675612818318939679e-1,
                               8.3146961230254523707878837761791e-1,
                               7.0710678118654752440084436210485e-1,
                               5.5557023301960222474283081394853e-1,
                               3.8268343236508977172845998403040e-1,
                               1.9509032201612826784828486847702e-1,
                               0.0};
    const auto cosine = [&cosines](unsigned m) {
      if (m <= 8u) return cosines[m];
      if (m <= 16u) return -cosines[16u - m];
      if (m <= 24u) return -cosines[m - 16u];
      return cosines[32u - m];
    };
    std::array<double, 64> b{};
    for (int n = 0; n < 8; ++n) b[n] = 3.5355339059327376220042218105242e-1;
    for (int k = 1; k < 8; ++k)
      for (int n = 0; n < 8; ++n) b[8 * k + n] = 0.5 * cosine(static_cast<unsigned>(((2 * n + 1) * k) % 32));
    return b;
  }();
  return basis;
}

void ComponentModel::coefficient(size_t b, int k, double& lower, double& upper, double& centre, double& own_lower,
                                 double& own_upper) const {
  const int16_t q = levels[64 * b + static_cast<size_t>(k)];
  const double level = static_cast<double>(q);
  const double shift = k == 0 ? 1024.0 : 0.0;
  lower = interval_end(level, -0.5, -slack, steps[k], shift);
  upper = interval_end(level, 0.5, slack, steps[k], shift);
  own_lower = interval_end(level, -0.5, 0.0, steps[k], shift);
  own_upper = interval_end(level, 0.5, 0.0, steps[k], shift);
  centre = level * steps[k] + shift;
  if (mmse && k > 0) {
    // sign(q) (|q| - delta/Q) Q, 0 for q = 0 (cpp/include/unround/laplace.hpp:55-60).
    const double size = (std::fabs(level) - shrink[k]) * steps[k];
    centre = q == 0 ? 0.0 : (q < 0 ? -size : size);
  }
}

bool FrameModel::free(size_t c) const {
  const ComponentModel& k = components[c];
  return k.cells() > 1u || 8u * k.rows * k.down < rows || 8u * k.columns * k.across < columns;
}

namespace {

// The problem of one component's kept levels (cpp/src/model.cpp:126-244).
bool make_problem(ComponentModel& k, const DataTerm& data, std::string& message) {
  if (!(at_least_zero(data.mu_scale) && at_least_zero(data.mu_chroma) && std::isfinite(data.mu_power))) {
    message = "the scale of mu and its factor in the chroma are at least 0 and finite, and its power finite";
    return false;
  }
  double mu = 0.0;
  if (data.mu) {
    mu = *data.mu;
  } else {
    uint32_t sum = 0;
    for (const uint16_t step : k.table) sum += step;
    const double mean = static_cast<double>(sum) / 64.0;
    mu = data.mu_power == 1.0 ? data.mu_scale * mean : data.mu_scale * std::pow(mean, data.mu_power);
  }
  if (!(at_least_zero(mu) && at_least_zero(data.slack) && at_least_zero(data.dc_weight))) {
    message = "mu, slack and the weight of DC are at least 0 and finite";
    return false;
  }
  if (!at_least_zero(data.power) || !at_least_zero(data.slack_cost) || !at_least_zero(data.flat_weight) ||
      data.flat_levels > 63u) {
    message = "the options of the data term are out of their ranges";
    return false;
  }
  for (const uint16_t step : k.table) {
    if (step == 0) {
      message = "a quantization step is at least 1";
      return false;
    }
  }
  for (int q = 0; q < 64; ++q) {
    k.steps[q] = static_cast<double>(k.table[q]);
    k.weights[q] = mu / power_of(k.steps[q], data.power);
  }
  k.weights[0] *= data.dc_weight;
  for (int q = 0; q < 64; ++q) k.flat_weights[q] = k.weights[q] * data.flat_weight;
  k.slack = data.slack;
  k.mmse = data.mmse;
  k.costed = data.slack > 0.0 && data.slack_cost > 0.0;
  for (int q = 0; q < 64; ++q) k.costs[q] = k.costed ? data.slack_cost / k.steps[q] : 0.0;
  // Flat blocks: at most flat_levels AC levels not 0, and a DC within flat_dc of the
  // blocks beside it (cpp/src/model.cpp:208-226).
  const size_t rows = k.rows, columns = k.columns;
  k.flat.assign(rows * columns, 0);
  const auto dc = [&k, columns](size_t y, size_t x) { return static_cast<int32_t>(k.levels[64 * (y * columns + x)]); };
  for (size_t by = 0; by < rows; ++by) {
    for (size_t bx = 0; bx < columns; ++bx) {
      const int16_t* q = k.levels.data() + 64 * (by * columns + bx);
      uint32_t nonzero = 0;
      for (int e = 1; e < 64; ++e) nonzero += q[e] != 0;
      bool flat = nonzero <= data.flat_levels;
      if (data.flat_dc) {
        const int32_t own = dc(by, bx);
        const auto within = [&](size_t y, size_t x) {
          const int32_t other = dc(y, x);
          return static_cast<uint32_t>(own > other ? own - other : other - own) <= *data.flat_dc;
        };
        flat = flat && (by == 0 || within(by - 1, bx)) && (by + 1 == rows || within(by + 1, bx)) &&
               (bx == 0 || within(by, bx - 1)) && (bx + 1 == columns || within(by, bx + 1));
      }
      k.flat[by * columns + bx] = flat ? 1 : 0;
    }
  }
  k.shrink.fill(0.0);
  if (data.mmse) {
    // The counts of every frequency over the blocks, then each AC frequency's scale and
    // shrinkage (cpp/src/laplace.cpp:123-160).
    std::array<uint64_t, 64> zeros{}, others{}, sums{};
    const size_t blocks = k.blocks();
    for (size_t b = 0; b < blocks; ++b) {
      const int16_t* q = k.levels.data() + 64 * b;
      for (int e = 0; e < 64; ++e) {
        const int64_t level = q[e];
        const uint64_t zero = level == 0 ? 1u : 0u;
        zeros[e] += zero;
        others[e] += 1u - zero;
        sums[e] += zero ? 0u : static_cast<uint64_t>(2 * (level < 0 ? -level : level) - 1);
      }
    }
    for (int e = 1; e < 64; ++e) {
      const double scale = laplace_scale(zeros[e], others[e], sums[e], k.steps[e]);
      const double rho = scale > 0.0 ? k.steps[e] / scale : std::numeric_limits<double>::infinity();
      k.shrink[e] = laplace_shrinkage(rho);
    }
  }
  return true;
}

}  // namespace

bool make_frame(uint32_t height, uint32_t width, int32_t color_space, const std::vector<ComponentInput>& inputs,
                const Settings& settings, FrameModel& frame, std::string& message, int32_t& status) {
  const size_t count = inputs.size();
  const size_t expected = color_space == 1 ? 1u : 3u;
  status = 3;
  if (count != expected) {
    message = "a file of this color space has " + std::to_string(expected) + " components, not " + std::to_string(count);
    return false;
  }
  status = 4;
  if (settings.data.size() != 1 && settings.data.size() != count) {
    message = "options of G for each of the " + std::to_string(count) + " components, or one for all, not " +
              std::to_string(settings.data.size());
    return false;
  }
  status = 3;
  if (height == 0 || width == 0) {
    message = "a picture without samples is refused";
    return false;
  }
  uint32_t hmax = 0, vmax = 0;
  for (const ComponentInput& in : inputs) {
    if (in.h < 1 || in.h > 4 || in.v < 1 || in.v > 4) {
      message = "sampling factors are 1 to 4, not " + std::to_string(in.h) + " by " + std::to_string(in.v);
      return false;
    }
    hmax = std::max(hmax, static_cast<uint32_t>(in.h));
    vmax = std::max(vmax, static_cast<uint32_t>(in.v));
  }
  frame = FrameModel{};
  frame.height = height;
  frame.width = width;
  frame.color_space = color_space;
  if (count == 1) {
    frame.rows = 8u * inputs[0].rows;
    frame.columns = 8u * inputs[0].columns;
    if (frame.rows < height || frame.columns < width) {
      message = "the blocks do not cover the picture";
      return false;
    }
  } else {
    frame.rows = static_cast<uint32_t>(8u * vmax * ceiling(height, 8u * vmax));
    frame.columns = static_cast<uint32_t>(8u * hmax * ceiling(width, 8u * hmax));
  }
  for (size_t c = 0; c < count; ++c) {
    const ComponentInput& in = inputs[c];
    if (hmax % static_cast<uint32_t>(in.h) != 0 || vmax % static_cast<uint32_t>(in.v) != 0) {
      message = "the sampling factors of component " + std::to_string(c) + " are not whole fractions of the largest";
      return false;
    }
    ComponentModel k;
    k.h = static_cast<uint32_t>(in.h);
    k.v = static_cast<uint32_t>(in.v);
    k.down = count == 1 ? 1u : vmax / k.v;
    k.across = count == 1 ? 1u : hmax / k.h;
    k.rows = count == 1 ? in.rows : static_cast<uint32_t>(ceiling(height, 8u * k.down));
    k.columns = count == 1 ? in.columns : static_cast<uint32_t>(ceiling(width, 8u * k.across));
    if (in.rows < k.rows || in.columns < k.columns) {
      status = 4;
      message = "component " + std::to_string(c) + " has fewer blocks than the picture covers";
      return false;
    }
    k.table = in.table;
    k.levels.resize(64 * k.blocks());
    for (size_t by = 0; by < k.rows; ++by) {
      std::copy_n(in.levels + 64 * by * in.columns, 64 * static_cast<size_t>(k.columns),
                  k.levels.begin() + static_cast<ptrdiff_t>(64 * by * k.columns));
    }
    DataTerm term = settings.data.size() == 1 ? settings.data[0] : settings.data[c];
    if (color_space == 2 && c > 0 && !term.mu) term.mu_scale = term.mu_scale * term.mu_chroma;
    if (settings.method == Method::mmse) term.mmse = true;
    std::string why;
    if (!make_problem(k, term, why)) {
      status = 4;
      message = "component " + std::to_string(c) + ": " + why;
      return false;
    }
    frame.components.push_back(std::move(k));
  }
  status = 0;
  return true;
}

namespace {

// The defaults of a plan: the ratio of the steps, the tolerance and the most iterations,
// each of TV and of TGV.
struct PlanDefaults {
  double tv_ratio, tgv_ratio;
  double tv_tolerance, tgv_tolerance;
  uint64_t tv_iterations, tgv_iterations;
};

bool plan_with(const Settings& s, size_t channels, const PlanDefaults& d, Plan& plan, std::string& message);

}  // namespace

bool make_plan(const Settings& s, size_t channels, Plan& plan, std::string& message) {
  return plan_with(s, channels, PlanDefaults{10.0, 3.0, 2e-4, 2e-3, 5000u, 10000u}, plan, message);
}

bool make_samples_plan(const Settings& s, size_t channels, Plan& plan, std::string& message) {
  return plan_with(s, channels, PlanDefaults{10.0, 0.3, 1e-4, 1e-4, 10000u, 10000u}, plan, message);
}

namespace {

bool plan_with(const Settings& s, size_t channels, const PlanDefaults& d, Plan& plan, std::string& message) {
  std::vector<std::string> errors;
  if (s.channel_weights && s.channel_weights->size() != channels) {
    errors.push_back("the channel weights are one for each of the " + std::to_string(channels) + " channels, not " +
                     std::to_string(s.channel_weights->size()));
  }
  if (s.method == Method::tgv && s.partial_tolerance > 0.0 && !s.partial_radius) {
    errors.push_back("the partial tolerance needs the radius of the partial gap");
  }
  if (!errors.empty()) {
    message.clear();
    for (const std::string& e : errors) message += (message.empty() ? "" : "; ") + e;
    return false;
  }
  const bool tgv = s.method == Method::tgv;
  plan.gamma = s.channel_weights ? *s.channel_weights : std::vector<double>(channels, 1.0);
  double largest = 0.0;
  for (const double w : plan.gamma) largest = std::fmax(largest, w);
  const double bound = tgv ? (17.0 + std::sqrt(33.0)) / 2.0 : 8.0;
  const double norm_squared = s.norm_squared ? *s.norm_squared : bound * (largest * largest);
  const double weight = tgv ? s.alpha1 : s.alpha;
  const double default_ratio = tgv ? d.tgv_ratio : d.tv_ratio;
  const double default_tolerance = tgv ? d.tgv_tolerance : d.tv_tolerance;
  const double ratio = s.step_ratio ? *s.step_ratio : (s.weight_scaling ? default_ratio / (weight * weight) : default_ratio);
  plan.tolerance = s.tolerance ? *s.tolerance : (s.weight_scaling ? default_tolerance * weight : default_tolerance);
  plan.iterations = s.iterations ? *s.iterations : (tgv ? d.tgv_iterations : d.tv_iterations);
  plan.tau = std::sqrt((ratio * s.step_product) / norm_squared);
  plan.sigma = std::sqrt(s.step_product / (ratio * norm_squared));
  plan.relative_tolerance = s.relative_tolerance;
  plan.partial_tolerance = s.partial_tolerance;
  plan.has_partial = tgv && s.partial_radius.has_value();
  plan.partial_radius = s.partial_radius ? *s.partial_radius : 0.0;
  plan.relaxation = s.relaxation ? *s.relaxation : 1.9;
  plan.record_every = s.record_every;
  plan.free_radius = s.free_radius;
  return true;
}

// Whether x < n / m exactly, for a double x within a factor 2 of n / m and whole n, m > 0
// below 2^53: x m = p + e with p = x m rounded and e = fma(x, m, -p) exactly; p - n is
// exact (Sterbenz) and, where not 0, a multiple of p's unit in the last place, so larger
// than |e| and its sign is the sign of x m - n; where 0, the sign is e's.
bool lies_below(double x, double n, double m) {
  const double p = x * m;
  const double e = std::fma(x, m, -p);
  const double d = p - n;
  return d < 0.0 || (d == 0.0 && e < 0.0);
}

// The least double at or above n / m, and the greatest double below it.
double at_or_above(double n, double m) {
  const double near = n / m;
  return lies_below(near, n, m) ? std::nextafter(near, std::numeric_limits<double>::infinity()) : near;
}
double below(double n, double m) {
  const double near = n / m;
  return lies_below(near, n, m) ? near : std::nextafter(near, -std::numeric_limits<double>::infinity());
}

}  // namespace

bool samples_data_term(const Settings& settings, double& mu, double& slack, std::string& message) {
  if (settings.data.size() != 1) {
    message = "a picture of rounded samples takes one value of each option of the data term, not one for each "
              "component";
    return false;
  }
  const DataTerm& term = settings.data[0];
  const DataTerm defaults;
  if (term.mu_scale != defaults.mu_scale || term.mu_power != defaults.mu_power ||
      term.mu_chroma != defaults.mu_chroma || term.dc_weight != defaults.dc_weight || term.mmse != defaults.mmse ||
      term.power != defaults.power || term.slack_cost != defaults.slack_cost ||
      term.flat_weight != defaults.flat_weight || term.flat_levels != defaults.flat_levels ||
      term.flat_dc != defaults.flat_dc) {
    message = "of the data term, a picture of rounded samples takes --mu and --slack alone: the others are for JPEG "
              "files";
    return false;
  }
  mu = term.mu.value_or(0.0);
  slack = term.slack;
  return true;
}

bool make_samples(const uint16_t* levels, uint32_t channels, uint32_t height, uint32_t width, uint32_t bits,
                  const Settings& settings, SamplesModel& model, std::string& message, int32_t& status) {
  status = 4;
  double mu = 0.0, slack = 0.0;
  if (!samples_data_term(settings, mu, slack, message)) return false;
  const size_t count = static_cast<size_t>(channels) * height * width;
  if (levels == nullptr || count == 0) {
    message = "a picture of " + std::to_string(channels) + " x " + std::to_string(height) + " x " +
              std::to_string(width) + " levels, not empty";
    return false;
  }
  if (bits != 1 && bits != 2 && bits != 4 && bits != 8 && bits != 16) {
    message = "samples of 1, 2, 4, 8 or 16 bits, not " + std::to_string(bits);
    return false;
  }
  if (!(at_least_zero(mu) && at_least_zero(slack))) {
    message = "the weight of the data term and the slack are at least 0 and finite";
    return false;
  }
  const uint32_t top = (1u << bits) - 1u;
  for (size_t k = 0; k < count; ++k) {
    if (levels[k] > top) {
      message = "a level of " + std::to_string(bits) + " bits is at most " + std::to_string(top) + ", not " +
                std::to_string(levels[k]);
      return false;
    }
  }
  model.channels = channels;
  model.height = height;
  model.width = width;
  model.bits = bits;
  model.mu = mu;
  model.slack = slack;
  model.levels.assign(levels, levels + count);
  model.low.resize(top + 1);
  model.high.resize(top + 1);
  model.centre.resize(top + 1);
  const double wide = static_cast<double>(top), scale = 255.0 / wide;
  for (uint32_t v = 0; v <= top; ++v) {
    const double own = static_cast<double>(v);
    double lower, upper;
    if (slack == 0.0) {
      lower = v == 0 ? 0.0 : at_or_above((2.0 * own - 1.0) * 255.0, 2.0 * wide);
      upper = v == top ? 255.0 : below((2.0 * own + 1.0) * 255.0, 2.0 * wide);
    } else {
      lower = ((own - 0.5) - slack) * scale;
      upper = ((own + 0.5) + slack) * scale;
    }
    model.low[v] = std::max(lower, 0.0);
    model.high[v] = std::min(upper, 255.0);
    model.centre[v] = (own * 255.0) / wide;
  }
  status = 0;
  return true;
}

}  // namespace unround
