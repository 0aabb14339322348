// Reading and writing the settings' options (rust/unround/src/cli.rs:338-720).
#include "settings.hpp"

#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string_view>

namespace unround {

namespace {

const std::set<std::string_view> valued = {
    "method",         "alpha",        "alpha1",         "alpha0",
    "channel-weights", "channels",    "mu",             "mu-scale",
    "mu-power",       "mu-chroma",    "weight-power",   "dc-weight",
    "centres",        "slack",        "slack-cost",     "flat-weight",
    "flat-levels",    "flat-dc",      "iterations",     "tolerance",
    "relative-tolerance", "partial-tolerance", "partial-radius", "step-ratio",
    "relaxation",     "step-product", "norm-squared",   "record-every",
    "free-radius",    "threads",      "subgradient-iterations", "subgradient-step",
    "subgradient-decay", "subgradient-record-every", "max-pixels", "max-scans",
    "precision",
};
const std::set<std::string_view> flags = {"no-weight-scaling", "no-momentum", "warnings-are-errors"};
// The command line's options of its output and reports, which are no settings
// (rust/unround/src/cli.rs:543-559).
const std::set<std::string_view> output_options = {
    "format", "bits", "dither", "dither-channels", "compression", "ycbcr", "orientation", "no-icc",
    "overwrite", "report", "verbose", "quiet", "version", "help", "project-only",
};

std::optional<double> real_of(std::string_view text) {
  if (text.empty()) return std::nullopt;
  std::string copy(text);
  char* end = nullptr;
  const double value = std::strtod(copy.c_str(), &end);
  if (end != copy.c_str() + copy.size()) return std::nullopt;
  // Words such as "inf" and "nan" read as numbers and are refused by the ranges.
  return value;
}

std::optional<uint64_t> whole_of(std::string_view text) {
  uint64_t value = 0;
  const auto [stop, error] = std::from_chars(text.data(), text.data() + text.size(), value);
  if (error != std::errc{} || stop != text.data() + text.size() || text.empty()) return std::nullopt;
  return value;
}

std::vector<std::string_view> items_of(std::string_view text) {
  std::vector<std::string_view> items;
  size_t start = 0;
  while (true) {
    const size_t comma = text.find(',', start);
    items.push_back(text.substr(start, comma == std::string_view::npos ? std::string_view::npos : comma - start));
    if (comma == std::string_view::npos) break;
    start = comma + 1;
  }
  return items;
}

bool positive(double v) { return v > 0.0 && std::isfinite(v); }
bool at_least_zero(double v) { return v >= 0.0 && std::isfinite(v); }

struct Reader {
  std::map<std::string, std::string, std::less<>> given;
  std::set<std::string, std::less<>> flags_given;
  std::vector<std::string> errors;

  std::optional<std::string_view> text(std::string_view name) const {
    const auto found = given.find(name);
    if (found == given.end()) return std::nullopt;
    return std::string_view(found->second);
  }

  template <typename Check>
  std::optional<double> real(std::string_view name, Check check, std::string_view range) {
    const auto value_text = text(name);
    if (!value_text) return std::nullopt;
    const auto value = real_of(*value_text);
    if (!value || !check(*value)) {
      errors.push_back("--" + std::string(name) + " is " + std::string(range) + ", not " + std::string(*value_text));
      return std::nullopt;
    }
    return value;
  }

  std::optional<uint64_t> count(std::string_view name, uint64_t least) {
    const auto value_text = text(name);
    if (!value_text) return std::nullopt;
    const auto value = whole_of(*value_text);
    if (!value || *value < least) {
      errors.push_back("--" + std::string(name) + " is a whole number of at least " + std::to_string(least) +
                       ", not " + std::string(*value_text));
      return std::nullopt;
    }
    return value;
  }

  template <typename T, typename Read>
  std::optional<std::vector<T>> list(std::string_view name, Read read, std::string_view range) {
    const auto value_text = text(name);
    if (!value_text) return std::nullopt;
    std::vector<T> values;
    for (const std::string_view item : items_of(*value_text)) {
      const std::optional<T> value = read(item);
      if (!value) {
        errors.push_back("--" + std::string(name) + " is " + std::string(range) + ", not " + std::string(item));
        return std::nullopt;
      }
      values.push_back(*value);
    }
    return values;
  }
};

std::optional<double> at_least_zero_item(std::string_view item) {
  const auto value = real_of(item);
  return value && at_least_zero(*value) ? value : std::nullopt;
}

}  // namespace

bool parse_settings(const std::vector<std::string>& arguments, Settings& settings, std::string& message) {
  Reader reader;
  for (size_t i = 0; i < arguments.size(); ++i) {
    const std::string_view argument = arguments[i];
    if (argument.size() < 2 || argument.substr(0, 2) != "--") {
      reader.errors.push_back("settings are options alone, without files: " + std::string(argument));
      continue;
    }
    const size_t equals = argument.find('=');
    const std::string name(argument.substr(2, equals == std::string_view::npos ? std::string_view::npos : equals - 2));
    if (output_options.count(name) != 0) {
      reader.errors.push_back("--" + name + " is an option of the command line, not a setting");
      if (equals == std::string_view::npos && i + 1 < arguments.size() && arguments[i + 1].rfind("--", 0) != 0 &&
          name != "ycbcr" && name != "no-icc" && name != "overwrite" && name != "verbose" && name != "quiet" &&
          name != "version" && name != "help") {
        ++i;
      }
      continue;
    }
    if (flags.count(name) != 0) {
      if (equals != std::string_view::npos) reader.errors.push_back("--" + name + " takes no value");
      if (!reader.flags_given.insert(name).second) reader.errors.push_back("--" + name + " is given twice");
      continue;
    }
    if (valued.count(name) == 0) {
      reader.errors.push_back("--" + name + " is not an option");
      if (equals == std::string_view::npos && i + 1 < arguments.size() && arguments[i + 1].rfind("--", 0) != 0) ++i;
      continue;
    }
    std::string value;
    if (equals != std::string_view::npos) {
      value = std::string(argument.substr(equals + 1));
    } else if (i + 1 < arguments.size()) {
      value = arguments[++i];
    } else {
      reader.errors.push_back("--" + name + " needs a value");
      continue;
    }
    if (!reader.given.emplace(name, value).second) reader.errors.push_back("--" + name + " is given twice");
  }

  Settings out;
  if (const auto method = reader.text("method")) {
    if (*method == "mmse") out.method = Method::mmse;
    else if (*method == "tv") out.method = Method::tv;
    else if (*method == "tgv") out.method = Method::tgv;
    else if (*method == "subgradient") out.method = Method::subgradient;
    else reader.errors.push_back("--method is one of mmse, tv, tgv, subgradient, not " + std::string(*method));
  }
  if (const auto v = reader.real("alpha", positive, "positive and finite")) out.alpha = *v;
  if (const auto v = reader.real("alpha1", positive, "positive and finite")) out.alpha1 = *v;
  if (const auto v = reader.real("alpha0", positive, "positive and finite")) out.alpha0 = *v;
  out.channel_weights = reader.list<double>(
      "channel-weights",
      [](std::string_view item) -> std::optional<double> {
        const auto v = real_of(item);
        return v && positive(*v) ? v : std::nullopt;
      },
      "a list of positive, finite weights");
  if (const auto coupling = reader.text("channels")) {
    if (*coupling == "coupled") out.coupled = true;
    else if (*coupling == "apart") out.coupled = false;
    else reader.errors.push_back("--channels is one of coupled, apart, not " + std::string(*coupling));
  }

  // The data term: each option one value for every component, or one for each.
  const auto mu = reader.list<std::optional<double>>(
      "mu",
      [](std::string_view item) -> std::optional<std::optional<double>> {
        if (item == "rule") return std::optional<double>{};
        const auto v = at_least_zero_item(item);
        if (!v) return std::nullopt;
        return std::optional<double>(*v);
      },
      "at least 0 and finite, or rule");
  const auto mu_scale = reader.list<double>("mu-scale", at_least_zero_item, "at least 0 and finite");
  const auto mu_power = reader.list<double>(
      "mu-power",
      [](std::string_view item) -> std::optional<double> {
        const auto v = real_of(item);
        return v && std::isfinite(*v) ? v : std::nullopt;
      },
      "finite");
  const auto mu_chroma = reader.list<double>("mu-chroma", at_least_zero_item, "at least 0 and finite");
  const auto power = reader.list<double>("weight-power", at_least_zero_item, "at least 0 and finite");
  const auto dc_weight = reader.list<double>("dc-weight", at_least_zero_item, "at least 0 and finite");
  const auto slack = reader.list<double>("slack", at_least_zero_item, "at least 0 and finite");
  const auto slack_cost = reader.list<double>("slack-cost", at_least_zero_item, "at least 0 and finite");
  const auto centres = reader.list<bool>(
      "centres",
      [](std::string_view item) -> std::optional<bool> {
        if (item == "mmse") return true;
        if (item == "midpoint") return false;
        return std::nullopt;
      },
      "mmse or midpoint");
  const auto flat_weight = reader.list<double>("flat-weight", at_least_zero_item, "at least 0 and finite");
  const auto flat_levels = reader.list<uint32_t>(
      "flat-levels",
      [](std::string_view item) -> std::optional<uint32_t> {
        const auto v = whole_of(item);
        if (!v || *v > 63u) return std::nullopt;
        return static_cast<uint32_t>(*v);
      },
      "an integer from 0 to 63");
  const auto flat_dc = reader.list<std::optional<uint32_t>>(
      "flat-dc",
      [](std::string_view item) -> std::optional<std::optional<uint32_t>> {
        if (item == "any") return std::optional<uint32_t>{};
        const auto v = whole_of(item);
        if (!v || *v > 0xffffffffull) return std::nullopt;
        return std::optional<uint32_t>(static_cast<uint32_t>(*v));
      },
      "a whole number or any");
  std::set<size_t> lengths;
  const auto note = [&lengths](const auto& values) {
    if (values && values->size() > 1) lengths.insert(values->size());
  };
  note(mu), note(mu_scale), note(mu_power), note(mu_chroma), note(power), note(dc_weight), note(slack);
  note(slack_cost), note(centres), note(flat_weight), note(flat_levels), note(flat_dc);
  if (lengths.size() > 1) {
    reader.errors.push_back("the lists of the data term have one value for each component, and so one length");
  }
  const size_t count = lengths.empty() ? 1 : *lengths.begin();
  const auto pick = [](const auto& values, size_t index, auto& into) {
    if (values) into = (*values)[values->size() == 1 ? 0 : std::min(index, values->size() - 1)];
  };
  out.data.assign(count, DataTerm{});
  for (size_t c = 0; c < count; ++c) {
    DataTerm& term = out.data[c];
    pick(mu, c, term.mu);
    pick(mu_scale, c, term.mu_scale);
    pick(mu_power, c, term.mu_power);
    pick(mu_chroma, c, term.mu_chroma);
    pick(power, c, term.power);
    pick(dc_weight, c, term.dc_weight);
    pick(slack, c, term.slack);
    pick(slack_cost, c, term.slack_cost);
    pick(centres, c, term.mmse);
    pick(flat_weight, c, term.flat_weight);
    pick(flat_levels, c, term.flat_levels);
    pick(flat_dc, c, term.flat_dc);
  }

  if (const auto v = reader.count("iterations", 0)) out.iterations = *v;
  out.tolerance = reader.real("tolerance", at_least_zero, "at least 0 and finite");
  if (const auto v = reader.real("relative-tolerance", at_least_zero, "at least 0 and finite")) {
    out.relative_tolerance = *v;
  }
  if (const auto v = reader.real("partial-tolerance", at_least_zero, "at least 0 and finite")) {
    out.partial_tolerance = *v;
  }
  out.partial_radius = reader.real("partial-radius", at_least_zero, "at least 0 and finite");
  out.step_ratio = reader.real("step-ratio", positive, "positive and finite");
  out.relaxation = reader.real("relaxation", [](double v) { return v > 0.0 && v < 2.0; }, "in (0, 2)");
  if (const auto v = reader.real("step-product", [](double x) { return x > 0.0 && x < 1.0; }, "in (0, 1)")) {
    out.step_product = *v;
  }
  out.norm_squared = reader.real("norm-squared", positive, "positive and finite");
  out.weight_scaling = reader.flags_given.count("no-weight-scaling") == 0;
  if (const auto v = reader.count("record-every", 0)) out.record_every = *v;
  if (const auto v = reader.real("free-radius", at_least_zero, "at least 0 and finite")) out.free_radius = *v;
  if (const auto v = reader.count("threads", 0)) out.threads = *v;
  if (const auto v = reader.count("subgradient-iterations", 0)) out.subgradient_iterations = *v;
  if (const auto v = reader.real("subgradient-step", positive, "positive and finite")) out.subgradient_step = *v;
  if (const auto v = reader.real("subgradient-decay", at_least_zero, "at least 0 and finite")) {
    out.subgradient_decay = *v;
  }
  out.momentum = reader.flags_given.count("no-momentum") == 0;
  if (const auto v = reader.count("subgradient-record-every", 0)) out.subgradient_record_every = *v;
  if (const auto v = reader.count("max-pixels", 1)) out.max_pixels = *v;
  if (const auto v = reader.count("max-scans", 1)) {
    if (*v > 0x7fffffffull) {
      reader.errors.push_back("--max-scans is a whole number of at least 1, not " + std::to_string(*v));
    } else {
      out.max_scans = *v;
    }
  }
  out.warnings_are_errors = reader.flags_given.count("warnings-are-errors") != 0;
  if (const auto precision = reader.text("precision")) {
    if (*precision == "single") out.precision = Precision::single;
    else if (*precision == "double") out.precision = Precision::dual;
    else if (*precision == "mixed") out.precision = Precision::mixed;
    else reader.errors.push_back("--precision is one of single, double, mixed, not " + std::string(*precision));
  }
  // What the method does not take (cpp/src/settings.cpp:396-401).
  if (out.method == Method::tv && (out.partial_radius || out.partial_tolerance > 0.0)) {
    reader.errors.push_back("the partial gap is TGV's alone, not TV's");
  }
  if (out.partial_tolerance > 0.0 && !out.partial_radius) {
    reader.errors.push_back("--partial-tolerance needs --partial-radius");
  }
  if (!reader.errors.empty()) {
    message.clear();
    for (const std::string& error : reader.errors) message += (message.empty() ? "" : "; ") + error;
    return false;
  }
  settings = std::move(out);
  return true;
}

std::string decimal(double value) {
  if (std::isnan(value)) return "NaN";
  if (std::isinf(value)) return value > 0 ? "inf" : "-inf";
  if (value == 0.0) return std::signbit(value) ? "-0.0" : "0.0";
  // The shortest digits and exponent that read back as the value.
  char buffer[64];
  const auto [end, error] = std::to_chars(buffer, buffer + sizeof buffer, value, std::chars_format::scientific);
  std::string text(buffer, end);
  const size_t e = text.find('e');
  std::string mantissa = text.substr(0, e);
  const int exponent = std::atoi(text.c_str() + e + 1);
  const bool negative = mantissa[0] == '-';
  if (negative) mantissa.erase(0, 1);
  std::string digits;
  for (const char ch : mantissa)
    if (ch != '.') digits += ch;
  std::string out;
  if (exponent < -5 || exponent >= 16) {
    // Rust writes 1e-7, 1.5e16.
    out = digits.substr(0, 1);
    if (digits.size() > 1) out += "." + digits.substr(1);
    out += "e" + std::to_string(exponent);
  } else if (exponent < 0) {
    out = "0." + std::string(static_cast<size_t>(-exponent - 1), '0') + digits;
  } else {
    const size_t point = static_cast<size_t>(exponent) + 1;
    if (digits.size() <= point) {
      out = digits + std::string(point - digits.size(), '0') + ".0";
    } else {
      out = digits.substr(0, point) + "." + digits.substr(point);
    }
  }
  return negative ? "-" + out : out;
}

std::string options_text(const Settings& s) {
  std::vector<std::string> options;
  const auto put = [&options](const std::string& name, const std::string& value) {
    options.push_back("--" + name + "=" + value);
  };
  static const char* const methods[] = {"mmse", "tv", "tgv", "subgradient"};
  put("method", methods[static_cast<int>(s.method)]);
  put("alpha", decimal(s.alpha));
  put("alpha1", decimal(s.alpha1));
  put("alpha0", decimal(s.alpha0));
  if (s.channel_weights) {
    std::string weights;
    for (const double w : *s.channel_weights) weights += (weights.empty() ? "" : ",") + decimal(w);
    put("channel-weights", weights);
  }
  put("channels", s.coupled ? "coupled" : "apart");
  const auto listed = [&s](auto text) {
    std::string first = text(s.data.front());
    bool alike = true;
    for (const DataTerm& term : s.data) alike = alike && text(term) == first;
    if (alike) return first;
    std::string all;
    for (const DataTerm& term : s.data) all += (all.empty() ? "" : ",") + text(term);
    return all;
  };
  put("mu", listed([](const DataTerm& t) { return t.mu ? decimal(*t.mu) : std::string("rule"); }));
  put("mu-scale", listed([](const DataTerm& t) { return decimal(t.mu_scale); }));
  put("mu-power", listed([](const DataTerm& t) { return decimal(t.mu_power); }));
  put("mu-chroma", listed([](const DataTerm& t) { return decimal(t.mu_chroma); }));
  put("weight-power", listed([](const DataTerm& t) { return decimal(t.power); }));
  put("dc-weight", listed([](const DataTerm& t) { return decimal(t.dc_weight); }));
  put("centres", listed([](const DataTerm& t) { return std::string(t.mmse ? "mmse" : "midpoint"); }));
  put("slack", listed([](const DataTerm& t) { return decimal(t.slack); }));
  put("slack-cost", listed([](const DataTerm& t) { return decimal(t.slack_cost); }));
  put("flat-weight", listed([](const DataTerm& t) { return decimal(t.flat_weight); }));
  put("flat-levels", listed([](const DataTerm& t) { return std::to_string(t.flat_levels); }));
  put("flat-dc", listed([](const DataTerm& t) { return t.flat_dc ? std::to_string(*t.flat_dc) : std::string("any"); }));
  if (s.iterations) put("iterations", std::to_string(*s.iterations));
  if (s.tolerance) put("tolerance", decimal(*s.tolerance));
  put("relative-tolerance", decimal(s.relative_tolerance));
  put("partial-tolerance", decimal(s.partial_tolerance));
  if (s.partial_radius) put("partial-radius", decimal(*s.partial_radius));
  if (s.step_ratio) put("step-ratio", decimal(*s.step_ratio));
  if (s.relaxation) put("relaxation", decimal(*s.relaxation));
  put("step-product", decimal(s.step_product));
  if (s.norm_squared) put("norm-squared", decimal(*s.norm_squared));
  put("record-every", std::to_string(s.record_every));
  put("free-radius", decimal(s.free_radius));
  put("threads", std::to_string(s.threads));
  put("subgradient-iterations", std::to_string(s.subgradient_iterations));
  put("subgradient-step", decimal(s.subgradient_step));
  put("subgradient-decay", decimal(s.subgradient_decay));
  put("subgradient-record-every", std::to_string(s.subgradient_record_every));
  if (s.max_pixels != 0) put("max-pixels", std::to_string(s.max_pixels));
  if (s.max_scans != 0) put("max-scans", std::to_string(s.max_scans));
  static const char* const precisions[] = {"single", "double", "mixed"};
  put("precision", precisions[static_cast<int>(s.precision)]);
  if (!s.weight_scaling) options.push_back("--no-weight-scaling");
  if (!s.momentum) options.push_back("--no-momentum");
  if (s.warnings_are_errors) options.push_back("--warnings-are-errors");
  std::string text;
  for (const std::string& option : options) text += (text.empty() ? "" : " ") + option;
  return text;
}

}  // namespace unround
