// The C interface (include/unround.h): settings from options, the solve of components
// given as arrays on the GPU, and the result's accessors (rust/capi/src/lib.rs:99-117,
// 459-694; rust/unround/src/decode.rs:269-358). No C++ exception crosses it: each
// function that computes catches them and returns UNROUND_ERROR_INTERNAL.
#include "unround.h"
#include "unround_ext.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "files.hpp"
#include "model.hpp"
#include "settings.hpp"
#include "solver.hpp"

struct unround_settings {
  unround::Settings settings;
};

struct unround_result {
  unround::FrameModel frame;
  unround::Settings settings;
  unround::SolveOutput out;
  bool solver = false;
  std::vector<double> planes;
  std::vector<double> picture;
  uint64_t picture_channels = 1;
  std::vector<uint64_t> record_iterations;
  std::array<std::vector<double>, 5> record_values;
  // The problem's arrays, made once on demand: lower, upper, centres, steps, weights.
  mutable std::mutex made_guard;
  mutable std::vector<std::array<std::vector<double>, 5>> problems;
  mutable std::vector<bool> made;
  std::string warning;
  // A picture of rounded samples (a PNG file's): the format, the bits of its samples, and
  // the alpha's solve where it has alpha; its frame has no components.
  int32_t format = UNROUND_FORMAT_JPEG;
  int32_t bits = 8;
  bool has_alpha = false;
  unround::SolveOutput alpha;
  std::vector<uint64_t> alpha_record_iterations;
  std::array<std::vector<double>, 5> alpha_record_values;
  std::vector<uint8_t> icc_profile;
  int32_t exif_orientation = 0;
  // Of a projection, each component's coefficients' excess; empty for a reconstruction.
  std::vector<std::vector<double>> excess;
};

// The unrolled solve's handle: its frames (with their settings) and plans, which the solve
// reads, and the solve.
struct unround_unrolled {
  std::vector<std::unique_ptr<unround_result>> frames;
  std::vector<unround::Plan> plans;
  std::unique_ptr<unround::UnrolledSolve> solve;
};

namespace {

const char* const kVersion = "unround 0.1.0 (MI355X, HIP gfx950, C interface)";

// The message NUL-terminated, cut at a UTF-8 boundary (rust/capi/src/lib.rs:84-96).
void write_message(char* buffer, size_t size, const std::string& text) {
  if (buffer == nullptr || size == 0) return;
  size_t end = std::min(text.size(), size - 1);
  while (end > 0 && end < text.size() && (static_cast<unsigned char>(text[end]) & 0xC0u) == 0x80u) --end;
  std::memcpy(buffer, text.data(), end);
  buffer[end] = '\0';
}

template <typename Body>
unround_status guarded(char* message, size_t size, Body body) {
  try {
    std::string text;
    const unround_status status = body(text);
    if (status != UNROUND_OK) write_message(message, size, text);
    return status;
  } catch (const std::exception& error) {
    write_message(message, size, std::string("an internal error of JPEG-Unround: ") + error.what());
    return UNROUND_ERROR_INTERNAL;
  } catch (...) {
    write_message(message, size, "an internal error of JPEG-Unround");
    return UNROUND_ERROR_INTERNAL;
  }
}

// JFIF's RGB of Y, Cb and Cr (docs/math.md:1014-1024): the doubles nearest to 701/500,
// 443/250, 25251/73375 and 209599/293500.
constexpr double kCr = 1.402;
constexpr double kCb = 1.772;
const double kGb = 25251.0 / 73375.0;
const double kGr = 209599.0 / 293500.0;

void rgb_of(const double* planes, size_t count, double* picture) {
  const double* y = planes;
  const double* cb = planes + count;
  const double* cr = planes + 2 * count;
  for (size_t k = 0; k < count; ++k) {
    const double b = cb[k] - 128.0, r = cr[k] - 128.0;
    picture[3 * k + 0] = y[k] + kCr * r;
    picture[3 * k + 1] = (y[k] - kGb * b) - kGr * r;
    picture[3 * k + 2] = y[k] + kCb * b;
  }
}

// A solve's records as the accessors give them: iterations, and seconds, primal, dual,
// theta and partial.
void records_of(const unround::SolveOutput& out, std::vector<uint64_t>& iterations,
                std::array<std::vector<double>, 5>& values) {
  iterations.clear();
  for (auto& column : values) column.clear();
  for (const unround::Record& r : out.records) {
    iterations.push_back(r.iteration);
    values[0].push_back(r.seconds);
    values[1].push_back(r.primal);
    values[2].push_back(r.dual);
    values[3].push_back(r.theta);
    values[4].push_back(r.partial);
  }
}

// The canvas cut to the picture, and the picture of the colour space
// (rust/unround/src/decode.rs:306-358).
void assemble(unround_result& result) {
  const unround::FrameModel& f = result.frame;
  const size_t C = f.channels(), h = f.height, w = f.width;
  result.planes.resize(C * h * w);
  for (size_t c = 0; c < C; ++c)
    for (size_t i = 0; i < h; ++i)
      std::memcpy(&result.planes[(c * h + i) * w], &result.out.canvas[(c * f.rows + i) * f.columns], sizeof(double) * w);
  if (f.color_space == UNROUND_COLOR_YCBCR) {
    result.picture.resize(3 * h * w);
    rgb_of(result.planes.data(), h * w, result.picture.data());
    result.picture_channels = 3;
  } else if (C == 1) {
    result.picture = result.planes;
    result.picture_channels = 1;
  } else {
    result.picture.resize(C * h * w);
    for (size_t k = 0; k < h * w; ++k)
      for (size_t c = 0; c < C; ++c) result.picture[C * k + c] = result.planes[c * h * w + k];
    result.picture_channels = C;
  }
  records_of(result.out, result.record_iterations, result.record_values);
  result.problems.assign(C, {});
  result.made.assign(C, false);
}

// The checks of one input and its frame and plan, into `made` and `plan`.
unround_status frame_of(const unround_input* input, const unround_settings* settings, unround_result& made,
                        unround::Plan& plan, std::string& message) {
  if (input == nullptr || settings == nullptr) {
    message = "the input, the settings and the result are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  if (input->color_space < 1 || input->color_space > 3) {
    message = "a color space of 1, 2 or 3, not " + std::to_string(input->color_space);
    return UNROUND_ERROR_ARGUMENT;
  }
  if (input->components == nullptr || input->count == 0) {
    message = "components are given";
    return UNROUND_ERROR_ARGUMENT;
  }
  const unround::Settings& s = settings->settings;
  if (s.method == unround::Method::subgradient) {
    message = "the subgradient method is not part of this library: use --method tv, tgv or mmse";
    return UNROUND_ERROR_UNSUPPORTED;
  }
  std::vector<unround::ComponentInput> inputs;
  for (uint32_t c = 0; c < input->count; ++c) {
    const unround_component& given = input->components[c];
    if (given.levels == nullptr) {
      message = "a component's levels are a null pointer";
      return UNROUND_ERROR_ARGUMENT;
    }
    if (given.h_samp_factor < 0 || given.v_samp_factor < 0) {
      message = "a sampling factor below 0";
      return UNROUND_ERROR_ARGUMENT;
    }
    unround::ComponentInput in;
    in.levels = given.levels;
    in.rows = given.rows;
    in.columns = given.columns;
    in.h = given.h_samp_factor;
    in.v = given.v_samp_factor;
    std::memcpy(in.table.data(), given.quant_table, sizeof(uint16_t) * 64);
    inputs.push_back(in);
  }
  int32_t status = UNROUND_OK;
  if (!unround::make_frame(input->height, input->width, input->color_space, inputs, s, made.frame, message, status)) {
    return static_cast<unround_status>(status);
  }
  const bool solving = s.method == unround::Method::tv || s.method == unround::Method::tgv;
  if (solving && !unround::make_plan(s, made.frame.channels(), plan, message)) return UNROUND_ERROR_OPTIONS;
  if (!solving) plan.gamma.assign(made.frame.channels(), 1.0);
  made.settings = s;
  made.solver = solving;
  return UNROUND_OK;
}

// The UNROUND_TIMING environment variable: 1 times the runs of iterations between records,
// 2 also every launch. UNROUND_TIMING_WARMUP=W leaves a solve's first W iterations untimed.
unround::Timing timing_asked() {
  const char* timing = std::getenv("UNROUND_TIMING");
  if (timing == nullptr) return unround::Timing::none;
  if (std::string(timing) == "2") return unround::Timing::kernels;
  if (std::string(timing) == "1") return unround::Timing::iterations;
  return unround::Timing::none;
}

// A record's callback to the caller's observer, for a canvas of C x H x W samples.
unround::Observer watcher(unround_observer observer, void* user, uint64_t channels, uint64_t height, uint64_t width) {
  if (observer == nullptr) return {};
  const double samples = static_cast<double>(channels) * static_cast<double>(height) * static_cast<double>(width);
  return [observer, user, channels, height, width, samples](const unround::Record& record, const double* canvas) {
    unround_record raw{};
    raw.iteration = record.iteration;
    raw.gap = (record.primal - record.dual) / samples;
    raw.primal = record.primal;
    raw.dual = record.dual;
    raw.channels = channels;
    raw.height = height;
    raw.width = width;
    raw.canvas = canvas;
    return observer(user, &raw) != 0;
  };
}

unround_status reconstruct(const unround_input* input, const unround_settings* settings, const unround_start* start,
                           unround_observer observer, void* user, unround_result** result, std::string& message) {
  if (result == nullptr) {
    message = "the input, the settings and the result are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  auto made = std::make_unique<unround_result>();
  unround::Plan plan;
  const unround_status checked = frame_of(input, settings, *made, plan, message);
  if (checked != UNROUND_OK) return checked;
  const unround::Settings& s = settings->settings;
  int32_t status = UNROUND_OK;
  unround::StartState state;
  if (start != nullptr) {
    if (start->coefficients != nullptr)
      for (uint32_t c = 0; c < input->count; ++c) state.coefficients.push_back(start->coefficients[c]);
    state.p = start->p;
    state.w = start->w;
    state.r = start->r;
    if (s.method == unround::Method::tv && (state.w != nullptr || state.r != nullptr)) {
      message = "a start of TV has no fields w and r";
      return UNROUND_ERROR_OPTIONS;
    }
  }
  const unround::FrameModel& frame = made->frame;
  const unround::Observer watch = watcher(observer, user, frame.channels(), frame.rows, frame.columns);
  if (!unround::solve_on_gpu(frame, s, plan, start != nullptr ? &state : nullptr, watch, timing_asked(), made->out,
                             message, status)) {
    return static_cast<unround_status>(status);
  }
  assemble(*made);
  *result = made.release();
  return UNROUND_OK;
}

// A picture of rounded samples, interleaved, reconstructed (rust/unround/src/rounded.rs:
// 71-170): the colours by the settings' method and weights of channels, the alpha alone.
unround_status reconstruct_samples(const uint16_t* levels, uint64_t height, uint64_t width, uint64_t channels,
                                   uint32_t bits, const unround_settings* settings, unround_observer observer,
                                   void* user, unround_result** result, std::string& message) {
  if (result == nullptr || settings == nullptr || levels == nullptr) {
    message = "the levels, the settings and the result are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  *result = nullptr;
  if (channels < 1 || channels > 4) {
    message = "a picture of 1 to 4 channels (grey, grey and alpha, RGB, RGB and alpha), not " +
              std::to_string(channels);
    return UNROUND_ERROR_ARGUMENT;
  }
  if (height == 0 || width == 0 || height > (uint64_t{1} << 31) || width > (uint64_t{1} << 31) ||
      height * width > (uint64_t{1} << 40)) {
    message = "a picture of at least one and at most 2^40 pixels";
    return UNROUND_ERROR_ARGUMENT;
  }
  unround::Settings s = settings->settings;
  if (s.method != unround::Method::tv && s.method != unround::Method::tgv) {
    message = "a picture of rounded samples is reconstructed by TV or TGV";
    return UNROUND_ERROR_OPTIONS;
  }
  const uint64_t colours = (channels == 2 || channels == 4) ? channels - 1 : channels;
  const size_t plane = static_cast<size_t>(height * width);
  const auto planes_of = [&](uint64_t first, uint64_t count) {
    std::vector<uint16_t> out(count * plane);
    for (uint64_t c = 0; c < count; ++c)
      for (size_t k = 0; k < plane; ++k) out[c * plane + k] = levels[k * channels + first + c];
    return out;
  };
  const auto h = static_cast<uint32_t>(height), w = static_cast<uint32_t>(width);
  int32_t status = UNROUND_OK;
  unround::SamplesModel colour_model;
  const std::vector<uint16_t> colour_levels = planes_of(0, colours);
  if (!unround::make_samples(colour_levels.data(), static_cast<uint32_t>(colours), h, w, bits, s, colour_model,
                             message, status)) {
    return static_cast<unround_status>(status);
  }
  unround::Plan plan;
  if (!unround::make_samples_plan(s, colours, plan, message)) return UNROUND_ERROR_OPTIONS;
  auto made = std::make_unique<unround_result>();
  if (!unround::solve_samples_on_gpu(colour_model, s, plan, watcher(observer, user, colours, height, width),
                                     timing_asked(), made->out, message, status)) {
    return static_cast<unround_status>(status);
  }
  if (colours < channels) {
    // The alpha alone: one channel, without the colours' weights of channels.
    unround::Settings alone = s;
    alone.channel_weights.reset();
    unround::SamplesModel alpha_model;
    const std::vector<uint16_t> alpha_levels = planes_of(channels - 1, 1);
    if (!unround::make_samples(alpha_levels.data(), 1, h, w, bits, alone, alpha_model, message, status)) {
      return static_cast<unround_status>(status);
    }
    unround::Plan alpha_plan;
    if (!unround::make_samples_plan(alone, 1, alpha_plan, message)) return UNROUND_ERROR_OPTIONS;
    if (!unround::solve_samples_on_gpu(alpha_model, alone, alpha_plan, watcher(observer, user, 1, height, width),
                                       timing_asked(), made->alpha, message, status)) {
      return static_cast<unround_status>(status);
    }
    made->has_alpha = true;
    records_of(made->alpha, made->alpha_record_iterations, made->alpha_record_values);
  }
  made->format = UNROUND_FORMAT_PNG;
  made->bits = static_cast<int32_t>(bits);
  made->settings = s;
  made->solver = true;
  made->frame.height = made->frame.rows = h;
  made->frame.width = made->frame.columns = w;
  made->frame.color_space = channels <= 2 ? UNROUND_COLOR_GRAYSCALE : UNROUND_COLOR_RGB;
  made->planes = made->out.canvas;
  if (made->has_alpha) made->planes.insert(made->planes.end(), made->alpha.canvas.begin(), made->alpha.canvas.end());
  made->picture.resize(channels * plane);
  for (uint64_t c = 0; c < channels; ++c)
    for (size_t k = 0; k < plane; ++k) made->picture[k * channels + c] = made->planes[c * plane + k];
  made->picture_channels = channels;
  records_of(made->out, made->record_iterations, made->record_values);
  *result = made.release();
  return UNROUND_OK;
}

// A JPEG file's components as the solve's input, over the image's arrays.
unround_input input_of(const unround::JpegImage& image, std::vector<unround_component>& components) {
  components.assign(image.components.size(), unround_component{});
  for (size_t c = 0; c < image.components.size(); ++c) {
    const unround::JpegComponent& from = image.components[c];
    unround_component& to = components[c];
    to.levels = from.coefficients.data();
    to.rows = from.blocks_down;
    to.columns = from.blocks_across;
    to.h_samp_factor = from.h;
    to.v_samp_factor = from.v;
    std::memcpy(to.quant_table, from.table.data(), sizeof(to.quant_table));
  }
  unround_input input{};
  input.height = image.height;
  input.width = image.width;
  input.color_space = image.color_space;
  input.count = static_cast<uint32_t>(components.size());
  input.components = components.data();
  return input;
}

// A JPEG file read with the settings' limits; a failure of the reading is a file that
// cannot be read (rust/unround/src/error.rs:46-50).
unround_status read_file(const uint8_t* data, size_t size, const unround::Settings& s, unround::JpegImage& image,
                         std::string& message) {
  int32_t status = 0;
  if (!unround::read_jpeg(data, size, s.max_pixels, static_cast<int32_t>(std::min<uint64_t>(s.max_scans, 0x7FFFFFFF)),
                          s.warnings_are_errors, image, message, status)) {
    return UNROUND_ERROR_READ;
  }
  return UNROUND_OK;
}

// A file reconstructed (rust/capi/src/lib.rs:532-565): a PNG file within the intervals of
// its rounded samples, any other as a JPEG file, its components by the settings.
unround_status decode_file(const uint8_t* data, size_t size, const unround_settings* settings,
                           unround_observer observer, void* user, unround_result** result, std::string& message) {
  if (data == nullptr || settings == nullptr || result == nullptr) {
    message = "the data, the settings and the result are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  *result = nullptr;
  if (!unround::is_png(data, size)) {
    unround::JpegImage image;
    const unround_status read = read_file(data, size, settings->settings, image, message);
    if (read != UNROUND_OK) return read;
    std::vector<unround_component> components;
    const unround_input input = input_of(image, components);
    const unround_status done = reconstruct(&input, settings, nullptr, observer, user, result, message);
    if (done == UNROUND_OK) {
      (*result)->icc_profile = std::move(image.icc_profile);
      (*result)->exif_orientation = image.exif_orientation;
      (*result)->warning = image.warnings > 0 ? image.warning : "";
    }
    return done;
  }
  // What a PNG file takes is checked before it is read (rust/unround/src/rounded.rs:77-95).
  const unround::Settings& s = settings->settings;
  double mu = 0.0, slack = 0.0;
  if (!unround::samples_data_term(s, mu, slack, message)) return UNROUND_ERROR_OPTIONS;
  if (s.method != unround::Method::tv && s.method != unround::Method::tgv) {
    message = "a PNG file is reconstructed by TV or TGV";
    return UNROUND_ERROR_OPTIONS;
  }
  unround::PngImage image;
  int32_t status = UNROUND_OK;
  if (!unround::read_png(data, size, s.max_pixels, image, message, status)) {
    // The file layer's statuses (rust/unround/src/rounded.rs:90-94): unsupported stays so,
    // a limit is the options', anything else a file that cannot be read.
    return status == 3 ? UNROUND_ERROR_UNSUPPORTED : (status == 4 ? UNROUND_ERROR_OPTIONS : UNROUND_ERROR_READ);
  }
  const unround_status done = reconstruct_samples(image.samples.data(), image.height, image.width, image.channels,
                                                  image.bits, settings, observer, user, result, message);
  if (done == UNROUND_OK) {
    (*result)->icc_profile = std::move(image.icc_profile);
    (*result)->warning = image.warning;
  }
  return done;
}

// A picture made elsewhere measured against a JPEG file's intervals and projected onto them
// (docs/math.md 12; rust/capi/src/lib.rs:708-752; rust/unround/src/picture.rs:60-137): its
// canvas as an encoder makes the file's (JFIF's Y, Cb and Cr of its RGB for a file in YCbCr,
// its own channels otherwise; each plane extended over whole MCUs by repeating its last
// column, then its last row), each coefficient's excess, and the projection as a
// reconstruction without a solver.
unround_status project_picture(const uint8_t* data, size_t size, const unround_settings* settings,
                               const double* picture, uint64_t height, uint64_t width, uint64_t channels,
                               unround_result** result, std::string& message) {
  if (data == nullptr || settings == nullptr || result == nullptr) {
    message = "the data, the settings and the result are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  *result = nullptr;
  if (unround::is_png(data, size)) {
    message = "a picture is projected onto the intervals of a JPEG file, not of a PNG file";
    return UNROUND_ERROR_OPTIONS;
  }
  const unsigned __int128 count = static_cast<unsigned __int128>(height) * width * channels;
  if (count > (static_cast<unsigned __int128>(1) << 62)) {
    message = "a picture too large";
    return UNROUND_ERROR_ARGUMENT;
  }
  if (picture == nullptr) {
    message = "the samples are a null pointer";
    return UNROUND_ERROR_ARGUMENT;
  }
  unround::JpegImage image;
  const unround_status read = read_file(data, size, settings->settings, image, message);
  if (read != UNROUND_OK) return read;
  if (height != image.height || width != image.width) {
    message = "a picture of the file is " + std::to_string(image.height) + " x " + std::to_string(image.width) +
              ", not " + std::to_string(height) + " x " + std::to_string(width);
    return UNROUND_ERROR_OPTIONS;
  }
  std::vector<unround_component> components;
  const unround_input input = input_of(image, components);
  auto made = std::make_unique<unround_result>();
  unround::Plan plan;
  const unround_status checked = frame_of(&input, settings, *made, plan, message);
  if (checked != UNROUND_OK) return checked;
  const unround::FrameModel& frame = made->frame;
  const size_t expected = image.color_space == UNROUND_COLOR_GRAYSCALE ? 1 : 3;
  const size_t h = image.height, w = image.width, plane = h * w;
  if (channels != expected) {
    message = "a picture of the file is " + std::to_string(h) + " x " + std::to_string(w) + " x " +
              std::to_string(expected) + " samples, not in " + std::to_string(channels) + " channels";
    return UNROUND_ERROR_OPTIONS;
  }
  for (size_t k = 0; k < plane * channels; ++k) {
    if (!std::isfinite(picture[k])) {
      message = "a picture's samples are finite";
      return UNROUND_ERROR_OPTIONS;
    }
  }
  std::vector<double> planes(plane * channels);
  if (image.color_space == UNROUND_COLOR_YCBCR) {
    unround_to_ycbcr(picture, h, w, planes.data());
  } else {
    for (size_t k = 0; k < plane; ++k)
      for (size_t c = 0; c < channels; ++c) planes[c * plane + k] = picture[k * channels + c];
  }
  const size_t rows = frame.rows, columns = frame.columns;
  std::vector<double> canvas(channels * rows * columns);
  for (size_t c = 0; c < channels; ++c) {
    for (size_t i = 0; i < rows; ++i) {
      const double* row = planes.data() + c * plane + std::min(i, h - 1) * w;
      double* out = canvas.data() + (c * rows + i) * columns;
      std::memcpy(out, row, sizeof(double) * w);
      std::fill(out + w, out + columns, row[w - 1]);
    }
  }
  int32_t status = UNROUND_OK;
  std::vector<std::vector<double>> excess;
  if (!unround::project_on_gpu(frame, settings->settings, canvas, made->out, excess, message, status)) {
    return static_cast<unround_status>(status);
  }
  made->solver = false;
  made->excess = std::move(excess);
  assemble(*made);
  made->icc_profile = std::move(image.icc_profile);
  made->exif_orientation = image.exif_orientation;
  made->warning = image.warnings > 0 ? image.warning : "";
  *result = made.release();
  return UNROUND_OK;
}

// body(k) for every k below count, on up to 8 threads, each a contiguous run of k (this
// thread takes the runs of any thread that cannot be had). body must not throw.
template <typename Body>
void for_each_index(size_t count, const Body& body) {
  const size_t workers = std::min<size_t>({8, count, std::max(1u, std::thread::hardware_concurrency())});
  const auto run = [&](size_t t) {
    for (size_t k = count * t / workers; k < count * (t + 1) / workers; ++k) body(k);
  };
  std::vector<std::thread> threads;
  size_t started = 1;
  try {
    for (; started < workers; ++started) threads.emplace_back(run, started);
  } catch (...) {
  }
  if (workers > 0) run(0);
  for (size_t t = started; t < workers; ++t) run(t);
  for (std::thread& thread : threads) thread.join();
}

unround_status reconstruct_batch(const unround_input* inputs, size_t count, const unround_settings* settings,
                                 unround_result** results, double* latencies, std::string& message) {
  if (results == nullptr || (count > 0 && inputs == nullptr) || settings == nullptr) {
    message = "the inputs, the settings and the results are not null pointers";
    return UNROUND_ERROR_ARGUMENT;
  }
  std::vector<std::unique_ptr<unround_result>> made(count);
  std::vector<unround::Plan> plans(count);
  for (size_t k = 0; k < count; ++k) made[k] = std::make_unique<unround_result>();
  // The frames (host work, each input alone) on a few threads; the first refusal reported.
  std::vector<unround_status> checked(count, UNROUND_OK);
  std::vector<std::string> messages(count);
  for_each_index(count, [&](size_t k) {
    try {
      checked[k] = frame_of(&inputs[k], settings, *made[k], plans[k], messages[k]);
    } catch (const std::bad_alloc&) {
      checked[k] = UNROUND_ERROR_INTERNAL;
      messages[k] = "out of host memory";
    } catch (...) {
      checked[k] = UNROUND_ERROR_INTERNAL;
      messages[k] = "the frame could not be made";
    }
  });
  for (size_t k = 0; k < count; ++k) {
    if (checked[k] != UNROUND_OK) {
      message = "input " + std::to_string(k) + ": " + messages[k];
      return checked[k];
    }
  }
  const unround::Settings& s = settings->settings;
  const bool solving = s.method == unround::Method::tv || s.method == unround::Method::tgv;
  if (!solving && count > 0) {
    message = "a batch iterates: use --method tv or tgv";
    return UNROUND_ERROR_UNSUPPORTED;
  }
  std::vector<const unround::FrameModel*> frames;
  std::vector<const unround::Plan*> plan_pointers;
  for (size_t k = 0; k < count; ++k) {
    frames.push_back(&made[k]->frame);
    plan_pointers.push_back(&plans[k]);
  }
  std::vector<unround::SolveOutput> outs;
  std::vector<double> seconds;
  int32_t status = UNROUND_OK;
  if (count > 0 && !unround::solve_batch_on_gpu(frames, plan_pointers, s, outs, seconds, message, status)) {
    return static_cast<unround_status>(status);
  }
  for (size_t k = 0; k < count; ++k) {
    made[k]->out = std::move(outs[k]);
    if (latencies != nullptr) latencies[k] = seconds[k];
  }
  std::vector<char> assembled(count, 1);
  for_each_index(count, [&](size_t k) {
    try {
      assemble(*made[k]);
    } catch (...) {
      assembled[k] = 0;
    }
  });
  for (size_t k = 0; k < count; ++k) {
    if (!assembled[k]) {
      message = "input " + std::to_string(k) + ": out of host memory for its result";
      return UNROUND_ERROR_INTERNAL;
    }
  }
  for (size_t k = 0; k < count; ++k) results[k] = made[k].release();
  return UNROUND_OK;
}

}  // namespace

extern "C" {

int32_t unround_abi_version(void) { return UNROUND_ABI_VERSION; }

const char* unround_version(void) { return kVersion; }

unround_settings* unround_settings_new(size_t count, const char* const* options, char* message, size_t message_size) {
  unround_settings* made = nullptr;
  guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (count > 0 && options == nullptr) {
      text = "the strings are a null pointer";
      return UNROUND_ERROR_ARGUMENT;
    }
    std::vector<std::string> arguments;
    for (size_t k = 0; k < count; ++k) {
      if (options[k] == nullptr) {
        text = "a string is a null pointer";
        return UNROUND_ERROR_ARGUMENT;
      }
      arguments.emplace_back(options[k]);
    }
    unround::Settings settings;
    if (!unround::parse_settings(arguments, settings, text)) return UNROUND_ERROR_OPTIONS;
    made = new unround_settings{std::move(settings)};
    return UNROUND_OK;
  });
  return made;
}

void unround_settings_free(unround_settings* settings) { delete settings; }

size_t unround_settings_options(const unround_settings* settings, char* buffer, size_t size) {
  if (settings == nullptr) return 0;
  try {
    const std::string text = unround::options_text(settings->settings);
    write_message(buffer, size, text);
    return text.size();
  } catch (...) {
    return 0;
  }
}

unround_status unround_decode(const uint8_t* data, size_t size, const unround_settings* settings,
                              unround_observer observer, void* user, unround_result** result, char* message,
                              size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return decode_file(data, size, settings, observer, user, result, text);
  });
}

unround_status unround_solve(const unround_input* input, const unround_settings* settings, unround_observer observer,
                             void* user, unround_result** result, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return reconstruct(input, settings, nullptr, observer, user, result, text);
  });
}

unround_status unround_solve_start(const unround_input* input, const unround_settings* settings,
                                   const unround_start* start, unround_observer observer, void* user,
                                   unround_result** result, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return reconstruct(input, settings, start, observer, user, result, text);
  });
}

unround_status unround_solve_samples(const uint16_t* levels, uint64_t height, uint64_t width, uint64_t channels,
                                     uint32_t bits, const unround_settings* settings, unround_observer observer,
                                     void* user, unround_result** result, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return reconstruct_samples(levels, height, width, channels, bits, settings, observer, user, result, text);
  });
}

unround_status unround_solve_scratch(const unround_input* input, const unround_settings* settings, double* planes,
                                     uint64_t count, uint64_t* iterations, int32_t* stop, char* message,
                                     size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (planes == nullptr) {
      text = "the input, the settings and the planes are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    unround_result made;
    unround::Plan plan;
    const unround_status checked = frame_of(input, settings, made, plan, text);
    if (checked != UNROUND_OK) return checked;
    const unround::FrameModel& frame = made.frame;
    const uint64_t wanted = static_cast<uint64_t>(frame.channels()) * frame.height * frame.width;
    if (count != wanted) {
      text = "the planes hold " + std::to_string(wanted) + " doubles (" + std::to_string(frame.channels()) + " x " +
             std::to_string(frame.height) + " x " + std::to_string(frame.width) + "), not " + std::to_string(count);
      return UNROUND_ERROR_ARGUMENT;
    }
    uint64_t taken = 0;
    int32_t reason = 0, status = UNROUND_OK;
    if (!unround::solve_in_scratch(frame, settings->settings, plan, planes, taken, reason, text, status)) {
      return static_cast<unround_status>(status);
    }
    if (iterations != nullptr) *iterations = taken;
    if (stop != nullptr) *stop = reason;
    return UNROUND_OK;
  });
}

unround_status unround_solve_batch(const unround_input* inputs, size_t count, const unround_settings* settings,
                                   unround_result** results, double* latencies, char* message,
                                   size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return reconstruct_batch(inputs, count, settings, results, latencies, text);
  });
}

unround_status unround_unrolled_new(const unround_input* inputs, size_t count, const unround_settings* settings,
                                    uint32_t iterations, uint32_t checkpoint_every, unround_unrolled** made,
                                    char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (made == nullptr || settings == nullptr || inputs == nullptr || count == 0) {
      text = "the inputs, the settings and the handle are not null pointers, and there is an input";
      return UNROUND_ERROR_ARGUMENT;
    }
    *made = nullptr;
    auto handle = std::make_unique<unround_unrolled>();
    handle->plans.resize(count);
    std::vector<const unround::FrameModel*> frames;
    std::vector<const unround::Plan*> plans;
    for (size_t k = 0; k < count; ++k) {
      handle->frames.push_back(std::make_unique<unround_result>());
      const unround_status checked = frame_of(&inputs[k], settings, *handle->frames[k], handle->plans[k], text);
      if (checked != UNROUND_OK) {
        text = "input " + std::to_string(k) + ": " + text;
        return checked;
      }
      frames.push_back(&handle->frames[k]->frame);
      plans.push_back(&handle->plans[k]);
    }
    int32_t status = UNROUND_OK;
    handle->solve = unround::make_unrolled(frames, plans, settings->settings, static_cast<int>(iterations),
                                           static_cast<int>(checkpoint_every), timing_asked(), text, status);
    if (handle->solve == nullptr) return static_cast<unround_status>(status);
    *made = handle.release();
    return UNROUND_OK;
  });
}

void unround_unrolled_free(unround_unrolled* handle) { delete handle; }

unround_status unround_unrolled_shape(const unround_unrolled* handle, uint64_t* count, uint64_t* height,
                                      uint64_t* width) {
  if (handle == nullptr) return UNROUND_ERROR_ARGUMENT;
  if (count != nullptr) *count = static_cast<uint64_t>(handle->solve->frames());
  if (height != nullptr) *height = static_cast<uint64_t>(handle->solve->height());
  if (width != nullptr) *width = static_cast<uint64_t>(handle->solve->width());
  return UNROUND_OK;
}

unround_status unround_unrolled_constants(const unround_unrolled* handle, double* values) {
  if (handle == nullptr || values == nullptr) return UNROUND_ERROR_ARGUMENT;
  handle->solve->constants(values);
  return UNROUND_OK;
}

unround_status unround_unrolled_start(unround_unrolled* handle, double* state, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (handle == nullptr || state == nullptr) {
      text = "the handle and the state are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    handle->solve->start_state(state);
    return UNROUND_OK;
  });
}

unround_status unround_unrolled_problem(const unround_unrolled* handle, uint64_t frame, double* values) {
  if (handle == nullptr || values == nullptr || frame >= handle->frames.size()) return UNROUND_ERROR_ARGUMENT;
  try {
    const unround::ComponentModel& k = handle->frames[frame]->frame.components[0];
    for (size_t b = 0; b < k.blocks(); ++b) {
      for (int q = 0; q < 64; ++q) {
        double* entry = values + 4 * (64 * b + q);
        double own_lower, own_upper;
        k.coefficient(b, q, entry[0], entry[1], entry[2], own_lower, own_upper);
        entry[3] = k.flat[b] != 0 ? k.flat_weights[q] : k.weights[q];
      }
    }
    return UNROUND_OK;
  } catch (...) {
    return UNROUND_ERROR_INTERNAL;
  }
}

unround_status unround_unrolled_forward(unround_unrolled* handle, const float* weight_map, double alpha0,
                                        double alpha1, float* canvas, void* stream, char* message,
                                        size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (handle == nullptr || weight_map == nullptr || canvas == nullptr) {
      text = "the handle, the map and the canvas are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    if (!(alpha0 > 0.0) || !(alpha1 > 0.0)) {
      text = "alpha0 and alpha1 above 0";
      return UNROUND_ERROR_ARGUMENT;
    }
    handle->solve->forward(weight_map, alpha0, alpha1, canvas, stream);
    return UNROUND_OK;
  });
}

unround_status unround_unrolled_backward(unround_unrolled* handle, const float* grad_canvas, float* grad_map,
                                         double* grad_alphas, void* stream, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (handle == nullptr || grad_canvas == nullptr || grad_map == nullptr || grad_alphas == nullptr) {
      text = "the handle, the gradients and the map's gradient are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    handle->solve->backward(grad_canvas, grad_map, grad_alphas, stream);
    return UNROUND_OK;
  });
}

int32_t unround_unrolled_kernel_times(unround_unrolled* handle, double* values, size_t count) {
  if (handle == nullptr || values == nullptr) return 0;
  double all[6] = {};
  try {
    if (!handle->solve->kernel_times(all)) return 0;
  } catch (const std::exception&) {
    return 0;
  }
  for (size_t k = 0; k < count && k < 6; ++k) values[k] = all[k];
  return 1;
}

unround_status unround_project(const uint8_t* data, size_t size, const unround_settings* settings,
                               const double* picture, uint64_t height, uint64_t width, uint64_t channels,
                               unround_result** result, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) {
    return project_picture(data, size, settings, picture, height, width, channels, result, text);
  });
}

void unround_result_free(unround_result* result) { delete result; }

const double* unround_result_picture(const unround_result* result, uint64_t* height, uint64_t* width,
                                     uint64_t* channels) {
  if (result == nullptr) return nullptr;
  if (height != nullptr) *height = result->frame.height;
  if (width != nullptr) *width = result->frame.width;
  if (channels != nullptr) *channels = result->picture_channels;
  return result->picture.data();
}

const double* unround_result_planes(const unround_result* result, uint64_t* count) {
  if (result == nullptr) return nullptr;
  if (count != nullptr) *count = result->planes.size();
  return result->planes.data();
}

int32_t unround_result_color_space(const unround_result* result) {
  return result == nullptr ? 0 : result->frame.color_space;
}

const double* unround_result_canvas(const unround_result* result, uint64_t* channels, uint64_t* height,
                                    uint64_t* width) {
  if (result == nullptr) return nullptr;
  if (result->format == UNROUND_FORMAT_PNG) {
    // The planes: nothing lies beyond the picture.
    if (channels != nullptr) *channels = result->picture_channels;
    if (height != nullptr) *height = result->frame.height;
    if (width != nullptr) *width = result->frame.width;
    return result->planes.data();
  }
  if (channels != nullptr) *channels = result->frame.channels();
  if (height != nullptr) *height = result->frame.rows;
  if (width != nullptr) *width = result->frame.columns;
  return result->out.canvas.data();
}

uint64_t unround_result_components(const unround_result* result) {
  return result == nullptr ? 0 : result->frame.channels();
}

const double* unround_result_coefficients(const unround_result* result, uint64_t component, uint64_t* count) {
  if (result == nullptr || component >= result->frame.channels()) return nullptr;
  if (count != nullptr) *count = result->out.coefficients[component].size();
  return result->out.coefficients[component].data();
}

const double* unround_result_problem(const unround_result* result, uint64_t component, int32_t which,
                                     uint64_t* count) {
  if (result == nullptr || component >= result->frame.channels() || which < 0 || which > 4) return nullptr;
  try {
    const std::lock_guard<std::mutex> lock(result->made_guard);
    if (!result->made[component]) {
      const unround::ComponentModel& k = result->frame.components[component];
      auto& arrays = result->problems[component];
      const size_t entries = 64 * k.blocks();
      for (int a = 0; a < 3; ++a) arrays[a].resize(entries);
      // Each block's intervals depend on that block alone: runs of blocks on a few threads.
      const auto fill = [&](size_t from, size_t to) {
        for (size_t b = from; b < to; ++b) {
          for (int q = 0; q < 64; ++q) {
            double own_lower, own_upper;
            k.coefficient(b, q, arrays[0][64 * b + q], arrays[1][64 * b + q], arrays[2][64 * b + q], own_lower,
                          own_upper);
          }
        }
      };
      const size_t blocks = k.blocks();
      const size_t workers = blocks < 4096 ? 1 : std::min<size_t>(8, std::max(1u, std::thread::hardware_concurrency()));
      std::vector<std::thread> threads;
      size_t started = 1;
      try {
        for (; started < workers; ++started) {
          threads.emplace_back(fill, blocks * started / workers, blocks * (started + 1) / workers);
        }
      } catch (...) {
        // No more threads to be had: this one takes the runs not started.
      }
      fill(0, blocks / workers);
      for (size_t t = started; t < workers; ++t) fill(blocks * t / workers, blocks * (t + 1) / workers);
      for (std::thread& thread : threads) thread.join();
      arrays[3].assign(k.steps.begin(), k.steps.end());
      arrays[4].assign(k.weights.begin(), k.weights.end());
      result->made[component] = true;
    }
    const std::vector<double>& values = result->problems[component][which];
    if (count != nullptr) *count = values.size();
    return values.data();
  } catch (...) {
    return nullptr;
  }
}

unround_status unround_result_component(const unround_result* result, uint64_t component, uint64_t* rows,
                                        uint64_t* columns, uint64_t* down, uint64_t* across) {
  if (result == nullptr || component >= result->frame.channels()) return UNROUND_ERROR_ARGUMENT;
  const unround::ComponentModel& k = result->frame.components[component];
  if (rows != nullptr) *rows = k.rows;
  if (columns != nullptr) *columns = k.columns;
  if (down != nullptr) *down = k.down;
  if (across != nullptr) *across = k.across;
  return UNROUND_OK;
}

int32_t unround_result_solver(const unround_result* result, uint64_t* iterations, int32_t* stop) {
  if (result == nullptr || !result->solver) return 0;
  if (iterations != nullptr) *iterations = result->out.iterations;
  if (stop != nullptr) *stop = result->out.stop;
  return 1;
}

const uint64_t* unround_result_record_iterations(const unround_result* result, uint64_t* count) {
  if (result == nullptr) return nullptr;
  if (count != nullptr) *count = result->record_iterations.size();
  return result->record_iterations.data();
}

const double* unround_result_record_values(const unround_result* result, int32_t which, uint64_t* count) {
  if (result == nullptr || which < 0 || which > 4) return nullptr;
  if (count != nullptr) *count = result->record_values[which].size();
  return result->record_values[which].data();
}

int32_t unround_result_alpha_solver(const unround_result* result, uint64_t* iterations, int32_t* stop) {
  if (result == nullptr || !result->has_alpha) return 0;
  if (iterations != nullptr) *iterations = result->alpha.iterations;
  if (stop != nullptr) *stop = result->alpha.stop;
  return 1;
}

const uint64_t* unround_result_alpha_record_iterations(const unround_result* result, uint64_t* count) {
  if (result == nullptr || !result->has_alpha) {
    if (count != nullptr) *count = 0;
    return nullptr;
  }
  if (count != nullptr) *count = result->alpha_record_iterations.size();
  return result->alpha_record_iterations.data();
}

const double* unround_result_alpha_record_values(const unround_result* result, int32_t which, uint64_t* count) {
  if (result == nullptr || !result->has_alpha || which < 0 || which > 4) {
    if (count != nullptr) *count = 0;
    return nullptr;
  }
  if (count != nullptr) *count = result->alpha_record_values[which].size();
  return result->alpha_record_values[which].data();
}

const uint8_t* unround_result_icc_profile(const unround_result* result, uint64_t* size) {
  if (result == nullptr || result->icc_profile.empty()) {
    if (size != nullptr) *size = 0;
    return nullptr;
  }
  if (size != nullptr) *size = result->icc_profile.size();
  return result->icc_profile.data();
}

int32_t unround_result_exif_orientation(const unround_result* result) {
  return result == nullptr ? 0 : result->exif_orientation;
}

int32_t unround_result_format(const unround_result* result) { return result == nullptr ? 0 : result->format; }

int32_t unround_result_bits(const unround_result* result) { return result == nullptr ? 0 : result->bits; }

const char* unround_result_warning(const unround_result* result) {
  return result == nullptr ? nullptr : result->warning.c_str();
}

const double* unround_result_excess(const unround_result* result, uint64_t component, uint64_t* count) {
  if (result == nullptr || component >= result->excess.size()) {
    if (count != nullptr) *count = 0;
    return nullptr;
  }
  if (count != nullptr) *count = result->excess[component].size();
  return result->excess[component].data();
}

int32_t unround_result_kernel_times(const unround_result* result, double* values, size_t count) {
  if (result == nullptr || values == nullptr) return 0;
  const unround::KernelTimes& t = result->out.times;
  const double all[7] = {t.primal_seconds,     static_cast<double>(t.primal_launches),
                         t.dual_seconds,       static_cast<double>(t.dual_launches),
                         t.iterations_seconds, static_cast<double>(t.iterations),
                         static_cast<double>(t.dual_rows)};
  for (size_t k = 0; k < count && k < 7; ++k) values[k] = all[k];
  return t.iterations > 0 ? 1 : 0;
}

namespace {

// Bytes handed over to the caller, who frees them with unround_bytes_free.
void hand_over(const std::vector<uint8_t>& made, uint8_t** bytes, uint64_t* size) {
  uint8_t* out = static_cast<uint8_t*>(std::malloc(std::max<size_t>(1, made.size())));
  if (out == nullptr) throw std::bad_alloc();
  if (!made.empty()) std::memcpy(out, made.data(), made.size());
  *bytes = out;
  *size = made.size();
}

// The picture's samples and sizes checked as the reference's C interface does
// (rust/capi/src/lib.rs:1259-1269).
unround_status picture_given(const double* samples, uint64_t height, uint64_t width, uint64_t channels,
                             std::string& message) {
  const unsigned __int128 count = static_cast<unsigned __int128>(height) * width * channels;
  if (count > (static_cast<unsigned __int128>(1) << 62)) {
    message = "a picture too large";
    return UNROUND_ERROR_ARGUMENT;
  }
  if (samples == nullptr) {
    message = "the samples are a null pointer";
    return UNROUND_ERROR_ARGUMENT;
  }
  return UNROUND_OK;
}

unround::Dither dither_of(int32_t dither) { return static_cast<unround::Dither>(dither); }

bool dither_known(int32_t dither, std::string& message) {
  if (dither >= 0 && dither <= 2) return true;
  message = "the dither is 0, 1 or 2, not " + std::to_string(dither);
  return false;
}

}  // namespace

unround_status unround_tiff(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t ycbcr,
                            const uint8_t* icc, uint64_t icc_size, uint8_t** bytes, uint64_t* size, char* message,
                            size_t message_size) {
  std::string warning;
  const unround_status status = guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (bytes == nullptr || size == nullptr) {
      text = "the bytes and the size are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    const unround_status given = picture_given(samples, height, width, channels, text);
    if (given != UNROUND_OK) return given;
    std::vector<uint8_t> file;
    if (!unround::tiff_bytes(samples, height, width, channels, ycbcr != 0, icc, icc != nullptr ? icc_size : 0, file,
                             warning, text)) {
      return UNROUND_ERROR_OPTIONS;
    }
    hand_over(file, bytes, size);
    return UNROUND_OK;
  });
  if (status == UNROUND_OK) write_message(message, message_size, warning);
  return status;
}

unround_status unround_png(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t sixteen,
                           int32_t dither, int32_t compression, const uint8_t* icc, uint64_t icc_size, uint8_t** bytes,
                           uint64_t* size, char* message, size_t message_size) {
  std::string warning;
  const unround_status status = guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (bytes == nullptr || size == nullptr) {
      text = "the bytes and the size are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    if (compression < -1 || compression > 9) {
      text = "zlib's level is 0 to 9, or -1, not " + std::to_string(compression);
      return UNROUND_ERROR_OPTIONS;
    }
    if (!dither_known(dither, text)) return UNROUND_ERROR_OPTIONS;
    const unround_status given = picture_given(samples, height, width, channels, text);
    if (given != UNROUND_OK) return given;
    if (height > 0x7FFFFFFFu || width > 0x7FFFFFFFu) {
      text = "a side of the picture is larger than PNG allows";
      return UNROUND_ERROR_OPTIONS;
    }
    std::vector<uint16_t> levels;
    if (!unround::quantize(samples, height, width, channels, sixteen != 0, dither_of(dither), levels, text)) {
      return UNROUND_ERROR_OPTIONS;
    }
    std::vector<uint8_t> eight;
    const void* data = levels.data();
    if (sixteen == 0) {
      eight.assign(levels.begin(), levels.end());
      data = eight.data();
    }
    std::vector<uint8_t> file;
    int32_t layer = 0;
    if (!unround::png_bytes(data, static_cast<uint32_t>(width), static_cast<uint32_t>(height),
                            static_cast<int32_t>(channels), sixteen != 0 ? 16 : 8, compression, icc,
                            icc != nullptr ? icc_size : 0, file, warning, text, layer)) {
      return layer == 1 ? UNROUND_ERROR_OPTIONS : UNROUND_ERROR_IO;
    }
    hand_over(file, bytes, size);
    return UNROUND_OK;
  });
  if (status == UNROUND_OK) write_message(message, message_size, warning);
  return status;
}

unround_status unround_pnm(const double* samples, uint64_t height, uint64_t width, uint64_t channels, int32_t sixteen,
                           int32_t dither, uint8_t** bytes, uint64_t* size, char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    if (bytes == nullptr || size == nullptr) {
      text = "the bytes and the size are not null pointers";
      return UNROUND_ERROR_ARGUMENT;
    }
    if (!dither_known(dither, text)) return UNROUND_ERROR_OPTIONS;
    const unround_status given = picture_given(samples, height, width, channels, text);
    if (given != UNROUND_OK) return given;
    std::vector<uint8_t> file;
    if (!unround::pnm_bytes(samples, height, width, channels, sixteen != 0, dither_of(dither), file, text)) {
      return UNROUND_ERROR_OPTIONS;
    }
    hand_over(file, bytes, size);
    return UNROUND_OK;
  });
}

void unround_bytes_free(uint8_t* bytes, uint64_t) { std::free(bytes); }

unround_status unround_orient(const double* samples, uint64_t height, uint64_t width, uint64_t channels,
                              int32_t orientation, double* turned, uint64_t* turned_height, uint64_t* turned_width,
                              char* message, size_t message_size) {
  return guarded(message, message_size, [&](std::string& text) -> unround_status {
    const unround_status given = picture_given(samples, height, width, channels, text);
    if (given != UNROUND_OK) return given;
    if (turned == nullptr) {
      text = "the turned picture is a null pointer";
      return UNROUND_ERROR_ARGUMENT;
    }
    size_t rows = 0, columns = 0;
    if (!unround::orient(samples, height, width, channels, orientation, turned, rows, columns, text)) {
      return UNROUND_ERROR_OPTIONS;
    }
    if (turned_height != nullptr) *turned_height = rows;
    if (turned_width != nullptr) *turned_width = columns;
    return UNROUND_OK;
  });
}

// The samples of one plane of height x width, where three planes of them fit in memory's
// addresses as binary64 (rust/capi/src/lib.rs:1568-1582 refuses what does not fit).
static bool plane_samples(uint64_t height, uint64_t width, size_t& count) {
  uint64_t samples = 0, bytes = 0;
  if (__builtin_mul_overflow(height, width, &samples)) return false;
  if (__builtin_mul_overflow(samples, uint64_t{3 * sizeof(double)}, &bytes)) return false;
  if (bytes > static_cast<uint64_t>(PTRDIFF_MAX)) return false;
  count = static_cast<size_t>(samples);
  return true;
}

unround_status unround_to_rgb(const double* planes, uint64_t height, uint64_t width, double* picture) {
  size_t count = 0;
  if (planes == nullptr || picture == nullptr || !plane_samples(height, width, count)) return UNROUND_ERROR_ARGUMENT;
  rgb_of(planes, count, picture);
  return UNROUND_OK;
}

unround_status unround_to_ycbcr(const double* picture, uint64_t height, uint64_t width, double* planes) {
  size_t count = 0;
  if (planes == nullptr || picture == nullptr || !plane_samples(height, width, count)) return UNROUND_ERROR_ARGUMENT;
  // Y = (kR R + kG G) + kB B, Cb = 128 + (250/443)(B - Y), Cr = 128 + (500/701)(R - Y)
  // (docs/math.md:1035-1040).
  const double kCbInverse = 250.0 / 443.0, kCrInverse = 500.0 / 701.0;
  for (size_t k = 0; k < count; ++k) {
    const double r = picture[3 * k], g = picture[3 * k + 1], b = picture[3 * k + 2];
    const double y = (0.299 * r + 0.587 * g) + 0.114 * b;
    planes[k] = y;
    planes[count + k] = 128.0 + kCbInverse * (b - y);
    planes[2 * count + k] = 128.0 + kCrInverse * (r - y);
  }
  return UNROUND_OK;
}

int32_t unround_main(size_t, const char* const*) { return 2; }

}  // extern "C"
