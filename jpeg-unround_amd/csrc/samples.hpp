// A picture of rounded samples on the GPU (docs/math.md 11; rust/unround/src/samples.rs:
// 656-1104): TV or TGV within the intervals of the samples' levels. The iteration is the
// JPEG frame's with the proximal step of the data term taken sample by sample, without a
// transform: x~ = clip((v + t c) r, a, b), t = tau mu, r = 1 / (1 + t), v = x + tau gamma
// div p. The dual kernel, the norms of the records and the run loop are the frame's
// (Solver); this file holds the data term's kernels and SamplesSolver, which plugs them
// into Solver's hooks. (Included by solver.hip, inside its anonymous namespace.)
//
// Layout: the levels as uint16 [C][H][W] (2 bytes a sample); a table of every level's low
// end, high end and centre in the iterate's type, 4 entries a level (the fourth 0), which
// the proximal step gathers from (256 levels of 8 bits: 4 KB, cached); and the same table
// in binary64, exact, for the records and the final projection.

// The data term on the device.
template <typename T>
struct SampleTerm {
  const uint16_t* levels;  // [C][H][W]
  const T* ends;           // [top + 1][4]: low, high, centre, 0
  const double* exact;     // [top + 1][4], binary64
};

// x and x~ at the start: the centres, which lie within their intervals
// (samples.rs:903-908).
template <typename T, typename S>
__global__ __launch_bounds__(256) void samples_start_kernel(Fields<T, S> F, SampleTerm<T> term, size_t count) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double* entry = term.exact + 4 * static_cast<size_t>(term.levels[k]);
  const T start = static_cast<T>(device::clamp_to(entry[2], entry[0], entry[1]));
  F.x[k] = start;
  F.tx[k] = start;
}

// The proximal step of every channel (samples.rs:960-966, 171-182): v = x + tau gamma_c
// div p (backward differences), x~ = clip((t c + v) r, a, b), then x- = 2 x~ - x and x
// relaxed in place, and x~ kept where a record follows. A work-group takes 64 columns and
// 4 rows, a lane a pixel and its C channels; the neighbours of div p on the left and above
// are loads of the lanes beside it and of the row above, which the caches serve. Every
// load of every channel is issued before the first store, so that they are in flight
// together.
template <typename T, typename S, int C>
__global__ __launch_bounds__(256) void samples_prox_kernel(Fields<T, S> F, SampleTerm<T> term, int H, int W,
                                                           Steps<T> steps, T t, T r, int keep) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const S* __restrict__ p = F.p;
  T* __restrict__ x = F.x;
  const uint16_t* __restrict__ levels = term.levels;
  // p as stored until it is taken (a BF16 sample widened as it arrives makes the compiler
  // wait for its load before issuing the next).
  S across[C], left[C], down[C], up[C];
  T old[C];
  uint16_t level[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const S* px = p + (2 * c) * N;
    const S* py = p + (2 * c + 1) * N;
    across[c] = px[at];
    left[c] = px[j > 0 ? at - 1 : at];
    down[c] = py[at];
    up[c] = py[i > 0 ? at - W : at];
    old[c] = x[c * N + at];
    level[c] = levels[c * N + at];
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    // backward_across_of and backward_down_of leave out the left and upper samples at the
    // canvas' edges.
    T div = backward_down_of(static_cast<T>(device::widen(down[c])), static_cast<T>(device::widen(up[c])), i, H);
    div += backward_across_of(static_cast<T>(device::widen(across[c])), static_cast<T>(device::widen(left[c])), j, W);
    const T v = old[c] + steps.descent[c] * div;
    const T* entry = term.ends + 4 * static_cast<size_t>(level[c]);
    const T tilde = device::clamp_to(fma(t, entry[2], v) * r, entry[0], entry[1]);
    const size_t k = c * N + at;
    F.xbar[k] = T(2) * tilde - old[c];
    x[k] = steps.rho * tilde + steps.rest * old[c];
    if (keep) F.tx[k] = tilde;
  }
}

// G's sum at x~ before its factor mu / 2: the squares of the distances from the centres
// (samples.rs:202-220), a work-group's sum into partials (64 columns x 4 rows, every
// channel).
template <typename T, typename S>
__global__ __launch_bounds__(256) void samples_value_kernel(Fields<T, S> F, SampleTerm<T> term, int C, int H, int W,
                                                            double* __restrict__ partials) {
  __shared__ double room[256];
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  double sum = 0.0;
  if (i < H && j < W) {
    const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
    for (int c = 0; c < C; ++c) {
      const size_t k = c * N + at;
      const double away = static_cast<double>(F.tx[k]) - term.exact[4 * static_cast<size_t>(term.levels[k]) + 2];
      sum += away * away;
    }
  }
  const double total = group_reduce<false>(sum, room);
  if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = total;
}

// g*(s) of a sample (docs/math.md 11): sup over z in [a, b] of s z - mu/2 (z - c)^2.
__device__ __forceinline__ double sample_conjugate(double s, double mu, double low, double high, double centre) {
  if (mu > 0.0) {
    const double best = device::clamp_to(centre + s / mu, low, high);
    const double difference = best - centre;
    return s * best - 0.5 * mu * (difference * difference);
  }
  return fmax(s * low, s * high);
}

// G* at xi = gamma_c div p~ (from_r = 0) or gamma_c div(-theta div_2 r~) (from_r = 1),
// theta from the largest norm of div_2 r~ at *most (samples.rs:222-248, 837-875): a
// work-group's sum into partials.
template <typename T, typename S>
__global__ __launch_bounds__(256) void samples_conjugate_kernel(Fields<T, S> F, SampleTerm<T> term, int C, int H,
                                                                int W, double g0, double g1, double g2, double mu,
                                                                int from_r, const double* __restrict__ most,
                                                                double alpha1, double* __restrict__ partials) {
  __shared__ double room[256];
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  const double gammas[3] = {g0, g1, g2};
  double theta = 1.0;
  if (from_r) theta = *most > alpha1 ? alpha1 / *most : 1.0;
  double sum = 0.0;
  if (i < H && j < W) {
    const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
    for (int c = 0; c < C; ++c) {
      const double xi = gammas[c] * dual_divergence_at(F, c, i, j, H, W, from_r, theta);
      const double* entry = term.exact + 4 * static_cast<size_t>(term.levels[c * N + at]);
      sum += sample_conjugate(xi, mu, entry[0], entry[1], entry[2]);
    }
  }
  const double total = group_reduce<false>(sum, room);
  if (threadIdx.x == 0) partials[blockIdx.y * gridDim.x + blockIdx.x] = total;
}

// The picture: the last x~ in binary64 clipped to the exact intervals, so that every
// sample rounds to its level.
template <typename T>
__global__ __launch_bounds__(256) void samples_final_kernel(const T* __restrict__ tilde, const double* __restrict__ exact,
                                                            const uint16_t* __restrict__ levels, size_t count,
                                                            double* __restrict__ picture) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double* entry = exact + 4 * static_cast<size_t>(levels[k]);
  picture[k] = device::clamp_to(static_cast<double>(tilde[k]), entry[0], entry[1]);
}

// The canvas's shape of a picture of samples, as Solver reads it: its channels without
// blocks, one sample of each channel to a pixel.
struct SamplesShape {
  FrameModel shape;
  explicit SamplesShape(const SamplesModel& model) {
    shape.height = shape.rows = model.height;
    shape.width = shape.columns = model.width;
    shape.color_space = model.channels == 3 ? 3 : 1;
    shape.components.resize(model.channels);
  }
};

template <typename T, typename S>
class SamplesSolver : private SamplesShape, public Solver<T, S> {
  using Base = Solver<T, S>;

 public:
  SamplesSolver(const SamplesModel& model, const Settings& settings, const Plan& plan, Timing timing)
      : SamplesShape(model), Base(SamplesShape::shape, settings, plan, timing), model_(model) {}

 protected:
  dim3 pixels() const { return dim3((this->W_ + 63) / 64, (this->H_ + 3) / 4); }
  size_t pixel_groups() const { return static_cast<size_t>((this->W_ + 63) / 64) * ((this->H_ + 3) / 4); }

  void data_sizes(size_t& zeta, size_t& value_groups, size_t& conjugate_groups) const override {
    zeta = 0;
    value_groups = conjugate_groups = pixel_groups();
  }

  void start_primal(const StartState*) override {
    hipStream_t s = this->stream();
    const size_t count = model_.count(), levels = model_.low.size();
    std::vector<double> exact(4 * levels, 0.0);
    std::vector<T> ends(4 * levels, T(0));
    for (size_t v = 0; v < levels; ++v) {
      const double entry[3] = {model_.low[v], model_.high[v], model_.centre[v]};
      for (int e = 0; e < 3; ++e) {
        exact[4 * v + e] = entry[e];
        ends[4 * v + e] = static_cast<T>(entry[e]);
      }
    }
    levels_ = DeviceBuffer(sizeof(uint16_t) * count);
    exact_ = DeviceBuffer(sizeof(double) * exact.size());
    ends_ = DeviceBuffer(sizeof(T) * ends.size());
    HIP_CHECK(hipMemcpyAsync(levels_.pointer, model_.levels.data(), sizeof(uint16_t) * count, hipMemcpyHostToDevice,
                             s));
    HIP_CHECK(hipMemcpyAsync(exact_.pointer, exact.data(), sizeof(double) * exact.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(ends_.pointer, ends.data(), sizeof(T) * ends.size(), hipMemcpyHostToDevice, s));
    term_ = SampleTerm<T>{levels_.as<uint16_t>(), ends_.as<T>(), exact_.as<double>()};
    samples_start_kernel<T, S><<<static_cast<unsigned>((count + 255) / 256), 256, 0, s>>>(this->fields_, term_, count);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(s));
  }

  // The fused sweep's proximal step is a JPEG frame's.
  bool fusable() const override { return false; }

  void primal_step(bool keep) override {
    hipStream_t s = this->stream();
    const bool per_kernel = this->timing_kernels();
    if (per_kernel) {
      this->primal_events_.emplace_back();
      HIP_CHECK(hipEventRecord(this->primal_events_.back().begin, s));
    }
    Steps<T> steps{};
    for (int c = 0; c < this->C_; ++c) steps.descent[c] = static_cast<T>(this->plan_.tau * this->plan_.gamma[c]);
    steps.rho = static_cast<T>(this->plan_.relaxation);
    steps.rest = static_cast<T>(1.0 - this->plan_.relaxation);
    const double t = this->plan_.tau * model_.mu;
    const auto prox = [&](auto kernel) {
      kernel<<<pixels(), 256, 0, s>>>(this->fields_, term_, this->H_, this->W_, steps, static_cast<T>(t),
                                      static_cast<T>(1.0 / (1.0 + t)), keep ? 1 : 0);
    };
    this->C_ == 1 ? prox(samples_prox_kernel<T, S, 1>) : prox(samples_prox_kernel<T, S, 3>);
    HIP_CHECK(hipGetLastError());
    if (per_kernel) HIP_CHECK(hipEventRecord(this->primal_events_.back().end, s));
  }

  void data_records(double* value_partials, double* conjugate_partials, double* partial_partials,
                    double* results) override {
    hipStream_t s = this->stream();
    const int C = this->C_, H = this->H_, W = this->W_;
    const int groups = static_cast<int>(pixel_groups());
    samples_value_kernel<T, S><<<pixels(), 256, 0, s>>>(this->fields_, term_, C, H, W, value_partials);
    HIP_CHECK(hipGetLastError());
    reduce_kernel<<<1, 256, 0, s>>>(value_partials, groups, 1, 0u, results + 4);
    HIP_CHECK(hipGetLastError());
    const std::vector<double>& gamma = this->plan_.gamma;
    const double g[3] = {gamma[0], C > 1 ? gamma[1] : 0.0, C > 2 ? gamma[2] : 0.0};
    const auto conjugate = [&](int from_r, double* into, double* result) {
      samples_conjugate_kernel<T, S><<<pixels(), 256, 0, s>>>(this->fields_, term_, C, H, W, g[0], g[1], g[2],
                                                              model_.mu, from_r, results + 2, this->settings_.alpha1,
                                                              into);
      HIP_CHECK(hipGetLastError());
      reduce_kernel<<<1, 256, 0, s>>>(into, groups, 1, 0u, result);
      HIP_CHECK(hipGetLastError());
    };
    conjugate(this->tgv_ ? 1 : 0, conjugate_partials, results + 5);
    if (this->tgv_ && this->plan_.has_partial) conjugate(0, partial_partials, results + 6);
  }

  double data_value(double sum) const override { return model_.mu == 0.0 ? 0.0 : 0.5 * model_.mu * sum; }

  void finish(SolveOutput& out) override {
    hipStream_t s = this->stream();
    const size_t count = model_.count();
    DeviceBuffer picture(sizeof(double) * count);
    samples_final_kernel<T><<<static_cast<unsigned>((count + 255) / 256), 256, 0, s>>>(
        this->fields_.tx, exact_.as<double>(), levels_.as<uint16_t>(), count, picture.as<double>());
    HIP_CHECK(hipGetLastError());
    out.canvas.resize(count);
    out.coefficients.clear();
    HIP_CHECK(hipMemcpyAsync(out.canvas.data(), picture.pointer, sizeof(double) * count, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }

 private:
  const SamplesModel& model_;
  DeviceBuffer levels_, exact_, ends_;
  SampleTerm<T> term_{};
};
