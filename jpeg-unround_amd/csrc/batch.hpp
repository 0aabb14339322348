// Many frames solved together (included by solver.hip, inside its anonymous namespace,
// after Solver).
//
// Each frame keeps its own fields, components and record buffers: a Solver prepares it
// (the start) and finishes it (the final projection). In between, one launch per kind of
// tile runs a step of every frame still iterating: a table lists the tiles of every frame
// (frame, component, work-group), and a tile whose frame has stopped returns at once. The
// records, the stopping test and the count of iterations run on the device too, so that
// a run of iterations up to the next record is one HIP graph, captured once and replayed;
// the host only reads which frames still iterate after each run.
//
// Every tile computes exactly what the frame's own solve computes (the same tile
// functions, the same reductions in the same order, the records' arithmetic unfused), so
// that each frame's iterates, records and stopping iteration are its own solve's.

struct BatchTile {
  int image;
  int part;   // the component (prox, value, conjugate); 0 elsewhere
  int group;  // the work-group within the frame's own launch
};

// A work-group of the batch's dual step: strip group `group` of picture `image` from row
// `first`, `rows` rows; image < 0 for a place that holds no work.
struct DualTile {
  int image;
  int first;
  int rows;
  int group;
};

// What a step of one frame reads, for one parity of the ping-pong buffers.
template <typename T, typename S>
struct ImageSlot {
  Fields<T, S> F;
  const Component* components;
  int H, W;
  int across;  // the dual step's strip groups across
  Steps<T> steps;
  T descent[3], step[3];  // the prox step's descent and coefficient step per component
};

// What a record of one frame reads and where it goes.
template <typename T, typename S>
struct RecordSlot {
  Fields<T, S> F;
  const Component* components;
  int C, H, W, coupled, tgv, has_partial;
  double gamma[3];
  double* norms;  // 4 per pixel group
  int norm_groups, norm_across;
  double* values;
  int value_groups, value_offset[3];
  double* conjugates;
  double* partials;
  int conjugate_groups, conjugate_offset[3];
  double* results;  // first, second, most, residual, data, conjugate, partial conjugate
  double alpha, alpha0, alpha1, free_radius, partial_radius;
  double tolerance, relative_tolerance, partial_tolerance, samples;
  double* log;  // per record: iteration, primal, dual, theta, partial
  int log_capacity;
};

// The device's state of the batch: per frame whether it iterates, its stop (1 converged),
// the iteration it stopped at and its count of records; and the iterations done.
struct BatchState {
  int* active;
  int* stop;
  unsigned long long* stopped_at;
  int* records;
  unsigned long long* counter;
};

// The wavefronts per SIMD the batch's prox is compiled for: a frame's, but 5 for binary32
// tiles of one sample a cell (7 at 69 VGPRs, where 8 at 62 VGPRs measured 11 % slower on
// the batch's short ragged launches: 322 against 358 us on the 64 pictures).
template <typename T, int DV, int DH>
constexpr int kProxBatchWaves = sizeof(T) == 4 && DV * DH == 1 ? 5 : kProxWaves<T, DV, DH>;

// PLAIN: every component of the batch has no slack cost and midpoint centres.
template <typename T, typename S, int DV, int DH, template <typename> class Planes, bool PLAIN = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kProxBatchWaves<T, DV, DH>))) void
prox_batch_kernel(const BatchTile* __restrict__ tiles, const ImageSlot<T, S>* __restrict__ slots,
                  const int* __restrict__ active, T rho, T rest, int keep) {
  const BatchTile t = tiles[device::ordered(blockIdx.x, gridDim.x)];
  if (active[t.image] == 0) return;
  const ImageSlot<T, S>& m = slots[t.image];
  // zeta's pointer read from the slot at its index: an index into a copy of the fields
  // would place the copy in private memory, written by every lane.
  prox_tile<T, S, DV, DH, Planes, PLAIN>(m.components + t.part, m.F, t.part, m.F.zeta[t.part], m.F.x, m.H, m.W,
                                         m.descent[t.part], m.step[t.part], rho, rest, keep, t.group);
}

// The table is laid out in the order the hardware hands out work-groups (each XCD's share
// its own run of places, long runs of rows first), so the tiles are taken as they stand.
// KEEPS false: for the iterations that keep no tilde point (keep 0).
template <typename T, typename S, int C, bool TGV, bool COUPLED, template <typename> class Planes, bool KEEPS = true>
__global__ __launch_bounds__(256) void dual_batch_kernel(const DualTile* __restrict__ tiles,
                                                         const ImageSlot<T, S>* __restrict__ slots,
                                                         const int* __restrict__ active, int keep) {
  const DualTile t = tiles[blockIdx.x];
  if (t.image < 0 || active[t.image] == 0) return;
  const ImageSlot<T, S>& m = slots[t.image];
  dual_run<T, S, C, TGV, COUPLED, Planes, false, KEEPS>(m.F, m.H, m.W, t.first, min(t.first + t.rows, m.H), t.group,
                                                        m.steps, keep);
}

template <typename T, typename S, bool TGV>
__global__ __launch_bounds__(256) void norms_batch_kernel(const BatchTile* __restrict__ tiles,
                                                          const RecordSlot<T, S>* __restrict__ slots,
                                                          const int* __restrict__ active) {
  const BatchTile t = tiles[blockIdx.x];
  if (active[t.image] == 0) return;
  const RecordSlot<T, S>& m = slots[t.image];
  norms_tile<T, S, TGV>(m.F, m.C, m.H, m.W, m.gamma[0], m.gamma[1], m.gamma[2], m.coupled, m.has_partial, m.norms,
                        m.norm_across, t.group);
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void value_batch_kernel(const BatchTile* __restrict__ tiles,
                                                          const RecordSlot<T, S>* __restrict__ slots,
                                                          const int* __restrict__ active) {
  const BatchTile t = tiles[blockIdx.x];
  if (active[t.image] == 0) return;
  const RecordSlot<T, S>& m = slots[t.image];
  value_tile<T>(m.components[t.part], m.F.zeta[t.part], m.values + m.value_offset[t.part], t.group);
}

// from_r as in conjugate_kernel: 1 into the conjugates (TGV's xi), 0 into `partials` when
// `partial` (the partial gap's) or into the conjugates (TV's). The batch's cells are 1 or 2
// samples each way: the fixed shapes' instance.
template <typename T, typename S>
__global__ __launch_bounds__(256) void conjugate_batch_kernel(const BatchTile* __restrict__ tiles,
                                                              const RecordSlot<T, S>* __restrict__ slots,
                                                              const int* __restrict__ active, int from_r,
                                                              int partial) {
  const BatchTile t = tiles[blockIdx.x];
  if (active[t.image] == 0) return;
  const RecordSlot<T, S>& m = slots[t.image];
  double* into = (partial ? m.partials : m.conjugates) + m.conjugate_offset[t.part];
  conjugate_tile<T, S, false>(m.components[t.part], m.F, t.part, m.H, m.W, m.gamma[t.part], from_r, m.results + 2,
                              m.alpha1, m.free_radius, into, t.group);
}

struct ReduceJob {
  int image;
  int count, columns, column;
  unsigned largest_mask;
  const double* partials;
  double* out;
};

__global__ __launch_bounds__(256) void reduce_batch_kernel(const ReduceJob* __restrict__ jobs,
                                                           const int* __restrict__ active) {
  const ReduceJob job = jobs[blockIdx.x];
  if (active[job.image] == 0) return;
  reduce_tile(job.partials, job.count, job.columns, job.largest_mask, job.out, job.column);
}

// A frame's record from its reduced sums, and its stopping test, as the host makes them
// for one frame (primal_dual.hpp:446-501, 984-998): one thread per frame. `length` is the
// iterations since the count was last advanced.
template <typename T, typename S>
__global__ __launch_bounds__(256) void stop_kernel(const RecordSlot<T, S>* __restrict__ slots, int images,
                                                   BatchState state, int length) {
#pragma clang fp contract(off)
  const int image = blockIdx.x * blockDim.x + threadIdx.x;
  if (image >= images || state.active[image] == 0) return;
  const RecordSlot<T, S>& m = slots[image];
  const double* h = m.results;
  const unsigned long long iteration = *state.counter + static_cast<unsigned long long>(length);
  const double data = h[4];
  double primal, theta, partial;
  if (!m.tgv) {
    primal = m.alpha * h[0] + data;
    theta = __builtin_nan("");
    partial = __builtin_nan("");
  } else {
    primal = (m.alpha1 * h[0] + m.alpha0 * h[1]) + data;
    const double most = h[2];
    theta = most > m.alpha1 ? m.alpha1 / most : 1.0;
    partial = m.has_partial ? primal + h[6] + m.partial_radius * h[3] : __builtin_nan("");
  }
  const double dual = -h[5];
  const double gap = primal - dual;
  bool met = false;
  if (m.tolerance > 0.0 && gap / m.samples <= m.tolerance) met = true;
  if (m.relative_tolerance > 0.0 && gap <= m.relative_tolerance * primal) met = true;
  if (m.partial_tolerance > 0.0 && m.has_partial && partial / m.samples <= m.partial_tolerance) met = true;
  const int at = state.records[image];
  if (at < m.log_capacity) {
    double* entry = m.log + 5 * static_cast<size_t>(at);
    entry[0] = static_cast<double>(iteration);
    entry[1] = primal;
    entry[2] = dual;
    entry[3] = theta;
    entry[4] = partial;
    state.records[image] = at + 1;
  }
  if (met) {
    state.active[image] = 0;
    state.stop[image] = 1;
    state.stopped_at[image] = iteration;
  }
}

__global__ void advance_kernel(unsigned long long* counter, int length) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter += static_cast<unsigned long long>(length);
}

// An event that orders streams and wakes the host, with no timing.
struct Mark {
  hipEvent_t event = nullptr;
  Mark() { HIP_CHECK(hipEventCreateWithFlags(&event, hipEventDisableTiming)); }
  ~Mark() {
    if (event != nullptr) (void)hipEventDestroy(event);
  }
  Mark(const Mark&) = delete;
  Mark& operator=(const Mark&) = delete;
};

// The dual step's rows per wavefront for a batch: the best of a few measured on the batch
// itself (its tiles change with the rows), kept per shape of batch for the process.
struct RowsChoice {
  std::mutex lock;
  std::map<std::vector<int>, int> chosen;
};
inline RowsChoice& rows_choices() {
  static RowsChoice choices;
  return choices;
}

template <typename T, typename S>
class Batch {
 public:
  Batch(const std::vector<const FrameModel*>& frames, const std::vector<const Plan*>& plans, const Settings& settings)
      : frames_(frames), plans_(plans), settings_(settings) {}

  ~Batch() {
    for (auto& entry : graphs_) (void)hipGraphExecDestroy(entry.second);
  }

  void run(std::vector<SolveOutput>& outs, std::vector<double>& latencies) {
    hipStream_t s = stream_.stream;
    const int M = static_cast<int>(frames_.size());
    const bool solving = settings_.method == Method::tv || settings_.method == Method::tgv;
    tgv_ = settings_.method == Method::tgv;
    plain_ = true;
    for (const auto& frame : frames_) {
      for (const ComponentModel& k : frame->components) plain_ = plain_ && !k.costed && !k.mmse;
    }
    using Clock = std::chrono::steady_clock;
    const auto began = Clock::now();
    for (int m = 0; m < M; ++m) {
      solvers_.push_back(std::make_unique<Solver<T, S>>(*frames_[m], settings_, *plans_[m], Timing::none, s));
      solvers_.back()->prepare(nullptr);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    const auto prepared = Clock::now();
    auto tabled = prepared, iterated = prepared;
    outs.assign(M, {});
    latencies.assign(M, 0.0);
    most_ = plans_[0]->iterations;
    every_ = plans_[0]->record_every;
    for (const Plan* plan : plans_) {
      if (plan->iterations != most_ || plan->record_every != every_) throw GpuError("frames of one batch share their plan's iterations");
    }
    if (solving && M > 0) {
      build_state();
      build_tables(choose_rows());
      tabled = Clock::now();
      iterate_all();
      iterated = Clock::now();
    }
    for (int m = 0; m < M; ++m) {
      Solver<T, S>& solver = *solvers_[m];
      SolveOutput& out = outs[m];
      out.tau = plans_[m]->tau;
      out.sigma = plans_[m]->sigma;
      if (solving) {
        out.iterations = stop_[m] == 1 ? stopped_at_[m] : reached_;
        out.stop = stop_[m] == 1 ? 1 : 2;
        for (int r = 0; r < record_counts_[m]; ++r) {
          const double* entry = logs_host_[m].data() + 5 * r;
          Record record;
          record.iteration = static_cast<uint64_t>(entry[0]);
          record.primal = entry[1];
          record.dual = entry[2];
          record.theta = entry[3];
          record.partial = entry[4];
          record.seconds = seconds_at(record.iteration);
          out.records.push_back(record);
        }
        latencies[m] = seconds_at(out.iterations);
      }
      solver.finish(out);
    }
    if (std::getenv("UNROUND_BATCH_LOG") != nullptr) {
      const auto seconds = [](Clock::time_point a, Clock::time_point b) {
        return std::chrono::duration<double>(b - a).count();
      };
      std::fprintf(stderr, "batch of %d: prepare %.3f s, tables %.3f s, iterations %.3f s, finish %.3f s\n", M,
                   seconds(began, prepared), seconds(prepared, tabled), seconds(tabled, iterated),
                   seconds(iterated, Clock::now()));
    }
  }

 private:
  using Key = std::vector<int>;

  // The seconds from the batch's start to the end of the run of iterations that reached
  // `iteration`.
  double seconds_at(uint64_t iteration) const {
    const auto found = ends_.lower_bound(iteration);
    return found == ends_.end() ? (ends_.empty() ? 0.0 : ends_.rbegin()->second) : found->second;
  }

  void build_state() {
    const int M = static_cast<int>(solvers_.size());
    state_memory_ = DeviceBuffer(sizeof(int) * 3 * M + sizeof(unsigned long long) * (M + 1) + 64);
    char* at = state_memory_.as<char>();
    state_.counter = reinterpret_cast<unsigned long long*>(at);
    state_.stopped_at = state_.counter + 1;
    state_.active = reinterpret_cast<int*>(state_.stopped_at + M);
    state_.stop = state_.active + M;
    state_.records = state_.stop + M;
    std::vector<int> ones(M, 1), zeros(M, 0);
    std::vector<unsigned long long> counts(M + 1, 0);
    hipStream_t s = stream_.stream;
    HIP_CHECK(hipMemcpyAsync(state_.counter, counts.data(), sizeof(unsigned long long) * (M + 1),
                             hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(state_.active, ones.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(state_.stop, zeros.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(state_.records, zeros.data(), sizeof(int) * M, hipMemcpyHostToDevice, s));
    // The record slots and logs.
    const uint64_t capacity = (every_ > 0 ? most_ / every_ : 0) + 2;
    std::vector<RecordSlot<T, S>> records(M);
    logs_.clear();
    for (int m = 0; m < M; ++m) {
      Solver<T, S>& v = *solvers_[m];
      const Plan& plan = *plans_[m];
      RecordSlot<T, S>& r = records[m];
      r.F = v.fields_;
      r.components = v.component(0);
      r.C = v.C_;
      r.H = v.H_;
      r.W = v.W_;
      r.coupled = settings_.coupled ? 1 : 0;
      r.tgv = tgv_ ? 1 : 0;
      r.has_partial = plan.has_partial ? 1 : 0;
      for (int c = 0; c < 3; ++c) r.gamma[c] = c < v.C_ ? plan.gamma[c] : 0.0;
      double* partials = v.partials_.template as<double>();
      r.norms = partials;
      r.norm_groups = static_cast<int>(v.norm_partials_count_);
      r.norm_across = (v.W_ + 63) / 64;
      r.values = r.norms + 4 * v.norm_partials_count_;
      r.value_groups = static_cast<int>(v.value_partials_count_);
      r.conjugates = r.values + v.value_partials_count_;
      r.partials = r.conjugates + v.conjugate_partials_count_;
      r.conjugate_groups = static_cast<int>(v.conjugate_partials_count_);
      int value_at = 0, conjugate_at = 0;
      for (int c = 0; c < 3; ++c) {
        r.value_offset[c] = value_at;
        r.conjugate_offset[c] = conjugate_at;
        if (c < v.C_) {
          value_at += static_cast<int>(value_groups_of(frames_[m]->components[c].blocks()));
          conjugate_at += v.prox_groups(c);
        }
      }
      r.results = v.record_out_.template as<double>();
      r.alpha = settings_.alpha;
      r.alpha0 = settings_.alpha0;
      r.alpha1 = settings_.alpha1;
      r.free_radius = plan.free_radius;
      r.partial_radius = plan.partial_radius;
      r.tolerance = plan.tolerance;
      r.relative_tolerance = plan.relative_tolerance;
      r.partial_tolerance = plan.partial_tolerance;
      r.samples = static_cast<double>(v.C_) * v.H_ * v.W_;
      logs_.emplace_back(sizeof(double) * 5 * capacity);
      r.log = logs_.back().template as<double>();
      r.log_capacity = static_cast<int>(capacity);
    }
    record_slots_ = upload(records);
  }

  // The dual step's rows per wavefront: measured on the batch (the dual step only reads the
  // state and writes the other buffers, so trying it changes nothing), or the environment's.
  int choose_rows() {
    if (const char* rows = std::getenv("UNROUND_DUAL_ROWS")) return std::max(1, std::atoi(rows));
    Key shape;
    for (const auto& v : solvers_) {
      shape.push_back(v->C_);
      shape.push_back(v->H_);
      shape.push_back(v->W_);
    }
    RowsChoice& choices = rows_choices();
    {
      std::lock_guard<std::mutex> hold(choices.lock);
      const auto found = choices.chosen.find(shape);
      if (found != choices.chosen.end()) return found->second;
    }
    hipStream_t s = stream_.stream;
    // Each choice timed in two passes, in opposite orders, the smaller time kept.
    constexpr int kChoices[4] = {8, 16, 32, 64};
    float least[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int pass = 0; pass < 2; ++pass) {
      for (int n = 0; n < 4; ++n) {
        const int k = pass == 0 ? n : 3 - n;
        build_tables(kChoices[k]);
        EventPair timing;
        launch_duals(0, 0);
        HIP_CHECK(hipEventRecord(timing.begin, s));
        for (int launch = 0; launch < 3; ++launch) launch_duals(0, 0);
        HIP_CHECK(hipEventRecord(timing.end, s));
        HIP_CHECK(hipEventSynchronize(timing.end));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, timing.begin, timing.end));
        least[k] = pass == 0 ? ms : std::min(least[k], ms);
        if (std::getenv("UNROUND_BATCH_LOG") != nullptr) {
          std::fprintf(stderr, "batch of %zu: dual step with %d rows a wavefront, %.3f ms\n", solvers_.size(),
                       kChoices[k], ms / 3.0f);
        }
      }
    }
    int pick = 0;
    for (int k = 1; k < 4; ++k) pick = least[k] < least[pick] ? k : pick;
    const int best = kChoices[pick];
    std::lock_guard<std::mutex> hold(choices.lock);
    choices.chosen[shape] = best;
    return best;
  }

  template <typename X>
  DeviceBuffer upload(const std::vector<X>& items) {
    DeviceBuffer buffer(sizeof(X) * std::max<size_t>(1, items.size()));
    if (!items.empty()) {
      HIP_CHECK(hipMemcpyAsync(buffer.pointer, items.data(), sizeof(X) * items.size(), hipMemcpyHostToDevice,
                               stream_.stream));
      HIP_CHECK(hipStreamSynchronize(stream_.stream));
    }
    return buffer;
  }

  struct Launch {
    Key key;  // the kernel's kind
    DeviceBuffer tiles;
    int count = 0;
  };

  void build_tables(int rows) {
    const int M = static_cast<int>(solvers_.size());
    std::vector<ImageSlot<T, S>> even(M), odd(M);
    std::map<Key, std::vector<BatchTile>> prox;
    std::map<Key, std::vector<std::array<int, 3>>> dual;  // per kind: (picture, rows, strip groups)
    std::vector<BatchTile> norms, values, conjugates, partial_conjugates;
    std::vector<ReduceJob> first_jobs, last_jobs;
    const T rho = static_cast<T>(plans_[0]->relaxation);
    rho_ = rho;
    rest_ = static_cast<T>(1.0 - plans_[0]->relaxation);
    for (int m = 0; m < M; ++m) {
      Solver<T, S>& v = *solvers_[m];
      const Plan& plan = *plans_[m];
      if (static_cast<T>(plan.relaxation) != rho) throw GpuError("frames of one batch share their relaxation");
      ImageSlot<T, S>& e = even[m];
      e.F = v.fields_;
      e.components = v.component(0);
      e.H = v.H_;
      e.W = v.W_;
      const int strips = (v.W_ + kStripColumns - 1) / kStripColumns;
      e.across = (strips + 3) / 4;
      e.steps = Steps<T>{};
      for (int c = 0; c < 3; ++c) {
        const bool here = c < v.C_;
        e.steps.descent[c] = here ? static_cast<T>(plan.tau * plan.gamma[c]) : T(0);
        e.steps.ascent[c] = here ? static_cast<T>(plan.sigma * plan.gamma[c]) : T(0);
        e.descent[c] = e.steps.descent[c];
        e.step[c] = here ? static_cast<T>(plan.tau / static_cast<double>(frames_[m]->components[c].cells())) : T(0);
      }
      e.steps.radius = static_cast<T>(tgv_ ? settings_.alpha1 : settings_.alpha);
      e.steps.radius0 = static_cast<T>(settings_.alpha0);
      e.steps.rho = rho;
      e.steps.rest = rest_;
      odd[m] = e;
      std::swap(odd[m].F.p, odd[m].F.p_next);
      std::swap(odd[m].F.w, odd[m].F.w_next);
      std::swap(odd[m].F.r, odd[m].F.r_next);
      // The tiles of each kind.
      const int buffers_prox = v.prox_buffers() ? 1 : 0;
      for (int c = 0; c < v.C_; ++c) {
        const ComponentModel& k = frames_[m]->components[c];
        const bool known = (k.down == 1 || k.down == 2) && (k.across == 1 || k.across == 2);
        if (!known) throw GpuError("a batch takes cells of 1 or 2 samples each way");
        auto& list = prox[{static_cast<int>(k.down), static_cast<int>(k.across), buffers_prox}];
        for (int g = 0; g < v.prox_groups(c); ++g) list.push_back({m, c, g});
        for (int g = 0; g < static_cast<int>(value_groups_of(k.blocks())); ++g) values.push_back({m, c, g});
        for (int g = 0; g < v.prox_groups(c); ++g) {
          conjugates.push_back({m, c, g});
          if (tgv_ && plan.has_partial) partial_conjugates.push_back({m, c, g});
        }
      }
      dual[{v.C_, v.coupled_ ? 1 : 0, v.narrow_spans() ? 1 : 0}].push_back({m, v.H_, e.across});
      for (int g = 0; g < static_cast<int>(v.norm_partials_count_); ++g) norms.push_back({m, 0, g});
      // The reductions, as a frame's record makes them.
      double* partials = v.partials_.template as<double>();
      double* results = v.record_out_.template as<double>();
      for (int column = 0; column < 4; ++column) {
        first_jobs.push_back({m, static_cast<int>(v.norm_partials_count_), 4, column, 1u << 2, partials, results});
      }
      double* value_partials = partials + 4 * v.norm_partials_count_;
      double* conjugate_partials = value_partials + v.value_partials_count_;
      double* partial_partials = conjugate_partials + v.conjugate_partials_count_;
      last_jobs.push_back({m, static_cast<int>(v.value_partials_count_), 1, 0, 0u, value_partials, results + 4});
      last_jobs.push_back(
          {m, static_cast<int>(v.conjugate_partials_count_), 1, 0, 0u, conjugate_partials, results + 5});
      if (tgv_ && plan.has_partial) {
        last_jobs.push_back(
            {m, static_cast<int>(v.conjugate_partials_count_), 1, 0, 0u, partial_partials, results + 6});
      }
    }
    slots_[0] = upload(even);
    slots_[1] = upload(odd);
    const auto launches = [this](std::map<Key, std::vector<BatchTile>>& kinds) {
      std::vector<Launch> made;
      for (auto& [key, list] : kinds) {
        Launch launch;
        launch.key = key;
        launch.count = static_cast<int>(list.size());
        launch.tiles = upload(list);
        made.push_back(std::move(launch));
      }
      return made;
    };
    prox_launches_ = launches(prox);
    dual_launches_.clear();
    for (auto& [key, pictures] : dual) {
      const std::vector<DualTile> list = dual_table(pictures, rows, tail_rows(key));
      Launch launch;
      launch.key = key;
      launch.count = static_cast<int>(list.size());
      launch.tiles = upload(list);
      dual_launches_.push_back(std::move(launch));
    }
    const auto single = [this](std::vector<BatchTile>& list) {
      Launch launch;
      launch.count = static_cast<int>(list.size());
      launch.tiles = upload(list);
      return launch;
    };
    norms_launch_ = single(norms);
    values_launch_ = single(values);
    conjugates_launch_ = single(conjugates);
    partials_launch_ = single(partial_conjugates);
    first_jobs_count_ = static_cast<int>(first_jobs.size());
    last_jobs_count_ = static_cast<int>(last_jobs.size());
    first_jobs_ = upload(first_jobs);
    last_jobs_ = upload(last_jobs);
  }

  // The dual step's table for the pictures of one kind (picture, rows, strip groups): runs of
  // `rows` rows, the pictures in turn, dealt to the 8 XCDs in shares of about equal work;
  // the last `tail` rows of strip groups of each share cut into runs of 8, so that the
  // share's work-groups end close together. Laid out in the order the hardware hands out
  // work-groups (group g to XCD g mod 8, each XCD's in turn); places past a share's end
  // hold no work.
  static std::vector<DualTile> dual_table(const std::vector<std::array<int, 3>>& pictures, int rows, long tail) {
    struct Run {
      int image, first, rows, across;
      long work() const { return static_cast<long>(rows) * across; }
    };
    std::vector<Run> runs;
    long total = 0;
    for (const auto& [image, H, across] : pictures) {
      for (int first = 0; first < H; first += rows) {
        runs.push_back({image, first, std::min(rows, H - first), across});
        total += runs.back().work();
      }
    }
    constexpr int kXcds = 8;
    std::vector<std::vector<DualTile>> shares(kXcds);
    size_t next = 0;
    long dealt = 0;
    for (int x = 0; x < kXcds; ++x) {
      const long goal = total * (x + 1) / kXcds;
      std::vector<Run> mine;
      while (next < runs.size() && (dealt < goal || x == kXcds - 1)) {
        dealt += runs[next].work();
        mine.push_back(runs[next++]);
      }
      size_t cut = mine.size();
      for (long kept = 0; cut > 0 && kept < tail && rows > 8;) kept += mine[--cut].work();
      std::vector<DualTile>& share = shares[x];
      const auto emit = [&share](int image, int first, int count, int across) {
        for (int g = 0; g < across; ++g) share.push_back({image, first, count, g});
      };
      for (size_t k = 0; k < cut; ++k) emit(mine[k].image, mine[k].first, mine[k].rows, mine[k].across);
      for (size_t k = cut; k < mine.size(); ++k) {
        const Run& r = mine[k];
        for (int first = r.first; first < r.first + r.rows; first += 8) {
          emit(r.image, first, std::min(8, r.first + r.rows - first), r.across);
        }
      }
    }
    size_t longest = 0;
    for (const auto& share : shares) longest = std::max(longest, share.size());
    std::vector<DualTile> table(kXcds * longest, DualTile{-1, 0, 0, 0});
    for (int x = 0; x < kXcds; ++x) {
      for (size_t k = 0; k < shares[x].size(); ++k) table[k * kXcds + x] = shares[x][k];
    }
    return table;
  }

  // The rows of strip groups to cut into short runs at the end of an XCD's share: two
  // rounds of the work-groups its CUs hold at once, 8 rows each.
  long tail_rows(const Key& key) const {
    int device = 0, units = 0, resident = 0;
    HIP_CHECK(hipGetDevice(&device));
    HIP_CHECK(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount, device));
    const int C = key[0], coupled = key[1], narrow = key[2];
    const auto held = [&resident](auto kernel) {
      HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, 256, 0));
    };
    if (C == 1) {
      narrow ? held(dual_batch_kernel<T, S, 1, true, false, device::CachedBufferPlanes>)
             : held(dual_batch_kernel<T, S, 1, true, false, device::WidePlanes>);
    } else if (coupled) {
      narrow ? held(dual_batch_kernel<T, S, 3, true, true, device::CachedBufferPlanes>)
             : held(dual_batch_kernel<T, S, 3, true, true, device::WidePlanes>);
    } else {
      narrow ? held(dual_batch_kernel<T, S, 3, true, false, device::CachedBufferPlanes>)
             : held(dual_batch_kernel<T, S, 3, true, false, device::WidePlanes>);
    }
    return 2L * std::max(resident, 1) * std::max(units / 8, 1) * 8;
  }

  void launch_proxes(int parity, int keep) {
    hipStream_t s = stream_.stream;
    const auto* slots = slots_[parity].template as<ImageSlot<T, S>>();
    for (const Launch& launch : prox_launches_) {
      const auto* tiles = launch.tiles.template as<BatchTile>();
      const int dv = launch.key[0], dh = launch.key[1], buffers = launch.key[2];
      const auto go = [&](auto kernel) {
        kernel<<<launch.count, 256, 0, s>>>(tiles, slots, state_.active, rho_, rest_, keep);
      };
      if (dv == 1 && dh == 1 && buffers && plain_) {
        go(prox_batch_kernel<T, S, 1, 1, device::StreamedBufferPlanes, true>);
      } else if (dv == 2 && dh == 2 && buffers && plain_) {
        go(prox_batch_kernel<T, S, 2, 2, device::StreamedBufferPlanes, true>);
      } else if (dv == 1 && dh == 1) {
        buffers ? go(prox_batch_kernel<T, S, 1, 1, device::StreamedBufferPlanes>)
                : go(prox_batch_kernel<T, S, 1, 1, device::WidePlanes>);
      } else if (dv == 2 && dh == 2) {
        buffers ? go(prox_batch_kernel<T, S, 2, 2, device::StreamedBufferPlanes>)
                : go(prox_batch_kernel<T, S, 2, 2, device::WidePlanes>);
      } else if (dv == 1 && dh == 2) {
        buffers ? go(prox_batch_kernel<T, S, 1, 2, device::StreamedBufferPlanes>)
                : go(prox_batch_kernel<T, S, 1, 2, device::WidePlanes>);
      } else {
        buffers ? go(prox_batch_kernel<T, S, 2, 1, device::StreamedBufferPlanes>)
                : go(prox_batch_kernel<T, S, 2, 1, device::WidePlanes>);
      }
      HIP_CHECK(hipGetLastError());
    }
  }

  template <int C, bool TGV, bool COUPLED>
  void launch_dual(const Launch& launch, const ImageSlot<T, S>* slots, int keep) {
    const auto* tiles = launch.tiles.template as<DualTile>();
    if (launch.key[2] && keep == 0) {
      dual_batch_kernel<T, S, C, TGV, COUPLED, device::CachedBufferPlanes, false>
          <<<launch.count, 256, 0, stream_.stream>>>(tiles, slots, state_.active, 0);
    } else if (launch.key[2]) {
      dual_batch_kernel<T, S, C, TGV, COUPLED, device::CachedBufferPlanes>
          <<<launch.count, 256, 0, stream_.stream>>>(tiles, slots, state_.active, keep);
    } else {
      dual_batch_kernel<T, S, C, TGV, COUPLED, device::WidePlanes>
          <<<launch.count, 256, 0, stream_.stream>>>(tiles, slots, state_.active, keep);
    }
  }

  void launch_duals(int parity, int keep) {
    const auto* slots = slots_[parity].template as<ImageSlot<T, S>>();
    for (const Launch& launch : dual_launches_) {
      const int C = launch.key[0], coupled = launch.key[1];
      if (C == 1) {
        tgv_ ? launch_dual<1, true, false>(launch, slots, keep) : launch_dual<1, false, false>(launch, slots, keep);
      } else if (coupled) {
        tgv_ ? launch_dual<3, true, true>(launch, slots, keep) : launch_dual<3, false, true>(launch, slots, keep);
      } else {
        tgv_ ? launch_dual<3, true, false>(launch, slots, keep) : launch_dual<3, false, false>(launch, slots, keep);
      }
      HIP_CHECK(hipGetLastError());
    }
  }

  // The records of every frame still iterating and their stopping tests on stream `s`, then
  // the count of iterations advanced by `length`, the iterations since the last record.
  void launch_records(hipStream_t s, int length) {
    const int M = static_cast<int>(solvers_.size());
    const auto* slots = record_slots_.template as<RecordSlot<T, S>>();
    const auto tiles = [](const Launch& launch) { return launch.tiles.template as<BatchTile>(); };
    if (norms_launch_.count > 0) {
      if (tgv_) {
        norms_batch_kernel<T, S, true><<<norms_launch_.count, 256, 0, s>>>(tiles(norms_launch_), slots, state_.active);
      } else {
        norms_batch_kernel<T, S, false>
            <<<norms_launch_.count, 256, 0, s>>>(tiles(norms_launch_), slots, state_.active);
      }
      HIP_CHECK(hipGetLastError());
    }
    reduce_batch_kernel<<<first_jobs_count_, 256, 0, s>>>(first_jobs_.template as<ReduceJob>(), state_.active);
    HIP_CHECK(hipGetLastError());
    if (values_launch_.count > 0) {
      value_batch_kernel<T, S><<<values_launch_.count, 256, 0, s>>>(tiles(values_launch_), slots, state_.active);
      HIP_CHECK(hipGetLastError());
    }
    if (conjugates_launch_.count > 0) {
      conjugate_batch_kernel<T, S><<<conjugates_launch_.count, 256, 0, s>>>(tiles(conjugates_launch_), slots,
                                                                           state_.active, tgv_ ? 1 : 0, 0);
      HIP_CHECK(hipGetLastError());
    }
    if (partials_launch_.count > 0) {
      conjugate_batch_kernel<T, S>
          <<<partials_launch_.count, 256, 0, s>>>(tiles(partials_launch_), slots, state_.active, 0, 1);
      HIP_CHECK(hipGetLastError());
    }
    reduce_batch_kernel<<<last_jobs_count_, 256, 0, s>>>(last_jobs_.template as<ReduceJob>(), state_.active);
    HIP_CHECK(hipGetLastError());
    stop_kernel<T, S><<<(M + 255) / 256, 256, 0, s>>>(slots, M, state_, length);
    HIP_CHECK(hipGetLastError());
    advance_kernel<<<1, 64, 0, s>>>(state_.counter, length);
    HIP_CHECK(hipGetLastError());
  }

  // `length` iterations from `parity`, the last keeping the tilde point where `keep`.
  void launch_iterations(int length, int parity, bool keep) {
    for (int t = 0; t < length; ++t) {
      const int p = (parity + t) & 1;
      const int kept = keep && t == length - 1 ? 1 : 0;
      launch_proxes(p, kept);
      launch_duals(p, kept ? kKeepTilde | (!tgv_ || plans_[0]->has_partial ? kKeepDualP : 0) : 0);
    }
  }

  // The graph of `length` iterations (kind 0) on the iterations' stream, or of the records
  // after `length` iterations (kind 1) on the records' stream: captured once, then replayed.
  hipGraphExec_t graph_of(int kind, int length, int parity, bool keep) {
    const Key key{kind, length, parity, keep ? 1 : 0};
    const auto found = graphs_.find(key);
    if (found != graphs_.end()) return found->second;
    hipStream_t s = kind == 0 ? stream_.stream : records_stream_.stream;
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (kind == 0) {
      launch_iterations(length, parity, keep);
    } else {
      launch_records(s, length);
    }
    HIP_CHECK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t made = nullptr;
    const hipError_t instantiated = hipGraphInstantiate(&made, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_CHECK(instantiated);
    graphs_[key] = made;
    return made;
  }

  // The iterations run on one stream and the records on another, so that a record's kernels
  // run beside the next run of iterations: the run's last iteration, which keeps the tilde
  // point anew, waits for the previous records (the readers of the point it overwrites). A
  // frame that a record stops may take some of the next run's iterations, which change
  // neither its tilde point nor its records; the host reads which frames still iterate one
  // record behind, so that the iterations' stream always holds a run.
  void iterate_all() {
    hipStream_t s = stream_.stream, r = records_stream_.stream;
    const int M = static_cast<int>(solvers_.size());
    EventPair whole;
    HIP_CHECK(hipEventRecord(whole.begin, s));
    const uint64_t most = most_, every = every_;
    if (every > 0 || most == 0) launch_records(s, 0);
    int* flags = nullptr;
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&flags), sizeof(int) * 2 * std::max(M, 1),
                            hipHostMallocDefault));
    const std::unique_ptr<int, hipError_t (*)(void*)> held(flags, hipHostFree);
    int* const pinned = held.get();
    const auto any_of = [M](const int* flags) {
      for (int m = 0; m < M; ++m)
        if (flags[m] != 0) return true;
      return false;
    };
    std::deque<std::pair<uint64_t, EventPair>> ends;
    uint64_t iteration = 0, since = 0;
    int parity = 0, recorded = 0;
    HIP_CHECK(hipMemcpyAsync(pinned, state_.active, sizeof(int) * M, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    bool going = any_of(pinned);
    while (going && iteration < most) {
      // Up to the next record, in runs of at most kRun iterations.
      constexpr uint64_t kRun = 10;
      uint64_t next_record = most;
      if (every > 0) next_record = std::min(most, (iteration / every + 1) * every);
      const uint64_t length = std::min(kRun, next_record - iteration);
      const bool record = iteration + length == next_record;
      since += length;
      if (!record) {
        HIP_CHECK(hipGraphLaunch(graph_of(0, static_cast<int>(length), parity, false), s));
      } else {
        const int last = (parity + static_cast<int>(length) - 1) & 1;
        if (length > 1) HIP_CHECK(hipGraphLaunch(graph_of(0, static_cast<int>(length) - 1, parity, false), s));
        if (recorded > 0) HIP_CHECK(hipStreamWaitEvent(s, recorded_[(recorded - 1) & 1].event, 0));
        HIP_CHECK(hipGraphLaunch(graph_of(0, 1, last, true), s));
        HIP_CHECK(hipEventRecord(kept_.event, s));
        HIP_CHECK(hipStreamWaitEvent(r, kept_.event, 0));
        HIP_CHECK(hipGraphLaunch(graph_of(1, static_cast<int>(since), 0, false), r));
        HIP_CHECK(hipMemcpyAsync(pinned + (recorded & 1) * M, state_.active, sizeof(int) * M,
                                 hipMemcpyDeviceToHost, r));
        HIP_CHECK(hipEventRecord(recorded_[recorded & 1].event, r));
        ++recorded;
        since = 0;
      }
      iteration += length;
      parity = (parity + static_cast<int>(length)) & 1;
      ends.emplace_back();
      ends.back().first = iteration;
      HIP_CHECK(hipEventRecord(ends.back().second.end, record ? r : s));
      if (record && recorded >= 2) {
        // The flags after the previous record (or a later one: a frame only ever stops).
        HIP_CHECK(hipEventSynchronize(recorded_[recorded & 1].event));
        going = any_of(pinned + (recorded & 1) * M);
      }
    }
    HIP_CHECK(hipStreamSynchronize(s));
    HIP_CHECK(hipStreamSynchronize(r));
    reached_ = iteration;
    for (const auto& [at, pair] : ends) {
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, whole.begin, pair.end));
      ends_[at] = ms * 1e-3;
    }
    stop_.assign(M, 0);
    stopped_at_.assign(M, 0);
    record_counts_.assign(M, 0);
    HIP_CHECK(hipMemcpy(stop_.data(), state_.stop, sizeof(int) * M, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(stopped_at_.data(), state_.stopped_at, sizeof(unsigned long long) * M, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(record_counts_.data(), state_.records, sizeof(int) * M, hipMemcpyDeviceToHost));
    logs_host_.assign(M, {});
    const uint64_t capacity = (every > 0 ? most / every : 0) + 2;
    for (int m = 0; m < M; ++m) {
      logs_host_[m].resize(5 * capacity);
      HIP_CHECK(hipMemcpy(logs_host_[m].data(), logs_[m].pointer, sizeof(double) * 5 * capacity,
                          hipMemcpyDeviceToHost));
    }
  }

  const std::vector<const FrameModel*>& frames_;
  const std::vector<const Plan*>& plans_;
  const Settings& settings_;
  bool tgv_ = false;
  bool plain_ = false;  // every component: no slack cost, midpoint centres
  uint64_t most_ = 0, every_ = 0, reached_ = 0;
  T rho_ = T(0), rest_ = T(0);
  StreamHolder stream_, records_stream_;
  Mark kept_, recorded_[2];
  std::vector<std::unique_ptr<Solver<T, S>>> solvers_;
  DeviceBuffer state_memory_, record_slots_, first_jobs_, last_jobs_;
  DeviceBuffer slots_[2];
  std::vector<DeviceBuffer> logs_;
  BatchState state_{};
  std::vector<Launch> prox_launches_, dual_launches_;
  Launch norms_launch_, values_launch_, conjugates_launch_, partials_launch_;
  int first_jobs_count_ = 0, last_jobs_count_ = 0;
  std::map<Key, hipGraphExec_t> graphs_;
  std::map<uint64_t, double> ends_;
  std::vector<int> stop_, record_counts_;
  std::vector<unsigned long long> stopped_at_;
  std::vector<std::vector<double>> logs_host_;
};
