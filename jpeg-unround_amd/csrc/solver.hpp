// The GPU solve of a frame: the primal-dual iterations of docs/math.md 5 on HIP kernels,
// their records (docs/math.md 6) reduced in binary64 on the device, and the final binary64
// projection of the last canvas onto the quantization constraint set.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "model.hpp"
#include "settings.hpp"

namespace unround {

// A record of the solver (cpp/include/unround/results.hpp:17-24): NaN where not taken.
struct Record {
  uint64_t iteration = 0;
  double seconds = 0.0;
  double primal = 0.0;
  double dual = 0.0;
  double theta = 0.0;
  double partial = 0.0;
};

// A state to start from (cpp/include/unround/pdhg.hpp:128-133): null pointers for the
// default. coefficients has one pointer per component (its blocks x 64), p and w are
// [C][2][H][W], r [C][3][H][W], in binary64.
struct StartState {
  std::vector<const double*> coefficients;
  const double* p = nullptr;
  const double* w = nullptr;
  const double* r = nullptr;
};

// Called at each record after the first with the tilde point's canvas (C x H x W
// binary64); true stops the solver.
using Observer = std::function<bool(const Record&, const double* canvas)>;

// How long the kernels of the iterations took, measured with HIP events on the solve's
// stream where timing is asked for: Timing::iterations brackets each run of iterations
// between two records with one pair of events (nothing between the launches);
// Timing::kernels brackets every iteration and every kernel launch.
enum class Timing { none = 0, iterations = 1, kernels = 2 };

struct KernelTimes {
  double primal_seconds = 0.0;
  uint64_t primal_launches = 0;
  double dual_seconds = 0.0;
  uint64_t dual_launches = 0;
  double iterations_seconds = 0.0;  // the iterations alone, records excluded
  uint64_t iterations = 0;
  int dual_rows = 0;  // the dual step's rows per wavefront (a frame's solve)
};

struct SolveOutput {
  std::vector<double> canvas;                     // C x H x W, the final projection
  std::vector<std::vector<double>> coefficients;  // per component, blocks x 64
  std::vector<Record> records;
  uint64_t iterations = 0;
  int32_t stop = 0;  // 0 no solver, 1 converged, 2 iterations, 3 observer
  double tau = 0.0, sigma = 0.0;
  KernelTimes times;
};

// Solves on the current HIP device. false with a message (and a status) on failure.
bool solve_on_gpu(const FrameModel& frame, const Settings& settings, const Plan& plan, const StartState* start,
                  const Observer& observer, Timing timing, SolveOutput& out, std::string& message, int32_t& status);

// A canvas (C x H x W binary64, the frame's) projected onto the frame's intervals on the
// current HIP device (docs/math.md 12): out.canvas and out.coefficients as a final
// projection makes them, without a margin, and excess[c] how far each coefficient of
// component c lay outside its interval, in steps.
bool project_on_gpu(const FrameModel& frame, const Settings& settings, const std::vector<double>& canvas,
                    SolveOutput& out, std::vector<std::vector<double>>& excess, std::string& message,
                    int32_t& status);

// A picture of rounded samples (TV or TGV within its intervals) on the current HIP device:
// out.canvas is the picture, channels x height x width, every sample within its interval;
// no coefficients.
bool solve_samples_on_gpu(const SamplesModel& model, const Settings& settings, const Plan& plan,
                          const Observer& observer, Timing timing, SolveOutput& out, std::string& message,
                          int32_t& status);

// Frames solved together on the GPU, with the same settings (and so the same iterations and
// records) and each its own plan: the outputs are each frame's own solve's (its iterates,
// records and stopping iteration), and latencies[m] the seconds from the first iteration
// of the batch to the end of the run of iterations in which frame m stopped. No observer.
// Unrolled TGV iterations on the GPU with their adjoint (grey frames of one size, blocks
// covering them, binary32): the forward from the start with p's ball of radius
// alpha1 m(i, j) at pixel (i, j) and r's of alpha0, keeping the state every `every`
// iterations; the backward of a loss's gradient on the last x to the gradients of m, alpha0
// and alpha1. Device pointers; `stream` a hipStream_t the calls are ordered after and
// before (null: the default stream).
class UnrolledSolve {
 public:
  virtual ~UnrolledSolve() = default;
  virtual int frames() const = 0;
  virtual int height() const = 0;
  virtual int width() const = 0;
  // tau, sigma, gamma, rho.
  virtual void constants(double* out) const = 0;
  // The start state [frame][x, w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy][H][W], host memory.
  virtual void start_state(double* out) = 0;
  // canvas: [frame][H][W] floats; map: [H][W] floats.
  virtual void forward(const float* map, double alpha0, double alpha1, float* canvas, void* stream) = 0;
  // Of the last forward: grad_canvas [frame][H][W] to grad_map [H][W] (device) and
  // grad_alphas[0] = dL/d alpha0, [1] = dL/d alpha1 (host).
  virtual void backward(const float* grad_canvas, float* grad_map, double* grad_alphas, void* stream) = 0;
  // Where launches are timed (timing == Timing::kernels at construction), the seconds and
  // launches since the last forward began of: the dual step's adjoint, the prox's adjoint,
  // the forward dual step (the forward's and the backward's recomputation); false where
  // nothing is timed.
  virtual bool kernel_times(double* out) = 0;
};

// Null with the reason and a status where the frames or the settings are refused or the
// GPU fails. The frames and plans must outlive the result.
std::unique_ptr<UnrolledSolve> make_unrolled(const std::vector<const FrameModel*>& frames,
                                             const std::vector<const Plan*>& plans, const Settings& settings,
                                             int iterations, int every, Timing timing, std::string& message,
                                             int32_t& status);

// The bytes of the fixed device region behind solve_in_scratch, on each GPU.
constexpr size_t kScratchBytes = size_t{64} << 20;

// solve_on_gpu from the default start with no observer, allocating no device memory:
// every buffer of the solve is a stretch of a fixed region of kScratchBytes that the
// library holds on each GPU (one solve at a time holds it: a lock per GPU orders them),
// the work on the calling thread's default stream, and the final projection's planes (the
// canvas cut to the picture, channel by channel, binary64) copied into `planes`, host
// memory of channels x height x width doubles. A frame whose solve needs more than the
// region is refused with status 1 before anything runs on the GPU.
bool solve_in_scratch(const FrameModel& frame, const Settings& settings, const Plan& plan, double* planes,
                      uint64_t& iterations, int32_t& stop, std::string& message, int32_t& status);

bool solve_batch_on_gpu(const std::vector<const FrameModel*>& frames, const std::vector<const Plan*>& plans,
                        const Settings& settings, std::vector<SolveOutput>& outs, std::vector<double>& latencies,
                        std::string& message, int32_t& status);

}  // namespace unround
