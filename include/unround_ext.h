/* Additions of the MI355X edition to the C interface of JPEG-Unround (include/unround.h).
 *
 * These entry points have no counterpart in the reference's C ABI: they expose what its
 * library API has (a start to resume from) or what this GPU edition adds (kernel timings,
 * batches of pictures). The reference's own declarations are unchanged, and so is
 * UNROUND_ABI_VERSION.
 */
#ifndef UNROUND_EXT_H
#define UNROUND_EXT_H

#include "unround.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A state to start the solver from (the reference's library start,
 * cpp/include/unround/pdhg.hpp:128-133 and rust/unround/src/pdhg.rs:354-363, which its C
 * ABI does not expose): any pointer may be null. coefficients holds one pointer per
 * component (its rows x columns x 64 coefficients, clipped to their intervals) or is
 * null; p and w are [C][2][H][W] and r [C][3][H][W] binary64 fields of the canvas (p and
 * r are projected onto their balls; w defaults to the gradient of x). */
typedef struct unround_start {
  const double* const* coefficients;
  const double* p;
  const double* w;
  const double* r;
} unround_start;

/* unround_solve from a given state. */
unround_status unround_solve_start(const unround_input* input, const unround_settings* settings,
                                   const unround_start* start, unround_observer observer, void* user,
                                   unround_result** result, char* message, size_t message_size);

/* unround_solve of `count` inputs together on the GPU, with one settings (so the same
 * iterations and records for all): results[k] is what unround_solve gives for inputs[k]
 * (the same iterates, records and stopping iteration), to be freed with unround_result_free
 * each. Every input's tiles run in the same launches, a run of iterations between two
 * records is one HIP graph, and the records and stopping tests run on the device (on a
 * second stream, beside the next run of iterations). No observer. Where latencies is not
 * null, latencies[k] is the seconds from the batch's first iteration to the end of the
 * record at which input k stopped (or of the last run of iterations). On an error no
 * result is made. No counterpart in the reference: the batch of the reference's
 * unround_solve calls (rust/capi/include/unround.h:138-140) a caller would otherwise make. */
unround_status unround_solve_batch(const unround_input* inputs, size_t count, const unround_settings* settings,
                                   unround_result** results, double* latencies, char* message,
                                   size_t message_size);

/* Unrolled TGV with its adjoint, for gradients with respect to alpha0, alpha1 and a map
 * m(i, j) weighting the first-order term per pixel: p's ball at pixel (i, j) has radius
 * alpha1 m(i, j) (the reference has one radius per channel, docs/math.md:544-566). Grey
 * inputs of one size whose blocks cover them, no slack cost, --method tgv; binary32; the
 * settings give everything but the alphas. No counterpart in the reference: for training
 * the method's weights through a loss on its result. */
typedef struct unround_unrolled unround_unrolled;

/* A handle for `count` inputs, `iterations` iterations, the state kept every
 * checkpoint_every iterations for the backward (memory: (iterations / checkpoint_every + 1
 * + checkpoint_every) x 8 planes per input). */
unround_status unround_unrolled_new(const unround_input* inputs, size_t count, const unround_settings* settings,
                                    uint32_t iterations, uint32_t checkpoint_every, unround_unrolled** made,
                                    char* message, size_t message_size);
void unround_unrolled_free(unround_unrolled* handle);
unround_status unround_unrolled_shape(const unround_unrolled* handle, uint64_t* count, uint64_t* height,
                                      uint64_t* width);
/* tau, sigma, the channel's gamma and rho. */
unround_status unround_unrolled_constants(const unround_unrolled* handle, double* values);
/* The start state, [count][x, w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy][height][width], into host
 * memory. */
unround_status unround_unrolled_start(unround_unrolled* handle, double* state, char* message, size_t message_size);
/* Input `frame`'s problem, [blocks][64][lower, upper, centre, weight], into host memory. */
unround_status unround_unrolled_problem(const unround_unrolled* handle, uint64_t frame, double* values);
/* The forward from the start: weight_map [height][width] and canvas [count][height][width]
 * are device memory (floats); the canvas gets the last x. `stream` is the hipStream_t the
 * call is ordered after and before (null: the default stream). */
unround_status unround_unrolled_forward(unround_unrolled* handle, const float* weight_map, double alpha0,
                                        double alpha1, float* canvas, void* stream, char* message,
                                        size_t message_size);
/* The backward of the last forward: grad_canvas [count][height][width] (device) to grad_map
 * [height][width] (device) and grad_alphas = {dL/d alpha0, dL/d alpha1} (host). */
unround_status unround_unrolled_backward(unround_unrolled* handle, const float* grad_canvas, float* grad_map,
                                         double* grad_alphas, void* stream, char* message, size_t message_size);
/* The durations of the handle's launches since its last forward began, where the
 * environment variable UNROUND_TIMING was 2 when the handle was made (HIP events around
 * every launch on the handle's stream): values[0] the seconds of the dual step's adjoint and
 * [1] its launches, [2] and [3] those of the prox's adjoint, [4] and [5] those of the
 * forward dual step (the forward's and the backward's recomputation); at most count values
 * are written. Waits for the handle's stream. Returns 1 where launches were timed, 0
 * otherwise. */
int32_t unround_unrolled_kernel_times(unround_unrolled* handle, double* values, size_t count);

/* A picture of rounded samples given as arrays, reconstructed as unround_decode does a PNG
 * file (rust/capi/include/unround.h:128-136; rust/unround/src/rounded.rs:71-170): levels of
 * `bits` bits (1, 2, 4, 8 or 16), height x width x channels interleaved as a PNG file
 * stores them (1 grey, 2 grey and alpha, 3 RGB, 4 RGB and alpha); the colours together by
 * TV or TGV within the intervals of their levels (docs/math.md 11) with --mu and --slack
 * alone of the data term and the defaults of rounded samples, and the alpha alone, without
 * the colours' weights of channels. The result is a PNG file's: unround_result_format 2,
 * unround_result_bits the bits, the picture interleaved as given and every sample within
 * its interval, the canvas its planes, no components, the alpha's solver through
 * unround_result_alpha_solver. observer, where it is not null, is called with user and
 * each record after the first: the colours', then the alpha's. No counterpart in the
 * reference's C ABI, whose unround_decode reads the levels from a PNG file: the same
 * solve for levels a caller already holds. */
unround_status unround_solve_samples(const uint16_t* levels, uint64_t height, uint64_t width, uint64_t channels,
                                     uint32_t bits, const unround_settings* settings, unround_observer observer,
                                     void* user, unround_result** result, char* message, size_t message_size);

/* The bytes of the device region behind unround_solve_scratch, held on each GPU. */
#define UNROUND_SCRATCH_BYTES (64u << 20)

/* No counterpart in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * unround_solve of one input from the default start, with no observer and no result
 * object, allocating no device memory: every buffer of the solve is a stretch of a
 * region of UNROUND_SCRATCH_BYTES that the library holds on each GPU from its first use
 * to the end of the process (a file-scope device array). The final canvas cut to the
 * picture, channel by channel (what unround_result_planes gives), is written to `planes`,
 * host memory of `count` = components x height x width doubles; iterations and stop (1
 * converged, 2 iterations; 0 for the decoder of the centres) as unround_result_solver
 * gives them, where not null. An input whose solve needs more than the region (about 110
 * bytes a canvas sample in binary32 TGV: 512 x 512 grey fits) is refused with
 * UNROUND_ERROR_ARGUMENT before anything runs. The region keeps the last call's contents
 * and is cleared where a solve needs zeros; calls that overlap on one GPU would race for
 * it, so the library makes them take turns (a lock per GPU). The work runs on the calling
 * thread's default stream. */
unround_status unround_solve_scratch(const unround_input* input, const unround_settings* settings, double* planes,
                                     uint64_t count, uint64_t* iterations, int32_t* stop, char* message,
                                     size_t message_size);

/* The durations of the solve's kernels, where the environment variable UNROUND_TIMING was
 * set during the solve (HIP events on the solve's stream): values[0] the seconds of the
 * primal kernels and [1] their launches, [2] and [3] those of the dual kernels, [4] the
 * seconds of the iterations (records excluded), [5] the iterations and [6] the dual step's
 * rows per wavefront; at most count values are written. UNROUND_TIMING=1 brackets each run of iterations between two
 * records with one pair of events (launches [1] and [3] are then 0); UNROUND_TIMING=2
 * brackets every iteration and every launch. UNROUND_TIMING_WARMUP=W leaves the first W
 * iterations untimed (they run straight before the timed ones). Returns 1 where timings
 * were taken, 0 otherwise. */
int32_t unround_result_kernel_times(const unround_result* result, double* values, size_t count);

#ifdef __cplusplus
}
#endif

#endif
