/* TEST INFRASTRUCTURE ONLY.
 *
 * The CPU oracle of the TV/TGV primal-dual reconstruction of a JPEG's quantized
 * coefficients: a plain-C, binary64 restatement of the reference's algorithm in the
 * natural layout, used only by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker. Nothing of the product links or calls it.
 *
 * It follows docs/math.md (sections Arithmetic, 1.1-1.3, 2, 3, 4.1-4.4, 5, 6.1-6.6)
 * and the natural-layout C++ implementation of the reference (cpp/src/{dct,model,
 * frames,operators,pdhg,exact,laplace}.cpp, cpp/src/primal_dual.hpp), whose order of
 * operations it keeps. Each function in unround_oracle.c cites the file and lines it
 * restates.
 */
#ifndef UNROUND_ORACLE_H
#define UNROUND_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A component as arrays: rows x columns blocks of 64 levels in natural order. */
typedef struct oracle_component {
  const int16_t* levels;
  uint32_t rows;
  uint32_t columns;
  int32_t h;
  int32_t v;
  uint16_t table[64];
} oracle_component;

/* The options of the data term G of one component (cpp/include/unround/model.hpp:39-63). */
typedef struct oracle_data {
  int32_t has_mu;
  double mu;
  double mu_scale;
  double mu_power;
  double mu_chroma;
  double slack;
  double slack_cost;
  double dc_weight;
  double power;
  double flat_weight;
  uint32_t flat_levels;
  int32_t has_flat_dc;
  uint32_t flat_dc;
  int32_t centres; /* 0 the MMSE centres, 1 the middles */
} oracle_data;

/* The method and the solver's options (cpp/include/unround/pdhg.hpp:29-75). */
typedef struct oracle_options {
  int32_t method; /* 0 the MMSE decoder, 1 TV, 2 TGV */
  double alpha;
  double alpha1;
  double alpha0;
  int32_t coupled;
  int32_t has_channel_weights;
  double channel_weights[3];
  int32_t has_iterations;
  uint64_t iterations;
  int32_t has_tolerance;
  double tolerance;
  double relative_tolerance;
  double partial_tolerance;
  int32_t has_partial_radius;
  double partial_radius;
  int32_t has_step_ratio;
  double step_ratio;
  int32_t has_relaxation;
  double relaxation;
  double step_product;
  int32_t has_norm_squared;
  double norm_squared;
  int32_t weight_scaling;
  uint64_t record_every;
  double free_radius;
} oracle_options;

/* A state to start from (cpp/include/unround/pdhg.hpp:128-133): any pointer may be
 * null. coefficients[c] has the component's blocks x 64; p is [C][2][H][W] (x then y),
 * w likewise, r [C][3][H][W] (xx, yy, xy). */
typedef struct oracle_start {
  const double* coefficients[3];
  const double* p;
  const double* w;
  const double* r;
} oracle_start;

typedef struct oracle_result oracle_result;

/* Solves; null with a message where the input or the options are refused. */
oracle_result* oracle_solve(uint32_t height, uint32_t width, int32_t color_space, uint32_t count,
                            const oracle_component* components, const oracle_data* data,
                            const oracle_options* options, const oracle_start* start, char* message,
                            size_t message_size);
void oracle_result_free(oracle_result* result);

/* The canvas, channels x rows x columns. */
const double* oracle_result_canvas(const oracle_result* result, uint64_t* channels, uint64_t* rows,
                                   uint64_t* columns);
/* The last iterate x (the canvas is the last tilde point x~, projected into the
 * intervals), the canvas's shape. */
const double* oracle_result_iterate(const oracle_result* result);
/* A component's blocks, and its cell (rows, columns of the canvas per sample). */
void oracle_result_component(const oracle_result* result, uint32_t c, uint64_t* block_rows,
                             uint64_t* block_columns, uint64_t* down, uint64_t* across);
/* which: 0 coefficients, 1 lower, 2 upper, 3 centres (blocks x 64); 4 steps, 5 weights,
 * 6 flat weights (64); 7 flat marks as doubles (blocks). */
const double* oracle_result_array(const oracle_result* result, uint32_t c, int32_t which, uint64_t* count);
/* tau, sigma, iterations, stop (1 converged, 2 iterations, 0 no solver). */
/* Wall seconds spent in the iterations of a solve (records and set-up excluded). */
double oracle_result_iteration_seconds(const oracle_result* result);
/* The least wall seconds of one iteration of the solve. */
double oracle_result_least_iteration_seconds(const oracle_result* result);
/* The threads the iterations take (0: every core), as bands of rows and blocks; each sample
 * is made by one thread alone, in the same operations whatever the count, so results do
 * not depend on it (cpp/src/workers.hpp:21-86 does the same with its strips). */
void oracle_set_threads(int32_t count);
int32_t oracle_threads(void);
void oracle_result_solver(const oracle_result* result, double* tau, double* sigma, uint64_t* iterations,
                          int32_t* stop);
/* Records: iterations, and which 0 primal, 1 dual, 2 theta (NaN for TV), 3 partial gap
 * (NaN where not taken). */
const uint64_t* oracle_result_record_iterations(const oracle_result* result, uint64_t* count);
const double* oracle_result_record_values(const oracle_result* result, int32_t which, uint64_t* count);

/* Pictures of rounded samples (docs/math.md, 11; rust/unround/src/samples.rs): the
 * intervals and centres of levels of `bits` bits (1, 2, 4, 8 or 16) with a slack in
 * levels, into low/high/centre (count each); 0 where the bits, a level or the slack are
 * refused. */
int oracle_samples_intervals(const uint16_t* levels, uint64_t count, uint32_t bits, double slack, double* low,
                             double* high, double* centre);
/* The proximal map of tau G at values (count), G's value at x and G* at xi, of a model of
 * `channels x height x width` levels (count = their product). */
void oracle_samples_prox(const uint16_t* levels, uint64_t count, uint32_t bits, double mu, double slack,
                         const double* values, double tau, double* out);
double oracle_samples_value(const uint16_t* levels, uint64_t channels, uint64_t height, uint64_t width,
                            uint32_t bits, double mu, double slack, const double* x);
double oracle_samples_conjugate(const uint16_t* levels, uint64_t channels, uint64_t height, uint64_t width,
                                uint32_t bits, double mu, double slack, const double* xi);
/* TV (options->method 1) or TGV (2) within the intervals: the picture is the canvas,
 * channels x height x width; the defaults of rounded samples replace the options' where
 * they are not given (has_...): tolerance 1e-4 and the step ratio 10 (TV) or 0.3 (TGV),
 * each with the first-order weight as JPEG's are, and 10000 iterations. */
oracle_result* oracle_samples_solve(const uint16_t* levels, uint32_t channels, uint32_t height, uint32_t width,
                                    uint32_t bits, double mu, double slack, const oracle_options* options,
                                    char* message, size_t message_size);

/* Exposed parts, for tests of their own. */
void oracle_basis(double out[64]);
void oracle_dct_forward(const double in[64], double out[64]);
void oracle_dct_inverse(const double in[64], double out[64]);
double oracle_laplace_scale(uint64_t zeros, uint64_t others, uint64_t sum, double step);
double oracle_laplace_shrinkage(double rho);
/* The MMSE centre of an AC level of step Q at the Laplace scale given. */
double oracle_mmse_centre(int16_t level, double step, double scale);
double oracle_sum_rows(const double* terms, uint64_t rows, uint64_t columns);

#ifdef __cplusplus
}
#endif

#endif
