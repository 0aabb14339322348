/* TEST INFRASTRUCTURE ONLY: the binary64 CPU oracle of the TV/TGV primal-dual
 * reconstruction (unround_oracle.h says what it is for). Every function names the
 * reference file and lines whose algorithm it restates; the order of each value's
 * operations is the reference's, so that only rounding (fused or not) separates the two.
 *
 * Whole fields are taken a pass at a time where the reference sweeps rows: every value
 * is computed from the same operands by the same operations either way
 * (cpp/src/primal_dual.hpp:8-16, docs/math.md 5.1).
 */
#include <omp.h>
#include <time.h>
#include "unround_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------------------
 * Arithmetic helpers */

static double clamp_to(double value, double low, double high) {
  /* std::min(std::max(value, low), high): cpp/src/model.cpp:50 */
  double up = value < low ? low : value;
  return high < up ? high : up;
}

static void say(char* message, size_t size, const char* format, ...) {
  if (message == NULL || size == 0) return;
  va_list args;
  va_start(args, format);
  vsnprintf(message, size, format, args);
  va_end(args);
}

/* The sum of a row's terms in eight lanes (docs/math.md:49-58; cpp/src/exact.cpp:246-261). */
static double row_sum(const double* terms, size_t count) {
  double lanes[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  size_t whole = count / 8u;
  for (size_t row = 0; row < whole; ++row)
    for (size_t lane = 0; lane < 8u; ++lane) lanes[lane] += terms[8u * row + lane];
  for (size_t lane = 0; lane < count - 8u * whole; ++lane) lanes[lane] += terms[8u * whole + lane];
  double total = lanes[0];
  for (size_t lane = 1; lane < 8u; ++lane) total += lanes[lane];
  return total;
}

/* Rows' sums in blocks of 128, the blocks in a pairwise tree (cpp/src/exact.cpp:263-291). */
double oracle_sum_rows(const double* terms, uint64_t rows, uint64_t columns) {
  if (columns == 0 || rows == 0) return 0.0;
  size_t blocks = (size_t)((rows + 127u) / 128u);
  double* level = (double*)malloc(sizeof(double) * blocks);
  double block = 0.0;
  size_t count = 0, filled = 0;
  for (uint64_t i = 0; i < rows; ++i) {
    block += row_sum(terms + i * columns, (size_t)columns);
    ++count;
    if (count % 128u == 0) {
      level[filled++] = block;
      block = 0.0;
    }
  }
  if (count % 128u != 0) level[filled++] = block;
  while (filled > 1) {
    size_t next = 0;
    for (size_t index = 0; index + 1u < filled; index += 2u) level[next++] = level[index] + level[index + 1u];
    if (filled % 2u != 0) level[next++] = level[filled - 1u];
    filled = next;
  }
  double total = level[0];
  free(level);
  return total;
}

/* ------------------------------------------------------------------------------------
 * The DCT (docs/math.md:87-122; cpp/src/dct.cpp:20-56, 80-184) */

/* cos(pi j / 16), j = 0..8, each literal's double the nearest to the cosine. */
static const double cosines[9] = {
    1.0,
    9.8078528040323044912618223613424e-1,
    9.2387953251128675612818318939679e-1,
    8.3146961230254523707878837761791e-1,
    7.0710678118654752440084436210485e-1,
    5.5557023301960222474283081394853e-1,
    3.8268343236508977172845998403040e-1,
    1.9509032201612826784828486847702e-1,
    0.0,
};
/* sqrt(1/8) */
static const double root_eighth = 3.5355339059327376220042218105242e-1;

static double cosine_of(unsigned m) {
  if (m <= 8u) return cosines[m];
  if (m <= 16u) return -cosines[16u - m];
  if (m <= 24u) return -cosines[m - 16u];
  return cosines[32u - m];
}

static double basis[8][8];
static int basis_made = 0;

static void make_basis(void) {
  if (basis_made) return;
  for (int n = 0; n < 8; ++n) basis[0][n] = root_eighth;
  for (int k = 1; k < 8; ++k)
    for (int n = 0; n < 8; ++n) basis[k][n] = 0.5 * cosine_of((unsigned)(((2 * n + 1) * k) % 32));
  basis_made = 1;
}

void oracle_basis(double out[64]) {
  make_basis();
  for (int k = 0; k < 8; ++k)
    for (int n = 0; n < 8; ++n) out[8 * k + n] = basis[k][n];
}

/* y = C x of eight values at stride sx into stride sy, by even and odd halves. */
static void forward8(const double* x, size_t sx, double* y, size_t sy) {
  const double s0 = x[0] + x[7 * sx], s1 = x[sx] + x[6 * sx], s2 = x[2 * sx] + x[5 * sx], s3 = x[3 * sx] + x[4 * sx];
  const double d0 = x[0] - x[7 * sx], d1 = x[sx] - x[6 * sx], d2 = x[2 * sx] - x[5 * sx], d3 = x[3 * sx] - x[4 * sx];
  for (int k = 0; k < 8; k += 2) {
    const double* c = basis[k];
    y[k * sy] = c[0] * s0 + c[1] * s1 + c[2] * s2 + c[3] * s3;
    const double* o = basis[k + 1];
    y[(k + 1) * sy] = o[0] * d0 + o[1] * d1 + o[2] * d2 + o[3] * d3;
  }
}

/* x = C^T y likewise: even and odd parts, then their sum and difference. */
static void inverse8(const double* y, size_t sy, double* x, size_t sx) {
  const double a0 = y[0], a1 = y[sy], a2 = y[2 * sy], a3 = y[3 * sy];
  const double a4 = y[4 * sy], a5 = y[5 * sy], a6 = y[6 * sy], a7 = y[7 * sy];
  for (int n = 0; n < 4; ++n) {
    const double even = basis[0][n] * a0 + basis[2][n] * a2 + basis[4][n] * a4 + basis[6][n] * a6;
    const double odd = basis[1][n] * a1 + basis[3][n] * a3 + basis[5][n] * a5 + basis[7][n] * a7;
    x[n * sx] = even + odd;
    x[(7 - n) * sx] = even - odd;
  }
}

/* A block of a grid of row pitch `pitch` into 64 coefficients [v][u]: down the columns,
 * then along the rows (cpp/src/dct.cpp:206-212). */
static void forward_block(const double* grid, size_t pitch, double* out) {
  double middle[64];
  for (int n = 0; n < 8; ++n) forward8(grid + n, pitch, middle + n, 8);
  for (int k = 0; k < 8; ++k) forward8(middle + 8 * k, 1, out + 8 * k, 1);
}

/* 64 coefficients back into a block: along the rows, then down the columns
 * (cpp/src/dct.cpp:214-220). */
static void inverse_block(const double* in, double* grid, size_t pitch) {
  double middle[64];
  for (int v = 0; v < 8; ++v) inverse8(in + 8 * v, 1, middle + 8 * v, 1);
  for (int n = 0; n < 8; ++n) inverse8(middle + n, 8, grid + n, pitch);
}

void oracle_dct_forward(const double in[64], double out[64]) {
  make_basis();
  forward_block(in, 8, out);
}

void oracle_dct_inverse(const double in[64], double out[64]) {
  make_basis();
  inverse_block(in, out, 8);
}

/* ------------------------------------------------------------------------------------
 * The Laplace model (docs/math.md:223-318; cpp/src/laplace.cpp:86-164) */

/* B_2k / (2k)!, k = 1..11, each the double nearest to the rational. */
static const double series[11] = {
    0x1.5555555555555p-4,  -0x1.6c16c16c16c17p-10, 0x1.1566abc011567p-15, -0x1.bbd779334ef0bp-21,
    0x1.66a8f2bf70ebep-26, -0x1.22805d644267fp-31, 0x1.d6db2c4e09162p-37, -0x1.7da4e1f79955cp-42,
    0x1.355871d652e9ep-47, -0x1.f57d968caacf1p-53, 0x1.967e1f09c376fp-58,
};

static double nearest(u128 value) { return (double)value; }

double oracle_laplace_scale(uint64_t zeros, uint64_t others, uint64_t sum, double step) {
  if (sum == 0) return 0.0;
  const u128 n0 = zeros, n1 = others, s = sum;
  const u128 a = n0 + s + 2u * n1;
  const u128 discriminant = n0 * n0 + 4u * a * s;
  const double root = sqrt(nearest(discriminant));
  const double n0_double = nearest(n0);
  const double t = nearest(2u * s) / (n0_double + root);
  double log_inverse;
  if (t >= 0.5) {
    const double top = nearest(8u * s * (n0 + n1));
    const double first = root + nearest(2u * s - n0);
    const double second = n0_double + root;
    const double complement = top / (first * second);
    log_inverse = -log1p(-complement);
  } else {
    log_inverse = -log(t);
  }
  return step / (2.0 * log_inverse);
}

double oracle_laplace_shrinkage(double rho) {
  if (!(rho > 0.0)) return 0.0;
  if (isinf(rho)) return 0.5;
  if (rho < 1.0) {
    const double square = rho * rho;
    double total = series[10];
    for (int k = 9; k >= 0; --k) total = total * square + series[k];
    return total * rho;
  }
  const double decay = exp(-rho);
  return (0.5 - 1.0 / rho) + decay / -expm1(-rho);
}

/* sign(q) (|q| - delta/Q) Q, 0 for q = 0 (cpp/include/unround/laplace.hpp:55-60). */
static double mmse_centre(int16_t level, double step, double shrink) {
  const double q = (double)level;
  const double size = (fabs(q) - shrink) * step;
  const double signed_size = q < 0.0 ? -size : size;
  return q == 0.0 ? 0.0 : signed_size;
}

double oracle_mmse_centre(int16_t level, double step, double scale) {
  const double rho = scale > 0.0 ? step / scale : INFINITY;
  return mmse_centre(level, step, oracle_laplace_shrinkage(rho));
}

/* ------------------------------------------------------------------------------------
 * The model of a component (docs/math.md:129-146, 377-492; cpp/src/model.cpp:126-262) */

typedef struct problem {
  size_t rows, columns; /* blocks */
  int16_t* levels;
  double* lower;
  double* upper;
  double* centres;
  double* inner_lower;
  double* inner_upper;
  uint8_t* flat;
  double steps[64];
  double weights[64];
  double flat_weights[64];
  double costs[64];
  int costed;
} problem;

static void problem_free(problem* p) {
  free(p->levels);
  free(p->lower);
  free(p->upper);
  free(p->centres);
  free(p->inner_lower);
  free(p->inner_upper);
  free(p->flat);
  memset(p, 0, sizeof *p);
}

static int at_least_zero(double value) { return value >= 0.0 && isfinite(value); }
static int positive(double value) { return value > 0.0 && isfinite(value); }

static double power_of(double step, double power) {
  if (power == 2.0) return step * step;
  if (power == 1.0) return step;
  return pow(step, power);
}

/* ((q + half) + widen) Q + shift, the product and the sum apart (model.cpp:44-48). */
static double interval_end(double level, double half, double widen, double step, double shift) {
  volatile double widened = (level + half) + widen;
  volatile double scaled = widened * step;
  return scaled + shift;
}

static double rule_mu(const oracle_data* data, const uint16_t table[64]) {
  if (data->has_mu) return data->mu;
  uint32_t sum = 0;
  for (int k = 0; k < 64; ++k) sum += table[k];
  const double mean = (double)sum / 64.0;
  if (data->mu_power == 1.0) return data->mu_scale * mean;
  return data->mu_scale * pow(mean, data->mu_power);
}

static int problem_make(problem* pr, const int16_t* levels, size_t rows, size_t columns, const uint16_t table[64],
                        const oracle_data* data, char* message, size_t size) {
  memset(pr, 0, sizeof *pr);
  if (!(at_least_zero(data->mu_scale) && at_least_zero(data->mu_chroma) && isfinite(data->mu_power))) {
    say(message, size, "the scale of mu and its factor in the chroma are at least 0 and finite");
    return 0;
  }
  const double mu = rule_mu(data, table);
  if (!(at_least_zero(mu) && at_least_zero(data->slack) && at_least_zero(data->dc_weight)) ||
      !at_least_zero(data->power) || !at_least_zero(data->slack_cost) || !at_least_zero(data->flat_weight) ||
      data->flat_levels > 63u) {
    say(message, size, "an option of the data term is out of its range");
    return 0;
  }
  for (int k = 0; k < 64; ++k) {
    if (table[k] == 0) {
      say(message, size, "a quantization step is at least 1");
      return 0;
    }
  }
  const size_t blocks = rows * columns;
  pr->rows = rows;
  pr->columns = columns;
  pr->levels = (int16_t*)malloc(sizeof(int16_t) * 64u * (blocks ? blocks : 1));
  memcpy(pr->levels, levels, sizeof(int16_t) * 64u * blocks);
  for (int k = 0; k < 64; ++k) {
    pr->steps[k] = (double)table[k];
    pr->weights[k] = mu / power_of(pr->steps[k], data->power);
  }
  pr->weights[0] *= data->dc_weight;
  for (int k = 0; k < 64; ++k) pr->flat_weights[k] = pr->weights[k] * data->flat_weight;
  /* Flat blocks (model.cpp:208-226). */
  pr->flat = (uint8_t*)malloc(blocks ? blocks : 1);
  for (size_t by = 0; by < rows; ++by) {
    for (size_t bx = 0; bx < columns; ++bx) {
      const int16_t* q = levels + 64u * (by * columns + bx);
      unsigned nonzero = 0;
      for (int k = 1; k < 64; ++k) nonzero += q[k] != 0;
      int flat = nonzero <= data->flat_levels;
      if (data->has_flat_dc) {
        const int32_t own = q[0];
#define DC_AT(y, x) ((int32_t)levels[64u * ((y) * columns + (x))])
#define WITHIN(y, x) ((uint32_t)(own > DC_AT(y, x) ? own - DC_AT(y, x) : DC_AT(y, x) - own) <= data->flat_dc)
        flat = flat && (by == 0 || WITHIN(by - 1, bx)) && (by + 1 == rows || WITHIN(by + 1, bx)) &&
               (bx == 0 || WITHIN(by, bx - 1)) && (bx + 1 == columns || WITHIN(by, bx + 1));
#undef WITHIN
#undef DC_AT
      }
      pr->flat[by * columns + bx] = (uint8_t)(flat ? 1 : 0);
    }
  }
  double shrink[64] = {0};
  if (data->centres == 0) {
    /* Counts of every frequency, then the scale and the shrinkage (laplace.cpp:123-160). */
    uint64_t zeros[64] = {0}, others[64] = {0}, sums[64] = {0};
    for (size_t b = 0; b < blocks; ++b) {
      for (int k = 0; k < 64; ++k) {
        const int64_t level = levels[64u * b + k];
        const uint64_t zero = level == 0 ? 1u : 0u;
        const int64_t magnitude = level < 0 ? -level : level;
        zeros[k] += zero;
        others[k] += 1u - zero;
        sums[k] += zero ? 0u : (uint64_t)(2 * magnitude - 1);
      }
    }
    for (int k = 1; k < 64; ++k) {
      const double scale = oracle_laplace_scale(zeros[k], others[k], sums[k], (double)table[k]);
      const double rho = scale > 0.0 ? (double)table[k] / scale : INFINITY;
      shrink[k] = oracle_laplace_shrinkage(rho);
    }
  }
  const size_t entries = 64u * (blocks ? blocks : 1);
  pr->lower = (double*)malloc(sizeof(double) * entries);
  pr->upper = (double*)malloc(sizeof(double) * entries);
  pr->centres = (double*)malloc(sizeof(double) * entries);
  pr->costed = data->slack > 0.0 && data->slack_cost > 0.0;
  if (pr->costed) {
    pr->inner_lower = (double*)malloc(sizeof(double) * entries);
    pr->inner_upper = (double*)malloc(sizeof(double) * entries);
    for (int k = 0; k < 64; ++k) pr->costs[k] = data->slack_cost / pr->steps[k];
  }
  for (size_t b = 0; b < blocks; ++b) {
    const int16_t* q = levels + 64u * b;
    for (int k = 0; k < 64; ++k) {
      const double level = (double)q[k];
      const double shift = k == 0 ? 1024.0 : 0.0;
      pr->lower[64u * b + k] = interval_end(level, -0.5, -data->slack, pr->steps[k], shift);
      pr->upper[64u * b + k] = interval_end(level, 0.5, data->slack, pr->steps[k], shift);
      pr->centres[64u * b + k] = level * pr->steps[k] + shift;
      if (k > 0 && data->centres == 0) pr->centres[64u * b + k] = mmse_centre(q[k], pr->steps[k], shrink[k]);
      if (pr->costed) {
        pr->inner_lower[64u * b + k] = interval_end(level, -0.5, 0.0, pr->steps[k], shift);
        pr->inner_upper[64u * b + k] = interval_end(level, 0.5, 0.0, pr->steps[k], shift);
      }
    }
  }
  return 1;
}

/* The proximal map of one coefficient (model.cpp:62-75, 346-398). */
static double prox_one(const problem* pr, size_t b, int k, double e, double step) {
  const double* weights = pr->flat[b] ? pr->flat_weights : pr->weights;
  const double t = step * weights[k];
  const double r = 1.0 / (1.0 + t);
  const double centre = pr->centres[64u * b + k];
  const double low = pr->lower[64u * b + k];
  const double high = pr->upper[64u * b + k];
  if (!pr->costed) return clamp_to((e + t * centre) * r, low, high);
  const double h = step * pr->costs[k] * r;
  const double own_low = pr->inner_lower[64u * b + k];
  const double own_high = pr->inner_upper[64u * b + k];
  const double z = (e + t * centre) * r;
  const double above = own_high > z - h ? own_high : z - h;
  const double below = own_low < z + h ? own_low : z + h;
  const double inside = z < own_low ? below : z;
  return clamp_to(z > own_high ? above : inside, low, high);
}

static double beyond_of(double c, double own_low, double own_high) {
  double most = c - own_high;
  if (own_low - c > most) most = own_low - c;
  return most > 0.0 ? most : 0.0;
}

/* g*(s) of one coefficient (model.cpp:80-109). */
static double conjugate_one(const problem* pr, size_t b, int k, double s) {
  const double m = (pr->flat[b] ? pr->flat_weights : pr->weights)[k];
  const double centre = pr->centres[64u * b + k];
  const double low = pr->lower[64u * b + k];
  const double high = pr->upper[64u * b + k];
  if (!pr->costed) {
    const double best = clamp_to(centre + s / m, low, high);
    const double away = best - centre;
    const double weighted = s * best - 0.5 * m * (away * away);
    const double a = s * low, c = s * high;
    const double support = a < c ? c : a;
    return m > 0.0 ? weighted : support;
  }
  const double lambda = pr->costs[k];
  const double own_low = pr->inner_lower[64u * b + k];
  const double own_high = pr->inner_upper[64u * b + k];
  const double free_at = centre + s / m;
  const double upward = clamp_to(centre + (s - lambda) / m, own_high, high);
  const double downward = clamp_to(centre + (s + lambda) / m, low, own_low);
  const double within = free_at < own_low ? downward : free_at;
  const double weighted = free_at > own_high ? upward : within;
  const double lower_end = s < -lambda ? low : own_low;
  const double upper_end = s > lambda ? high : own_high;
  const double flat = s > 0.0 ? upper_end : lower_end;
  const double best = m > 0.0 ? weighted : flat;
  const double away = best - centre;
  return s * best - 0.5 * m * (away * away) - lambda * beyond_of(best, own_low, own_high);
}

/* G's term of one coefficient within its interval (model.cpp:411-436). */
static double value_one(const problem* pr, size_t b, int k, double c) {
  const double* weights = pr->flat[b] ? pr->flat_weights : pr->weights;
  const double away = c - pr->centres[64u * b + k];
  double term = 0.5 * weights[k] * (away * away);
  if (pr->costed) term = term + pr->costs[k] * beyond_of(c, pr->inner_lower[64u * b + k], pr->inner_upper[64u * b + k]);
  return term;
}

/* The terms of a component's blocks summed as block rows of 64 x columns terms
 * (model.cpp:438-440). */
static double blocks_sum(const problem* pr, const double* terms) {
  return oracle_sum_rows(terms, pr->rows, pr->columns * 64u);
}

/* ------------------------------------------------------------------------------------
 * The frame: components on one canvas (docs/math.md:162-221, 510-542, 911-949;
 * cpp/src/frames.cpp:343-675) */

typedef struct component {
  problem pr;
  size_t down, across;
} component;

typedef struct frame {
  size_t rows, columns, channels;
  component comps[3];
} frame;

static size_t ceiling(size_t n, size_t d) { return d == 0 ? 0 : (n + d - 1u) / d; }

static int frame_free(const frame* f, size_t c) {
  const component* k = &f->comps[c];
  return k->down * k->across > 1u || 8u * k->pr.rows * k->down < f->rows ||
         8u * k->pr.columns * k->across < f->columns;
}

/* dst = src, count doubles, in bands over the threads. */
static void copy_doubles(double* dst, const double* src, size_t count) {
  const size_t band = 1u << 16;
  const long bands = (long)((count + band - 1) / band);
#pragma omp parallel for schedule(static)
  for (long b = 0; b < bands; ++b) {
    const size_t at = (size_t)b * band;
    memcpy(dst + at, src + at, sizeof(double) * (count - at < band ? count - at : band));
  }
}

/* The cells of the first 8 by x 8 bx samples summed (or their means): rows of a cell
 * added down first, then the samples across (frames.cpp:244-339). */
static void cells_of(const frame* f, size_t c, const double* x, double* out, int divide) {
  const component* k = &f->comps[c];
  const size_t rows = 8u * k->pr.rows, columns = 8u * k->pr.columns;
  const size_t span = columns * k->across;
  const double cells = (double)(k->down * k->across);
#pragma omp parallel
  {
    double* sums = (double*)malloc(sizeof(double) * (span ? span : 1));
#pragma omp for schedule(static)
    for (size_t i = 0; i < rows; ++i) {
      memcpy(sums, x + (k->down * i) * f->columns, sizeof(double) * span);
      for (size_t a = 1; a < k->down; ++a) {
        const double* next = x + (k->down * i + a) * f->columns;
        for (size_t j = 0; j < span; ++j) sums[j] = sums[j] + next[j];
      }
      for (size_t j = 0; j < columns; ++j) {
        double sum = sums[k->across * j];
        for (size_t b = 1; b < k->across; ++b) sum += sums[k->across * j + b];
        out[i * columns + j] = divide ? sum / cells : sum;
      }
    }
    free(sums);
  }
}

/* The coefficients of a grid of 8 rows x 8 columns blocks, pitch `pitch`. */
static void forward_grid(const double* grid, size_t pitch, size_t brows, size_t bcols, double* coefficients) {
#pragma omp parallel for schedule(static)
  for (size_t by = 0; by < brows; ++by)
    for (size_t bx = 0; bx < bcols; ++bx)
      forward_block(grid + 8u * by * pitch + 8u * bx, pitch, coefficients + 64u * (by * bcols + bx));
}

static void inverse_grid(const double* coefficients, size_t brows, size_t bcols, double* grid, size_t pitch) {
#pragma omp parallel for schedule(static)
  for (size_t by = 0; by < brows; ++by)
    for (size_t bx = 0; bx < bcols; ++bx)
      inverse_block(coefficients + 64u * (by * bcols + bx), grid + 8u * by * pitch + 8u * bx, pitch);
}

/* The proximal map of tau G_c (frames.cpp:490-548): zeta into kept, the canvas into out. */
static void frame_prox(const frame* f, size_t c, const double* v, double tau, double* out, double* kept) {
  const component* k = &f->comps[c];
  const problem* pr = &k->pr;
  const size_t blocks = pr->rows * pr->columns;
  const double step = tau / (double)(k->down * k->across);
  double* at = (double*)malloc(sizeof(double) * 64u * (blocks ? blocks : 1));
  if (k->down * k->across == 1u) {
    forward_grid(v, f->columns, pr->rows, pr->columns, at);
#pragma omp parallel for schedule(static)
    for (size_t b = 0; b < blocks; ++b)
      for (int q = 0; q < 64; ++q) kept[64u * b + q] = prox_one(pr, b, q, at[64u * b + q], step);
    copy_doubles(out, v, f->rows * f->columns);
    inverse_grid(kept, pr->rows, pr->columns, out, f->columns);
    free(at);
    return;
  }
  const size_t crows = 8u * pr->rows, ccols = 8u * pr->columns;
  double* means = (double*)malloc(sizeof(double) * crows * ccols);
  double* change = (double*)malloc(sizeof(double) * crows * ccols);
  cells_of(f, c, v, means, 1);
  forward_grid(means, ccols, pr->rows, pr->columns, at);
#pragma omp parallel for schedule(static)
  for (size_t b = 0; b < blocks; ++b)
    for (int q = 0; q < 64; ++q) kept[64u * b + q] = prox_one(pr, b, q, at[64u * b + q], step);
  inverse_grid(kept, pr->rows, pr->columns, change, ccols);
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < crows * ccols; ++e) change[e] -= means[e];
  /* out = v + the change spread over each cell, v beyond the blocks (frames.cpp:445-476). */
  copy_doubles(out, v, f->rows * f->columns);
  const size_t span = ccols * k->across;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < crows; ++i) {
    for (size_t a = 0; a < k->down; ++a) {
      const size_t row = k->down * i + a;
      for (size_t j = 0; j < span; ++j) out[row * f->columns + j] = v[row * f->columns + j] + change[i * ccols + j / k->across];
    }
  }
  free(means);
  free(change);
  free(at);
}

/* The canvas of coefficients, `beyond` outside the blocks (frames.cpp:569-587). */
static void frame_canvas(const frame* f, size_t c, const double* coefficients, double beyond, double* out) {
  const component* k = &f->comps[c];
  for (size_t e = 0; e < f->rows * f->columns; ++e) out[e] = beyond;
  if (k->down * k->across == 1u) {
    inverse_grid(coefficients, k->pr.rows, k->pr.columns, out, f->columns);
    return;
  }
  const size_t crows = 8u * k->pr.rows, ccols = 8u * k->pr.columns;
  double* samples = (double*)malloc(sizeof(double) * crows * ccols);
  inverse_grid(coefficients, k->pr.rows, k->pr.columns, samples, ccols);
  for (size_t i = 0; i < crows; ++i)
    for (size_t a = 0; a < k->down; ++a)
      for (size_t j = 0; j < ccols * k->across; ++j) out[(k->down * i + a) * f->columns + j] = samples[i * ccols + j / k->across];
  free(samples);
}

/* G_c* at the field xi, the free samples in the box of radius R (frames.cpp:595-681). */
static double frame_conjugate(const frame* f, size_t c, const double* xi, double radius) {
  const component* k = &f->comps[c];
  const problem* pr = &k->pr;
  const size_t blocks = pr->rows * pr->columns;
  const size_t entries = 64u * (blocks ? blocks : 1);
  double* s = (double*)malloc(sizeof(double) * entries);
  double* terms = (double*)malloc(sizeof(double) * entries);
  const size_t crows = 8u * pr->rows, ccols = 8u * pr->columns;
  if (k->down * k->across == 1u) {
    forward_grid(xi, f->columns, pr->rows, pr->columns, s);
  } else {
    double* sums = (double*)malloc(sizeof(double) * crows * ccols);
    cells_of(f, c, xi, sums, 0);
    forward_grid(sums, ccols, pr->rows, pr->columns, s);
    free(sums);
  }
  for (size_t b = 0; b < blocks; ++b)
    for (int q = 0; q < 64; ++q) terms[64u * b + q] = conjugate_one(pr, b, q, s[64u * b + q]);
  double value = blocks_sum(pr, terms);
  free(s);
  free(terms);
  if (!frame_free(f, c)) return value;
  double* means = (double*)malloc(sizeof(double) * crows * ccols);
  double* samples = (double*)malloc(sizeof(double) * f->rows * f->columns);
  cells_of(f, c, xi, means, 1);
  const size_t span = ccols * k->across;
  for (size_t i = 0; i < f->rows; ++i) {
    const int within = i < k->down * crows;
    for (size_t j = 0; j < f->columns; ++j) {
      const double sample = xi[i * f->columns + j];
      double term;
      if (within && j < span) {
        term = radius * fabs(sample - means[(i / k->down) * ccols + j / k->across]);
      } else {
        term = 128.0 * sample + radius * fabs(sample);
      }
      samples[i * f->columns + j] = term;
    }
  }
  value = value + oracle_sum_rows(samples, f->rows, f->columns);
  free(means);
  free(samples);
  return value;
}

/* ------------------------------------------------------------------------------------
 * Differences (docs/math.md:320-371; cpp/src/operators.cpp:49-198) */

/* Forward differences, 0 out of the grid. */
static void gradient(const double* f, size_t H, size_t W, double* gx, double* gy) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < H; ++i) {
    for (size_t j = 0; j < W; ++j) {
      const double here = f[i * W + j];
      gx[i * W + j] = j + 1 < W ? f[i * W + j + 1] - here : 0.0;
      gy[i * W + j] = i + 1 < H ? f[(i + 1) * W + j] - here : 0.0;
    }
  }
}

/* delta_x: f[j] (not in the last column) - f[j-1] (not in the first). */
static double backward_across(const double* f, size_t W, size_t i, size_t j) {
  if (W == 1) return 0.0;
  if (j == 0) return f[i * W];
  if (j + 1 == W) return -f[i * W + j - 1];
  return f[i * W + j] - f[i * W + j - 1];
}

static double backward_down(const double* f, size_t H, size_t W, size_t i, size_t j) {
  const int first = i == 0, last = i + 1 == H;
  if (!first && !last) return f[i * W + j] - f[(i - 1) * W + j];
  if (!last) return f[i * W + j];
  if (!first) return -f[(i - 1) * W + j];
  return 0.0;
}

static double forward_across(const double* f, size_t W, size_t i, size_t j) {
  return j + 1 < W ? f[i * W + j + 1] - f[i * W + j] : 0.0;
}

static double forward_down(const double* f, size_t H, size_t W, size_t i, size_t j) {
  return i + 1 < H ? f[(i + 1) * W + j] - f[i * W + j] : 0.0;
}

/* div p = delta_y p_y, then + delta_x p_x. */
static void divergence(const double* px, const double* py, size_t H, size_t W, double* out) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < H; ++i)
    for (size_t j = 0; j < W; ++j) {
      double value = backward_down(py, H, W, i, j);
      value += backward_across(px, W, i, j);
      out[i * W + j] = value;
    }
}

/* E w = (delta_x w_x, delta_y w_y, (delta_y w_x + delta_x w_y) / 2). */
static void symmetrized_gradient(const double* wx, const double* wy, size_t H, size_t W, double* xx, double* yy,
                                 double* xy) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < H; ++i)
    for (size_t j = 0; j < W; ++j) {
      xx[i * W + j] = backward_across(wx, W, i, j);
      yy[i * W + j] = backward_down(wy, H, W, i, j);
      double sum = backward_down(wx, H, W, i, j);
      sum += backward_across(wy, W, i, j);
      xy[i * W + j] = 0.5 * sum;
    }
}

/* div_2 r = (d_x r_xx + d_y r_xy, d_x r_xy + d_y r_yy). */
static void tensor_divergence(const double* rxx, const double* ryy, const double* rxy, size_t H, size_t W, double* x,
                              double* y) {
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < H; ++i)
    for (size_t j = 0; j < W; ++j) {
      double a = forward_across(rxx, W, i, j);
      a += forward_down(rxy, H, W, i, j);
      double b = forward_across(rxy, W, i, j);
      b += forward_down(ryy, H, W, i, j);
      x[i * W + j] = a;
      y[i * W + j] = b;
    }
}

/* radius / |p| beyond the ball, 1 within (operators.hpp:112-115). */
static double ball_scale(double square, double radius) {
  const double norm = sqrt(square);
  return norm > radius ? radius / norm : 1.0;
}

/* ------------------------------------------------------------------------------------
 * The result */

struct oracle_result {
  size_t channels, rows, columns;
  double* canvas;
  double* iterate;
  size_t block_rows[3], block_columns[3], down[3], across[3];
  double* arrays[3][8];
  uint64_t counts[3][8];
  double tau, sigma;
  uint64_t iterations;
  int32_t stop;
  double iteration_seconds, least_iteration_seconds;
  uint64_t records;
  uint64_t* record_iterations;
  double* record_values[4];
};

void oracle_result_free(oracle_result* result) {
  if (result == NULL) return;
  free(result->canvas);
  free(result->iterate);
  for (int c = 0; c < 3; ++c)
    for (int k = 0; k < 8; ++k) free(result->arrays[c][k]);
  free(result->record_iterations);
  for (int k = 0; k < 4; ++k) free(result->record_values[k]);
  free(result);
}

const double* oracle_result_canvas(const oracle_result* r, uint64_t* channels, uint64_t* rows, uint64_t* columns) {
  if (channels) *channels = r->channels;
  if (rows) *rows = r->rows;
  if (columns) *columns = r->columns;
  return r->canvas;
}

const double* oracle_result_iterate(const oracle_result* r) { return r->iterate; }

void oracle_result_component(const oracle_result* r, uint32_t c, uint64_t* block_rows, uint64_t* block_columns,
                             uint64_t* down, uint64_t* across) {
  if (block_rows) *block_rows = r->block_rows[c];
  if (block_columns) *block_columns = r->block_columns[c];
  if (down) *down = r->down[c];
  if (across) *across = r->across[c];
}

const double* oracle_result_array(const oracle_result* r, uint32_t c, int32_t which, uint64_t* count) {
  if (c >= r->channels || which < 0 || which > 7) return NULL;
  if (count) *count = r->counts[c][which];
  return r->arrays[c][which];
}

double oracle_result_iteration_seconds(const oracle_result* r) { return r->iteration_seconds; }

double oracle_result_least_iteration_seconds(const oracle_result* r) { return r->least_iteration_seconds; }

void oracle_set_threads(int32_t count) { omp_set_num_threads(count > 0 ? count : omp_get_num_procs()); }

int32_t oracle_threads(void) { return omp_get_max_threads(); }

void oracle_result_solver(const oracle_result* r, double* tau, double* sigma, uint64_t* iterations, int32_t* stop) {
  if (tau) *tau = r->tau;
  if (sigma) *sigma = r->sigma;
  if (iterations) *iterations = r->iterations;
  if (stop) *stop = r->stop;
}

const uint64_t* oracle_result_record_iterations(const oracle_result* r, uint64_t* count) {
  if (count) *count = r->records;
  return r->record_iterations;
}

const double* oracle_result_record_values(const oracle_result* r, int32_t which, uint64_t* count) {
  if (which < 0 || which > 3) return NULL;
  if (count) *count = r->records;
  return r->record_values[which];
}

static double* copy_of(const double* from, size_t count) {
  double* to = (double*)malloc(sizeof(double) * (count ? count : 1));
  if (count) memcpy(to, from, sizeof(double) * count);
  return to;
}

static void push_record(oracle_result* r, uint64_t iteration, double primal, double dual, double theta,
                        double partial) {
  const uint64_t n = r->records + 1;
  r->record_iterations = (uint64_t*)realloc(r->record_iterations, sizeof(uint64_t) * n);
  for (int k = 0; k < 4; ++k) r->record_values[k] = (double*)realloc(r->record_values[k], sizeof(double) * n);
  r->record_iterations[n - 1] = iteration;
  r->record_values[0][n - 1] = primal;
  r->record_values[1][n - 1] = dual;
  r->record_values[2][n - 1] = theta;
  r->record_values[3][n - 1] = partial;
  r->records = n;
}

/* ------------------------------------------------------------------------------------
 * The primal-dual method (docs/math.md:568-682, 742-811;
 * cpp/src/pdhg.cpp:147-233, cpp/src/primal_dual.hpp:401-1013) */

/* ---------------------------------------------------------------------------------
 * Pictures of rounded samples (docs/math.md, 11; rust/unround/src/samples.rs:69-345). */

typedef struct samples_model {
  size_t C, H, W;
  double *low, *high, *centre; /* C x H x W */
  double mu;
} samples_model;

/* A double as mantissa / 2^power, exactly (samples.rs:286-300): 0, or a normal double. */
static void dyadic_of(double value, __int128* mantissa, unsigned* power) {
  uint64_t bits;
  memcpy(&bits, &value, sizeof bits);
  const unsigned exponent = (unsigned)((bits >> 52) & 0x7FFu);
  if (exponent == 0) {
    *mantissa = 0;
    *power = 0;
    return;
  }
  const __int128 m = (__int128)((bits & ((1ull << 52) - 1u)) | (1ull << 52));
  *power = 1075u - exponent;
  *mantissa = value < 0.0 ? -m : m;
}

/* Whether value < numerator / denominator exactly (samples.rs:302-307). */
static int lies_below(double value, __int128 numerator, __int128 denominator) {
  __int128 mantissa;
  unsigned power;
  dyadic_of(value, &mantissa, &power);
  return mantissa * denominator < numerator * ((__int128)1 << power);
}

/* The least double at or above numerator / denominator, and the greatest double below it
 * (samples.rs:309-330); the nearest is the correctly rounded quotient of two doubles. */
static double at_or_above(__int128 numerator, __int128 denominator) {
  const double near = (double)numerator / (double)denominator;
  return lies_below(near, numerator, denominator) ? nextafter(near, INFINITY) : near;
}
static double under(__int128 numerator, __int128 denominator) {
  const double near = (double)numerator / (double)denominator;
  return lies_below(near, numerator, denominator) ? near : nextafter(near, -INFINITY);
}

/* The ends and centre of every level 0..top (samples.rs:252-284): without slack the
 * doubles within max(0, (v - 1/2) s) <= x < (v + 1/2) s, or <= 255 at the last level;
 * with a slack, max(0, ((v - 1/2) - slack) s) and min(255, ((v + 1/2) + slack) s) in
 * floating point, s the double nearest 255 / top; the centre v s, the nearest double. */
static void level_ends(uint32_t level, uint32_t top, double slack, double* low, double* high, double* centre) {
  const __int128 wide = top, value = level;
  double below, above;
  if (slack == 0.0) {
    below = at_or_above((2 * value - 1) * 255, 2 * wide);
    above = under((2 * value + 1) * 255, 2 * wide);
  } else {
    const double step = 255.0 / (double)top;
    const double own = (double)level;
    below = ((own - 0.5) - slack) * step;
    above = ((own + 0.5) + slack) * step;
  }
  *low = below > 0.0 ? below : 0.0;
  *high = above < 255.0 ? above : 255.0;
  *centre = (double)(value * 255) / (double)wide;
}

int oracle_samples_intervals(const uint16_t* levels, uint64_t count, uint32_t bits, double slack, double* low,
                             double* high, double* centre) {
  if (!(bits == 1 || bits == 2 || bits == 4 || bits == 8 || bits == 16) || !(slack >= 0.0 && isfinite(slack)))
    return 0;
  const uint32_t top = (1u << bits) - 1u;
  for (uint64_t k = 0; k < count; ++k) {
    if (levels[k] > top) return 0;
    level_ends(levels[k], top, slack, &low[k], &high[k], &centre[k]);
  }
  return 1;
}

static int samples_make(samples_model* m, const uint16_t* levels, size_t C, size_t H, size_t W, uint32_t bits,
                        double mu, double slack) {
  const size_t count = C * H * W;
  m->C = C;
  m->H = H;
  m->W = W;
  m->mu = mu;
  m->low = (double*)malloc(sizeof(double) * (count ? count : 1));
  m->high = (double*)malloc(sizeof(double) * (count ? count : 1));
  m->centre = (double*)malloc(sizeof(double) * (count ? count : 1));
  if (count == 0 || !(mu >= 0.0 && isfinite(mu)) ||
      !oracle_samples_intervals(levels, count, bits, slack, m->low, m->high, m->centre)) {
    free(m->low);
    free(m->high);
    free(m->centre);
    return 0;
  }
  return 1;
}

static void samples_free(samples_model* m) {
  free(m->low);
  free(m->high);
  free(m->centre);
}

/* prox_{tau G}(v) of channel c's samples: clip((t c + v) r, a, b), t = tau mu, r = 1/(1+t)
 * (samples.rs:171-182; t c + v fused, as the reference's multiply_add is with FMA). */
static void samples_prox(const samples_model* m, size_t c, const double* v, double tau, double* out) {
  const size_t N = m->H * m->W, at = c * N;
  const double t = tau * m->mu;
  const double r = 1.0 / (1.0 + t);
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < N; ++e) out[e] = clamp_to(fma(t, m->centre[at + e], v[e]) * r, m->low[at + e], m->high[at + e]);
}

/* G(x) = mu / 2 the sum of the squares from the centres, row by row (samples.rs:205-222). */
static double samples_value(const samples_model* m, const double* x) {
  if (m->mu == 0.0) return 0.0;
  const size_t count = m->C * m->H * m->W;
  double* terms = (double*)malloc(sizeof(double) * count);
  for (size_t k = 0; k < count; ++k) {
    const double away = x[k] - m->centre[k];
    terms[k] = away * away;
  }
  const double total = oracle_sum_rows(terms, m->C * m->H, m->W);
  free(terms);
  return 0.5 * m->mu * total;
}

/* g*(s) of a sample (kernels.rs:1572-1590 without a charge). */
static double samples_term(double s, double mu, double centre, double low, double high) {
  if (mu > 0.0) {
    const double best = clamp_to(centre + s / mu, low, high);
    const double difference = best - centre;
    return s * best - 0.5 * mu * (difference * difference);
  }
  return fmax(s * low, s * high);
}

/* G*_c at xi (one channel's plane), row by row (samples.rs:224-248). */
static double samples_conjugate(const samples_model* m, size_t c, const double* xi) {
  const size_t N = m->H * m->W, at = c * N;
  double* terms = (double*)malloc(sizeof(double) * N);
  for (size_t e = 0; e < N; ++e) terms[e] = samples_term(xi[e], m->mu, m->centre[at + e], m->low[at + e], m->high[at + e]);
  const double total = oracle_sum_rows(terms, m->H, m->W);
  free(terms);
  return total;
}

void oracle_samples_prox(const uint16_t* levels, uint64_t count, uint32_t bits, double mu, double slack,
                         const double* values, double tau, double* out) {
  samples_model m;
  if (!samples_make(&m, levels, 1, 1, count, bits, mu, slack)) return;
  samples_prox(&m, 0, values, tau, out);
  samples_free(&m);
}

double oracle_samples_value(const uint16_t* levels, uint64_t channels, uint64_t height, uint64_t width,
                            uint32_t bits, double mu, double slack, const double* x) {
  samples_model m;
  if (!samples_make(&m, levels, channels, height, width, bits, mu, slack)) return NAN;
  const double value = samples_value(&m, x);
  samples_free(&m);
  return value;
}

double oracle_samples_conjugate(const uint16_t* levels, uint64_t channels, uint64_t height, uint64_t width,
                                uint32_t bits, double mu, double slack, const double* xi) {
  samples_model m;
  if (!samples_make(&m, levels, channels, height, width, bits, mu, slack)) return NAN;
  double total = 0.0;
  for (size_t c = 0; c < channels; ++c) total += samples_conjugate(&m, c, xi + c * height * width);
  samples_free(&m);
  return total;
}


typedef struct machine {
  const frame* f;
  const samples_model* sm; /* in place of f for a picture of rounded samples */
  int tgv, coupled;
  size_t C, H, W, N;
  double tau, sigma, rho, rest;
  double alpha, alpha1, alpha0;
  double gamma[3];
  double free_radius;
  int has_partial;
  double partial_radius;
  /* state and tilde point, [c][entry][N] */
  double *x, *p, *w, *r;
  double *tx, *tp, *tw, *tr;
  double* zeta[3];
  /* The iteration's work fields, made at its first iteration and kept (v, three
   * differences, x-, w- and the steps). */
  double* work;
  /* Wall seconds inside the iterations (records and set-up excluded), and the least of
   * one iteration. */
  double iteration_seconds, least_iteration_seconds;
} machine;

#define FIELD(base, c, e, m) ((base) + ((c) * (m) + (e)) * N)

/* Projects the duals of every channel onto their balls (operators.cpp:378-425). */
static void project_duals(machine* M, double* p, double* r) {
  const size_t C = M->C, N = M->N;
  const double radius = M->tgv ? M->alpha1 : M->alpha;
  const int coupled = M->coupled;
  for (size_t e = 0; e < N; ++e) {
    if (coupled) {
      double square = 0.0;
      for (size_t c = 0; c < C; ++c) square += FIELD(p, c, 0, 2)[e] * FIELD(p, c, 0, 2)[e] + FIELD(p, c, 1, 2)[e] * FIELD(p, c, 1, 2)[e];
      const double scale = ball_scale(square, radius);
      for (size_t c = 0; c < C; ++c) {
        FIELD(p, c, 0, 2)[e] = FIELD(p, c, 0, 2)[e] * scale;
        FIELD(p, c, 1, 2)[e] = FIELD(p, c, 1, 2)[e] * scale;
      }
      if (M->tgv) {
        double tsq = 0.0;
        for (size_t c = 0; c < C; ++c) {
          const double a = FIELD(r, c, 0, 3)[e], b = FIELD(r, c, 1, 3)[e], d = FIELD(r, c, 2, 3)[e];
          tsq += a * a + b * b + 2.0 * (d * d);
        }
        const double tscale = ball_scale(tsq, M->alpha0);
        for (size_t c = 0; c < C; ++c)
          for (int k = 0; k < 3; ++k) FIELD(r, c, k, 3)[e] = FIELD(r, c, k, 3)[e] * tscale;
      }
    } else {
      for (size_t c = 0; c < C; ++c) {
        double square = 0.0;
        square += FIELD(p, c, 0, 2)[e] * FIELD(p, c, 0, 2)[e] + FIELD(p, c, 1, 2)[e] * FIELD(p, c, 1, 2)[e];
        const double scale = ball_scale(square, radius);
        FIELD(p, c, 0, 2)[e] = FIELD(p, c, 0, 2)[e] * scale;
        FIELD(p, c, 1, 2)[e] = FIELD(p, c, 1, 2)[e] * scale;
        if (M->tgv) {
          const double a = FIELD(r, c, 0, 3)[e], b = FIELD(r, c, 1, 3)[e], d = FIELD(r, c, 2, 3)[e];
          double tsq = 0.0;
          tsq += a * a + b * b + 2.0 * (d * d);
          const double tscale = ball_scale(tsq, M->alpha0);
          for (int k = 0; k < 3; ++k) FIELD(r, c, k, 3)[e] = FIELD(r, c, k, 3)[e] * tscale;
        }
      }
    }
  }
}

/* One iteration (primal_dual.hpp:525-735). */
static void iterate(machine* M) {
  const size_t C = M->C, H = M->H, W = M->W, N = M->N;
  if (M->work == NULL) M->work = (double*)malloc(sizeof(double) * (4 + 8 * C) * N);
  double* v = M->work;
  double* one = v + N;
  double* two = one + N;
  double* three = two + N;
  double* bar = three + N;
  double* barw = bar + C * N;
  double* steps = barw + 2 * C * N;
  for (size_t c = 0; c < C; ++c) {
    const double descent = M->tau * M->gamma[c];
    divergence(FIELD(M->p, c, 0, 2), FIELD(M->p, c, 1, 2), H, W, one);
    const double* x = FIELD(M->x, c, 0, 1);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) v[e] = x[e] + descent * one[e];
    if (M->sm != NULL) {
      samples_prox(M->sm, c, v, M->tau, FIELD(M->tx, c, 0, 1));
    } else {
      frame_prox(M->f, c, v, M->tau, FIELD(M->tx, c, 0, 1), M->zeta[c]);
    }
    if (M->tgv) {
      tensor_divergence(FIELD(M->r, c, 0, 3), FIELD(M->r, c, 1, 3), FIELD(M->r, c, 2, 3), H, W, one, two);
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < N; ++e) {
        FIELD(M->tw, c, 0, 2)[e] = FIELD(M->w, c, 0, 2)[e] + descent * (FIELD(M->p, c, 0, 2)[e] + one[e]);
        FIELD(M->tw, c, 1, 2)[e] = FIELD(M->w, c, 1, 2)[e] + descent * (FIELD(M->p, c, 1, 2)[e] + two[e]);
      }
    }
  }
  /* The duals: extrapolations, steps, projections, relaxation. */
  for (size_t c = 0; c < C; ++c) {
    const double ascent = M->sigma * M->gamma[c];
    double* x = FIELD(M->x, c, 0, 1);
    const double* tx = FIELD(M->tx, c, 0, 1);
    double* b = bar + c * N;
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) b[e] = 2.0 * tx[e] - x[e];
    gradient(b, H, W, one, two);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) x[e] = M->rho * tx[e] + M->rest * x[e];
    double* sx = steps + (5 * c + 0) * N;
    double* sy = steps + (5 * c + 1) * N;
    double* px = FIELD(M->p, c, 0, 2);
    double* py = FIELD(M->p, c, 1, 2);
    if (!M->tgv) {
#pragma omp parallel for schedule(static)
      for (size_t e = 0; e < N; ++e) {
        sx[e] = px[e] + ascent * one[e];
        sy[e] = py[e] + ascent * two[e];
      }
      continue;
    }
    double* bwx = barw + (2 * c + 0) * N;
    double* bwy = barw + (2 * c + 1) * N;
    double* wx = FIELD(M->w, c, 0, 2);
    double* wy = FIELD(M->w, c, 1, 2);
    const double* twx = FIELD(M->tw, c, 0, 2);
    const double* twy = FIELD(M->tw, c, 1, 2);
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) {
      bwx[e] = 2.0 * twx[e] - wx[e];
      bwy[e] = 2.0 * twy[e] - wy[e];
      wx[e] = M->rho * twx[e] + M->rest * wx[e];
      wy[e] = M->rho * twy[e] + M->rest * wy[e];
      sx[e] = px[e] + ascent * (one[e] - bwx[e]);
      sy[e] = py[e] + ascent * (two[e] - bwy[e]);
    }
    symmetrized_gradient(bwx, bwy, H, W, one, two, three);
    double* sxx = steps + (5 * c + 2) * N;
    double* syy = steps + (5 * c + 3) * N;
    double* sxy = steps + (5 * c + 4) * N;
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < N; ++e) {
      sxx[e] = FIELD(M->r, c, 0, 3)[e] + ascent * one[e];
      syy[e] = FIELD(M->r, c, 1, 3)[e] + ascent * two[e];
      sxy[e] = FIELD(M->r, c, 2, 3)[e] + ascent * three[e];
    }
  }
  const double radius = M->tgv ? M->alpha1 : M->alpha;
#pragma omp parallel for schedule(static)
  for (size_t e = 0; e < N; ++e) {
    double scale_all = 1.0, tscale_all = 1.0;
    if (M->coupled) {
      double square = 0.0, tsq = 0.0;
      for (size_t c = 0; c < C; ++c) {
        const double a = steps[(5 * c) * N + e], b = steps[(5 * c + 1) * N + e];
        square += a * a + b * b;
        if (M->tgv) {
          const double xx = steps[(5 * c + 2) * N + e], yy = steps[(5 * c + 3) * N + e], xy = steps[(5 * c + 4) * N + e];
          tsq += xx * xx + yy * yy + 2.0 * (xy * xy);
        }
      }
      scale_all = ball_scale(square, radius);
      if (M->tgv) tscale_all = ball_scale(tsq, M->alpha0);
    }
    for (size_t c = 0; c < C; ++c) {
      const double a = steps[(5 * c) * N + e], b = steps[(5 * c + 1) * N + e];
      double scale = scale_all;
      if (!M->coupled) {
        double square = 0.0;
        square += a * a + b * b;
        scale = ball_scale(square, radius);
      }
      const double pa = a * scale, pb = b * scale;
      FIELD(M->tp, c, 0, 2)[e] = pa;
      FIELD(M->tp, c, 1, 2)[e] = pb;
      FIELD(M->p, c, 0, 2)[e] = M->rho * pa + M->rest * FIELD(M->p, c, 0, 2)[e];
      FIELD(M->p, c, 1, 2)[e] = M->rho * pb + M->rest * FIELD(M->p, c, 1, 2)[e];
      if (!M->tgv) continue;
      const double xx = steps[(5 * c + 2) * N + e], yy = steps[(5 * c + 3) * N + e], xy = steps[(5 * c + 4) * N + e];
      double tscale = tscale_all;
      if (!M->coupled) {
        double tsq = 0.0;
        tsq += xx * xx + yy * yy + 2.0 * (xy * xy);
        tscale = ball_scale(tsq, M->alpha0);
      }
      const double q[3] = {xx * tscale, yy * tscale, xy * tscale};
      for (int k = 0; k < 3; ++k) {
        FIELD(M->tr, c, k, 3)[e] = q[k];
        FIELD(M->r, c, k, 3)[e] = M->rho * q[k] + M->rest * FIELD(M->r, c, k, 3)[e];
      }
    }
  }
}

/* The record of the tilde point (primal_dual.hpp:446-501, 791-895). */
static void record(machine* M, double* primal, double* dual, double* theta_out, double* partial_out) {
  const size_t C = M->C, H = M->H, W = M->W, N = M->N;
  const frame* f = M->f;
  double data = 0.0;
  for (size_t c = 0; c < C && M->sm == NULL; ++c) {
    const problem* pr = &f->comps[c].pr;
    const size_t blocks = pr->rows * pr->columns;
    double* terms = (double*)malloc(sizeof(double) * 64u * (blocks ? blocks : 1));
    for (size_t b = 0; b < blocks; ++b)
      for (int q = 0; q < 64; ++q) terms[64u * b + q] = value_one(pr, b, q, M->zeta[c][64u * b + q]);
    data += blocks_sum(pr, terms);
    free(terms);
  }
  if (M->sm != NULL) data = samples_value(M->sm, M->tx);
  double* first = (double*)calloc(N, sizeof(double));
  double* second = (double*)calloc(N, sizeof(double));
  double* squares = (double*)calloc(N, sizeof(double));
  double* one = (double*)malloc(sizeof(double) * N);
  double* two = (double*)malloc(sizeof(double) * N);
  double* three = (double*)malloc(sizeof(double) * N);
  double* divergences = M->tgv ? (double*)malloc(sizeof(double) * 2 * C * N) : NULL;
  double* xi = (double*)malloc(sizeof(double) * N);
  double* g1 = (double*)malloc(sizeof(double) * N);
  double* g2 = (double*)malloc(sizeof(double) * N);
  /* The norms of the first order, each channel weighted by gamma. */
  for (size_t c = 0; c < C; ++c) {
    const double gamma = M->gamma[c];
    gradient(FIELD(M->tx, c, 0, 1), H, W, one, two);
    if (M->tgv)
      for (size_t e = 0; e < N; ++e) {
        one[e] = one[e] + -1.0 * FIELD(M->tw, c, 0, 2)[e];
        two[e] = two[e] + -1.0 * FIELD(M->tw, c, 1, 2)[e];
      }
    const double weight = M->coupled ? gamma * gamma : 1.0;
    for (size_t e = 0; e < N; ++e) {
      squares[e] += weight * (one[e] * one[e] + two[e] * two[e]);
      if (!M->coupled) {
        first[e] += gamma * sqrt(squares[e]);
        squares[e] = 0.0;
      }
    }
  }
  if (M->coupled)
    for (size_t e = 0; e < N; ++e) {
      first[e] = sqrt(squares[e]);
      squares[e] = 0.0;
    }
  if (!M->tgv) {
    const double variation = M->alpha * oracle_sum_rows(first, H, W);
    *primal = variation + data;
    double conjugates = 0.0;
    for (size_t c = 0; c < C; ++c) {
      divergence(FIELD(M->tp, c, 0, 2), FIELD(M->tp, c, 1, 2), H, W, one);
      for (size_t e = 0; e < N; ++e) xi[e] = M->gamma[c] * one[e];
      conjugates += (M->sm != NULL ? samples_conjugate(M->sm, c, xi) : frame_conjugate(f, c, xi, M->free_radius));
    }
    *dual = -conjugates;
    *theta_out = NAN;
    *partial_out = NAN;
  } else {
    for (size_t c = 0; c < C; ++c) {
      const double gamma = M->gamma[c];
      symmetrized_gradient(FIELD(M->tw, c, 0, 2), FIELD(M->tw, c, 1, 2), H, W, one, two, three);
      const double weight = M->coupled ? gamma * gamma : 1.0;
      for (size_t e = 0; e < N; ++e) {
        squares[e] += weight * (one[e] * one[e] + two[e] * two[e] + 2.0 * (three[e] * three[e]));
        if (!M->coupled) {
          second[e] += gamma * sqrt(squares[e]);
          squares[e] = 0.0;
        }
      }
    }
    if (M->coupled)
      for (size_t e = 0; e < N; ++e) {
        second[e] = sqrt(squares[e]);
        squares[e] = 0.0;
      }
    double most = 0.0;
    for (size_t c = 0; c < C; ++c) {
      double* dx = divergences + (2 * c) * N;
      double* dy = divergences + (2 * c + 1) * N;
      tensor_divergence(FIELD(M->tr, c, 0, 3), FIELD(M->tr, c, 1, 3), FIELD(M->tr, c, 2, 3), H, W, dx, dy);
      for (size_t e = 0; e < N; ++e) {
        squares[e] += 1.0 * (dx[e] * dx[e] + dy[e] * dy[e]);
        if (!M->coupled) {
          most = fmax(most, sqrt(squares[e]));
          squares[e] = 0.0;
        }
      }
    }
    if (M->coupled)
      for (size_t e = 0; e < N; ++e) most = fmax(most, sqrt(squares[e]));
    const double firsts = M->alpha1 * oracle_sum_rows(first, H, W);
    const double seconds = M->alpha0 * oracle_sum_rows(second, H, W);
    *primal = (firsts + seconds) + data;
    const double theta = most > M->alpha1 ? M->alpha1 / most : 1.0;
    *theta_out = theta;
    double conjugates = 0.0;
    for (size_t c = 0; c < C; ++c) {
      const double* dx = divergences + (2 * c) * N;
      const double* dy = divergences + (2 * c + 1) * N;
      for (size_t e = 0; e < N; ++e) {
        g1[e] = -theta * dx[e];
        g2[e] = -theta * dy[e];
      }
      divergence(g1, g2, H, W, one);
      for (size_t e = 0; e < N; ++e) xi[e] = M->gamma[c] * one[e];
      conjugates += (M->sm != NULL ? samples_conjugate(M->sm, c, xi) : frame_conjugate(f, c, xi, M->free_radius));
    }
    *dual = -conjugates;
    *partial_out = NAN;
    if (M->has_partial) {
      double partial = 0.0, residual = 0.0;
      for (size_t c = 0; c < C; ++c) {
        divergence(FIELD(M->tp, c, 0, 2), FIELD(M->tp, c, 1, 2), H, W, one);
        for (size_t e = 0; e < N; ++e) xi[e] = M->gamma[c] * one[e];
        partial += (M->sm != NULL ? samples_conjugate(M->sm, c, xi) : frame_conjugate(f, c, xi, M->free_radius));
        const double* dx = divergences + (2 * c) * N;
        const double* dy = divergences + (2 * c + 1) * N;
        for (size_t e = 0; e < N; ++e) {
          const double a = FIELD(M->tp, c, 0, 2)[e] + 1.0 * dx[e];
          const double b = FIELD(M->tp, c, 1, 2)[e] + 1.0 * dy[e];
          double sq = 0.0;
          sq += 1.0 * (a * a + b * b);
          g1[e] = sqrt(sq);
        }
        residual += M->gamma[c] * oracle_sum_rows(g1, H, W);
      }
      *partial_out = *primal + partial + M->partial_radius * residual;
    }
  }
  free(first);
  free(second);
  free(squares);
  free(one);
  free(two);
  free(three);
  free(divergences);
  free(xi);
  free(g1);
  free(g2);
}

/* The run (primal_dual.hpp:958-1013; samples.rs:613-626): records at 0 (where any are
 * taken), every `every` iterations and at the last; a tolerance met stops it. The relative
 * tolerance compares with |primal| where `absolute` (the samples' solvers), with primal
 * otherwise. */
static void run_machine(machine* M, oracle_result* result, double tolerance, double relative, double partial_tolerance,
                        uint64_t every, uint64_t most, int absolute) {
  const double samples = (double)M->C * (double)M->H * (double)M->W;
  int stopped = 0;
  uint64_t iteration = 0;
#define TAKE(at)                                                                                      \
  do {                                                                                                \
    double primal, dual, theta, partial;                                                              \
    record(M, &primal, &dual, &theta, &partial);                                                      \
    push_record(result, (at), primal, dual, theta, partial);                                          \
    const double gap = primal - dual;                                                                 \
    int met = 0;                                                                                      \
    if (tolerance > 0.0 && gap / samples <= tolerance) met = 1;                                       \
    if (relative > 0.0 && gap <= relative * (absolute ? fabs(primal) : primal)) met = 1;              \
    if (partial_tolerance > 0.0 && M->has_partial && partial / samples <= partial_tolerance) met = 1; \
    if (met) stopped = 1;                                                                             \
  } while (0)
  if (every > 0 || most == 0) TAKE(0);
  while (!stopped && iteration < most) {
    struct timespec began, ended;
    clock_gettime(CLOCK_MONOTONIC, &began);
    iterate(M);
    clock_gettime(CLOCK_MONOTONIC, &ended);
    const double seconds = (double)(ended.tv_sec - began.tv_sec) + 1e-9 * (double)(ended.tv_nsec - began.tv_nsec);
    M->iteration_seconds += seconds;
    if (iteration == 0 || seconds < M->least_iteration_seconds) M->least_iteration_seconds = seconds;
    iteration += 1;
    if ((every > 0 && iteration % every == 0) || iteration == most) TAKE(iteration);
  }
#undef TAKE
  result->iterations = iteration;
  result->stop = stopped ? 1 : 2;
  result->iteration_seconds = M->iteration_seconds;
  result->least_iteration_seconds = M->least_iteration_seconds;
}

oracle_result* oracle_solve(uint32_t height, uint32_t width, int32_t color_space, uint32_t count,
                            const oracle_component* components, const oracle_data* data,
                            const oracle_options* options, const oracle_start* start, char* message,
                            size_t message_size) {
  make_basis();
  const uint32_t expected = color_space == 1 ? 1u : 3u;
  if (color_space < 1 || color_space > 3 || count != expected || height == 0 || width == 0) {
    say(message, message_size, "a picture of this colour space has %u components, not %u", expected, count);
    return NULL;
  }
  frame f;
  memset(&f, 0, sizeof f);
  f.channels = count;
  uint32_t hmax = 0, vmax = 0;
  for (uint32_t c = 0; c < count; ++c) {
    if (components[c].h < 1 || components[c].h > 4 || components[c].v < 1 || components[c].v > 4) {
      say(message, message_size, "sampling factors are 1 to 4");
      return NULL;
    }
    if ((uint32_t)components[c].h > hmax) hmax = (uint32_t)components[c].h;
    if ((uint32_t)components[c].v > vmax) vmax = (uint32_t)components[c].v;
  }
  if (count == 1) {
    f.rows = 8u * components[0].rows;
    f.columns = 8u * components[0].columns;
    if (f.rows < height || f.columns < width) {
      say(message, message_size, "the blocks do not cover the picture");
      return NULL;
    }
  } else {
    f.rows = 8u * vmax * ceiling(height, 8u * vmax);
    f.columns = 8u * hmax * ceiling(width, 8u * hmax);
  }
  for (uint32_t c = 0; c < count; ++c) {
    const oracle_component* in = &components[c];
    if (hmax % (uint32_t)in->h != 0 || vmax % (uint32_t)in->v != 0) {
      say(message, message_size, "sampling factors are not whole fractions of the largest");
      for (uint32_t k = 0; k < c; ++k) problem_free(&f.comps[k].pr);
      return NULL;
    }
    const size_t down = count == 1 ? 1u : vmax / (uint32_t)in->v;
    const size_t across = count == 1 ? 1u : hmax / (uint32_t)in->h;
    const size_t brows = count == 1 ? in->rows : ceiling(height, 8u * down);
    const size_t bcols = count == 1 ? in->columns : ceiling(width, 8u * across);
    if (in->rows < brows || in->columns < bcols) {
      say(message, message_size, "component %u has fewer blocks than the picture covers", c);
      for (uint32_t k = 0; k < c; ++k) problem_free(&f.comps[k].pr);
      return NULL;
    }
    int16_t* kept = (int16_t*)malloc(sizeof(int16_t) * 64u * ((brows * bcols) != 0 ? brows * bcols : 1));
    for (size_t by = 0; by < brows; ++by)
      memcpy(kept + 64u * by * bcols, in->levels + 64u * by * in->columns, sizeof(int16_t) * 64u * bcols);
    oracle_data term = data[c];
    if (color_space == 2 && c > 0 && !term.has_mu) term.mu_scale = term.mu_scale * term.mu_chroma;
    /* The MMSE decoder takes the MMSE centres, whatever the data term's are (decode.rs:273-281). */
    if (options->method == 0) term.centres = 0;
    const int made = problem_make(&f.comps[c].pr, kept, brows, bcols, in->table, &term, message, message_size);
    free(kept);
    if (!made) {
      for (uint32_t k = 0; k < c; ++k) problem_free(&f.comps[k].pr);
      return NULL;
    }
    f.comps[c].down = down;
    f.comps[c].across = across;
  }
  oracle_result* result = (oracle_result*)calloc(1, sizeof(oracle_result));
  const size_t C = count, H = f.rows, W = f.columns, N = H * W;
  result->channels = C;
  result->rows = H;
  result->columns = W;
  for (size_t c = 0; c < C; ++c) {
    const problem* pr = &f.comps[c].pr;
    const size_t blocks = pr->rows * pr->columns;
    result->block_rows[c] = pr->rows;
    result->block_columns[c] = pr->columns;
    result->down[c] = f.comps[c].down;
    result->across[c] = f.comps[c].across;
    result->arrays[c][1] = copy_of(pr->lower, 64u * blocks);
    result->arrays[c][2] = copy_of(pr->upper, 64u * blocks);
    result->arrays[c][3] = copy_of(pr->centres, 64u * blocks);
    result->arrays[c][4] = copy_of(pr->steps, 64);
    result->arrays[c][5] = copy_of(pr->weights, 64);
    result->arrays[c][6] = copy_of(pr->flat_weights, 64);
    result->arrays[c][7] = (double*)malloc(sizeof(double) * (blocks ? blocks : 1));
    for (size_t b = 0; b < blocks; ++b) result->arrays[c][7][b] = pr->flat[b];
    result->counts[c][0] = result->counts[c][1] = result->counts[c][2] = result->counts[c][3] = 64u * blocks;
    result->counts[c][4] = result->counts[c][5] = result->counts[c][6] = 64;
    result->counts[c][7] = blocks;
  }

  /* The plan (pdhg.cpp:147-233). */
  const oracle_options* o = options;
  const int tgv = o->method == 2;
  double gamma[3] = {1.0, 1.0, 1.0};
  if (o->method != 0) {
    int bad = !positive(o->alpha) || !positive(o->alpha1) || !positive(o->alpha0);
    if (o->has_channel_weights)
      for (size_t c = 0; c < C; ++c) bad = bad || !positive(o->channel_weights[c]);
    bad = bad || !(o->step_product > 0.0 && o->step_product < 1.0);
    bad = bad || (o->has_relaxation && !(o->relaxation > 0.0 && o->relaxation < 2.0));
    bad = bad || (o->has_norm_squared && !positive(o->norm_squared));
    bad = bad || (o->has_step_ratio && !positive(o->step_ratio));
    bad = bad || (o->has_tolerance && !at_least_zero(o->tolerance));
    bad = bad || !at_least_zero(o->relative_tolerance) || !at_least_zero(o->partial_tolerance);
    bad = bad || (o->has_partial_radius && !at_least_zero(o->partial_radius));
    bad = bad || (!tgv && (o->has_partial_radius || o->partial_tolerance > 0.0));
    bad = bad || (tgv && o->partial_tolerance > 0.0 && !o->has_partial_radius);
    bad = bad || !at_least_zero(o->free_radius);
    if (bad) {
      say(message, message_size, "an option of the method is out of its range");
      for (size_t c = 0; c < C; ++c) problem_free(&f.comps[c].pr);
      oracle_result_free(result);
      return NULL;
    }
  }
  if (o->has_channel_weights)
    for (size_t c = 0; c < C; ++c) gamma[c] = o->channel_weights[c];
  double largest = 0.0;
  for (size_t c = 0; c < C; ++c) largest = fmax(largest, gamma[c]);
  const double bound = tgv ? (17.0 + sqrt(33.0)) / 2.0 : 8.0;
  const double norm_squared = o->has_norm_squared ? o->norm_squared : bound * (largest * largest);
  const double weight = tgv ? o->alpha1 : o->alpha;
  const double default_ratio = tgv ? 3.0 : 10.0;
  const double default_tolerance = tgv ? 2e-3 : 2e-4;
  const double ratio = o->has_step_ratio ? o->step_ratio
                       : o->weight_scaling ? default_ratio / (weight * weight)
                                           : default_ratio;
  const double tolerance = o->has_tolerance ? o->tolerance
                           : o->weight_scaling ? default_tolerance * weight
                                               : default_tolerance;
  const uint64_t most = o->has_iterations ? o->iterations : (tgv ? 10000u : 5000u);
  const double product = o->step_product;
  result->tau = sqrt((ratio * product) / norm_squared);
  result->sigma = sqrt(product / (ratio * norm_squared));

  machine M;
  memset(&M, 0, sizeof M);
  M.f = &f;
  M.tgv = tgv;
  M.coupled = o->coupled;
  M.C = C;
  M.H = H;
  M.W = W;
  M.N = N;
  M.tau = result->tau;
  M.sigma = result->sigma;
  M.rho = o->has_relaxation ? o->relaxation : 1.9;
  M.rest = 1.0 - M.rho;
  M.alpha = o->alpha;
  M.alpha1 = o->alpha1;
  M.alpha0 = o->alpha0;
  for (size_t c = 0; c < 3; ++c) M.gamma[c] = gamma[c];
  M.free_radius = o->free_radius;
  M.has_partial = tgv && o->has_partial_radius;
  M.partial_radius = o->partial_radius;
  M.x = (double*)calloc(C * N, sizeof(double));
  M.tx = (double*)calloc(C * N, sizeof(double));
  M.p = (double*)calloc(2 * C * N, sizeof(double));
  M.tp = (double*)calloc(2 * C * N, sizeof(double));
  if (tgv) {
    M.w = (double*)calloc(2 * C * N, sizeof(double));
    M.tw = (double*)calloc(2 * C * N, sizeof(double));
    M.r = (double*)calloc(3 * C * N, sizeof(double));
    M.tr = (double*)calloc(3 * C * N, sizeof(double));
  }
  for (size_t c = 0; c < C; ++c) {
    const problem* pr = &f.comps[c].pr;
    const size_t entries = 64u * pr->rows * pr->columns;
    M.zeta[c] = (double*)malloc(sizeof(double) * (entries ? entries : 1));
    const double* from = (start != NULL && start->coefficients[c] != NULL) ? start->coefficients[c] : pr->centres;
    for (size_t e = 0; e < entries; ++e) M.zeta[c][e] = clamp_to(from[e], pr->lower[e], pr->upper[e]);
    frame_canvas(&f, c, M.zeta[c], 128.0, FIELD(M.x, c, 0, 1));
    if (start != NULL && start->p != NULL) memcpy(FIELD(M.p, c, 0, 2), start->p + 2 * c * N, sizeof(double) * 2 * N);
    if (tgv) {
      if (start != NULL && start->w != NULL) {
        memcpy(FIELD(M.w, c, 0, 2), start->w + 2 * c * N, sizeof(double) * 2 * N);
      } else {
        gradient(FIELD(M.x, c, 0, 1), H, W, FIELD(M.w, c, 0, 2), FIELD(M.w, c, 1, 2));
      }
      if (start != NULL && start->r != NULL) memcpy(FIELD(M.r, c, 0, 3), start->r + 3 * c * N, sizeof(double) * 3 * N);
    }
  }
  if (o->method != 0) {
    project_duals(&M, M.p, M.r);
    memcpy(M.tx, M.x, sizeof(double) * C * N);
    memcpy(M.tp, M.p, sizeof(double) * 2 * C * N);
    if (tgv) {
      memcpy(M.tw, M.w, sizeof(double) * 2 * C * N);
      memcpy(M.tr, M.r, sizeof(double) * 3 * C * N);
    }
  } else {
    memcpy(M.tx, M.x, sizeof(double) * C * N);
  }

  /* The run (primal_dual.hpp:958-1013). */
  if (o->method != 0) {
    run_machine(&M, result, tolerance, o->relative_tolerance, o->partial_tolerance, o->record_every, most, 0);
  } else {
    result->iterations = 0;
    result->stop = 0;
  }
  result->canvas = copy_of(M.tx, C * N);
  result->iterate = copy_of(M.x, C * N);
  for (size_t c = 0; c < C; ++c) {
    result->arrays[c][0] = M.zeta[c];
    M.zeta[c] = NULL;
  }
  free(M.work);
  free(M.x);
  free(M.tx);
  free(M.p);
  free(M.tp);
  free(M.w);
  free(M.tw);
  free(M.r);
  free(M.tr);
  for (size_t c = 0; c < C; ++c) problem_free(&f.comps[c].pr);
  return result;
}

/* TV or TGV within the intervals of a picture of rounded samples (samples.rs:674-1080):
 * the method of 5 with the data term of 11, from the centres (projected), w their gradient
 * and the duals 0; the picture is the last tilde x. */
oracle_result* oracle_samples_solve(const uint16_t* levels, uint32_t channels, uint32_t height, uint32_t width,
                                    uint32_t bits, double mu, double slack, const oracle_options* options,
                                    char* message, size_t message_size) {
  const oracle_options* o = options;
  if (o->method != 1 && o->method != 2) {
    say(message, message_size, "the samples' methods are TV and TGV");
    return NULL;
  }
  samples_model model;
  if (channels == 0 || channels > 3 ||
      !samples_make(&model, levels, channels, height, width, bits, mu, slack)) {
    say(message, message_size, "a picture of 1 to 3 channels, levels of 1, 2, 4, 8 or 16 bits within range, "
                               "and a weight and slack at least 0");
    return NULL;
  }
  const int tgv = o->method == 2;
  const size_t C = channels, H = height, W = width, N = H * W;
  double gamma[3] = {1.0, 1.0, 1.0};
  if (o->has_channel_weights)
    for (size_t c = 0; c < C; ++c) gamma[c] = o->channel_weights[c];
  double largest = 0.0;
  for (size_t c = 0; c < C; ++c) largest = fmax(largest, gamma[c]);
  const double bound = tgv ? (17.0 + sqrt(33.0)) / 2.0 : 8.0;
  const double norm_squared = o->has_norm_squared ? o->norm_squared : bound * (largest * largest);
  const double weight = tgv ? o->alpha1 : o->alpha;
  const double scale = o->weight_scaling ? weight : 1.0;
  const double default_ratio = tgv ? 0.3 : 10.0;
  const double ratio = o->has_step_ratio ? o->step_ratio : default_ratio / (scale * scale);
  const double tolerance = o->has_tolerance ? o->tolerance : 1e-4 * scale;
  const uint64_t most = o->has_iterations ? o->iterations : 10000u;
  oracle_result* result = (oracle_result*)calloc(1, sizeof(oracle_result));
  result->channels = C;
  result->rows = H;
  result->columns = W;
  result->tau = sqrt((ratio * o->step_product) / norm_squared);
  result->sigma = sqrt(o->step_product / (ratio * norm_squared));

  machine M;
  memset(&M, 0, sizeof M);
  M.sm = &model;
  M.tgv = tgv;
  M.coupled = o->coupled;
  M.C = C;
  M.H = H;
  M.W = W;
  M.N = N;
  M.tau = result->tau;
  M.sigma = result->sigma;
  M.rho = o->has_relaxation ? o->relaxation : 1.9;
  M.rest = 1.0 - M.rho;
  M.alpha = o->alpha;
  M.alpha1 = o->alpha1;
  M.alpha0 = o->alpha0;
  for (size_t c = 0; c < 3; ++c) M.gamma[c] = gamma[c];
  M.has_partial = tgv && o->has_partial_radius;
  M.partial_radius = o->partial_radius;
  M.x = (double*)calloc(C * N, sizeof(double));
  M.tx = (double*)calloc(C * N, sizeof(double));
  M.p = (double*)calloc(2 * C * N, sizeof(double));
  M.tp = (double*)calloc(2 * C * N, sizeof(double));
  if (tgv) {
    M.w = (double*)calloc(2 * C * N, sizeof(double));
    M.tw = (double*)calloc(2 * C * N, sizeof(double));
    M.r = (double*)calloc(3 * C * N, sizeof(double));
    M.tr = (double*)calloc(3 * C * N, sizeof(double));
  }
  for (size_t k = 0; k < C * N; ++k) M.x[k] = clamp_to(model.centre[k], model.low[k], model.high[k]);
  for (size_t c = 0; c < C && tgv; ++c) gradient(FIELD(M.x, c, 0, 1), H, W, FIELD(M.w, c, 0, 2), FIELD(M.w, c, 1, 2));
  memcpy(M.tx, M.x, sizeof(double) * C * N);
  if (tgv) memcpy(M.tw, M.w, sizeof(double) * 2 * C * N);
  run_machine(&M, result, tolerance, o->relative_tolerance, o->partial_tolerance, o->record_every, most, 1);
  result->canvas = copy_of(M.tx, C * N);
  result->iterate = copy_of(M.x, C * N);
  free(M.work);
  free(M.x);
  free(M.tx);
  free(M.p);
  free(M.tp);
  free(M.w);
  free(M.tw);
  free(M.r);
  free(M.tr);
  samples_free(&model);
  return result;
}
