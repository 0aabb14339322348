// HIP kernels of the TV/TGV primal-dual iteration for gfx950, and the host driver.
//
// Layout in HBM: every field is planes of the canvas, row-major, pitch = canvas width:
// x[C][H][W], p[C][2][H][W] (x then y), w[C][2][H][W], r[C][3][H][W] (xx, yy, xy), plus
// the extrapolations x- and w- that the primal kernel hands to the dual kernel, and the
// tilde point kept for records. Levels stay int16 (rows x columns x 64 per component):
// the kernels make the ends, centres and weights of each coefficient from its level and
// 64 per-frequency constants, so the data term costs 2 bytes a coefficient to read.
//
// One iteration (docs/math.md:599-611) is two kernels:
//  * k_primal, one 64-lane wavefront per 8x8 block of a component (its cells of
//    canvas samples): v = x + tau gamma div p at every sample, the means of the cells, the
//    block DCT in LDS, the proximal map of each coefficient, the inverse DCT, the canvas
//    moved by the change; and for TGV w~ = w + tau gamma (p + div_2 r). It writes
//    x- = 2x~ - x and w- = 2w~ - w, and relaxes x and w in place (each is read only at
//    its own sample there).
//  * k_dual, one lane per pixel and every channel: the steps of p and r from the
//    differences of x- and w-, their projections onto the balls (coupled over the
//    channels or apart), and the relaxation of p and r in place.
// Records reduce in binary64 on the device, each work-group's terms in a fixed tree and
// the work-groups' partial sums in a fixed order, so that a run repeats to the last bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <mutex>
#include <set>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "solver.hpp"

namespace unround {

namespace {

struct GpuError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expression)                                                                             \
  do {                                                                                                    \
    const hipError_t error_ = (expression);                                                               \
    if (error_ != hipSuccess) {                                                                           \
      throw GpuError(std::string(#expression) + ": " + hipGetErrorString(error_));                        \
    }                                                                                                     \
  } while (0)

__constant__ double basis64[64];
__constant__ float basis32[64];

template <typename T>
__device__ __forceinline__ T basis(int k, int n);
template <>
__device__ __forceinline__ double basis<double>(int k, int n) {
  return basis64[8 * k + n];
}
template <>
__device__ __forceinline__ float basis<float>(int k, int n) {
  return basis32[8 * k + n];
}

// What the kernels know of a component: its levels and flat marks, its blocks and cells,
// and its per-frequency constants (cpp/include/unround/model.hpp:78-164).
struct DeviceComponent {
  const int16_t* levels;
  const uint8_t* flat;
  int rows, columns;  // blocks
  int down, across;   // the cell
  int tiles_y, tiles_x;
  int mmse, costed, free_samples;
  double slack;
  double steps[64];
  double weights[64];
  double flat_weights[64];
  double shrink[64];
  double costs[64];
};

// The fields of the method on the device, [channel][entry][H][W].
template <typename T>
struct Fields {
  T* x;
  T* xbar;
  T* tx;
  T* p;
  T* tp;
  T* w;
  T* wbar;
  T* tw;
  T* r;
  T* tr;
  T* zeta[3];
};

// The ends, centre and file's own ends of coefficient k of block b, in binary64, each
// step rounded apart as the reference rounds it (cpp/src/model.cpp:44-48, 242-262).
__device__ void ends_of(const DeviceComponent& K, int b, int k, double& lower, double& upper, double& centre,
                        double& own_lower, double& own_upper) {
#pragma clang fp contract(off)
  const int16_t q = K.levels[64 * b + k];
  const double level = static_cast<double>(q);
  const double shift = k == 0 ? 1024.0 : 0.0;
  const double step = K.steps[k];
  lower = ((level - 0.5) - K.slack) * step + shift;
  upper = ((level + 0.5) + K.slack) * step + shift;
  own_lower = (level - 0.5) * step + shift;
  own_upper = (level + 0.5) * step + shift;
  centre = level * step + shift;
  if (K.mmse && k > 0) {
    const double size = (fabs(level) - K.shrink[k]) * step;
    centre = q == 0 ? 0.0 : (q < 0 ? -size : size);
  }
}

template <typename T>
__device__ __forceinline__ T clamp_to(T value, T low, T high) {
  const T up = value < low ? low : value;
  return high < up ? high : up;
}

// The proximal map of one coefficient (docs/math.md:457-491; cpp/src/model.cpp:62-75).
template <typename T>
__device__ T prox_of(const DeviceComponent& K, int b, int k, T e, T step) {
  double lower, upper, centre, own_lower, own_upper;
  ends_of(K, b, k, lower, upper, centre, own_lower, own_upper);
  const T weight = static_cast<T>(K.flat[b] ? K.flat_weights[k] : K.weights[k]);
  const T t = step * weight;
  const T r = T(1) / (T(1) + t);
  const T z = (e + t * static_cast<T>(centre)) * r;
  if (!K.costed) return clamp_to(z, static_cast<T>(lower), static_cast<T>(upper));
  const T h = step * static_cast<T>(K.costs[k]) * r;
  const T high_own = static_cast<T>(own_upper), low_own = static_cast<T>(own_lower);
  const T above = high_own > z - h ? high_own : z - h;
  const T below = low_own < z + h ? low_own : z + h;
  const T inside = z < low_own ? below : z;
  return clamp_to(z > high_own ? above : inside, static_cast<T>(lower), static_cast<T>(upper));
}

// g*(s) of one coefficient, in binary64 (cpp/src/model.cpp:80-109).
__device__ double conjugate_of(const DeviceComponent& K, int b, int k, double s) {
  double low, high, centre, own_low, own_high;
  ends_of(K, b, k, low, high, centre, own_low, own_high);
  const double m = K.flat[b] ? K.flat_weights[k] : K.weights[k];
  if (!K.costed) {
    const double best = clamp_to(centre + s / m, low, high);
    const double away = best - centre;
    const double weighted = s * best - 0.5 * m * (away * away);
    const double support = fmax(s * low, s * high);
    return m > 0.0 ? weighted : support;
  }
  const double lambda = K.costs[k];
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
  const double beyond = fmax(fmax(best - own_high, own_low - best), 0.0);
  return s * best - 0.5 * m * (away * away) - lambda * beyond;
}

// G's term of one coefficient (cpp/src/model.cpp:411-436).
__device__ double value_of(const DeviceComponent& K, int b, int k, double c) {
  double low, high, centre, own_low, own_high;
  ends_of(K, b, k, low, high, centre, own_low, own_high);
  const double m = K.flat[b] ? K.flat_weights[k] : K.weights[k];
  const double away = c - centre;
  double term = 0.5 * m * (away * away);
  if (K.costed) term = term + K.costs[k] * fmax(fmax(c - own_high, own_low - c), 0.0);
  return term;
}

// ---------------------------------------------------------------------------------------
// The 8x8 DCT of a block in LDS, one lane per entry (m, n) = (lane / 8, lane % 8): the
// 8-point transforms by their even and odd halves (docs/math.md:108-122), forward down
// the columns then along the rows, inverse along the rows then down the columns.

// Output (m, n) of the forward transform down column n of block[8 row + column].
template <typename T>
__device__ __forceinline__ T forward_down(const T* block, int m, int n) {
  const bool odd = m & 1;
  T a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const T top = block[8 * j + n], bottom = block[8 * (7 - j) + n];
    a[j] = odd ? top - bottom : top + bottom;
  }
  return basis<T>(m, 0) * a[0] + basis<T>(m, 1) * a[1] + basis<T>(m, 2) * a[2] + basis<T>(m, 3) * a[3];
}

// Output (m, n) of the forward transform along row m.
template <typename T>
__device__ __forceinline__ T forward_along(const T* block, int m, int n) {
  const bool odd = n & 1;
  T a[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const T left = block[8 * m + j], right = block[8 * m + 7 - j];
    a[j] = odd ? left - right : left + right;
  }
  return basis<T>(n, 0) * a[0] + basis<T>(n, 1) * a[1] + basis<T>(n, 2) * a[2] + basis<T>(n, 3) * a[3];
}

// Output n of the inverse transform along row m: even part e_n' and odd part o_n' of
// n' = min(n, 7 - n), their sum below 4 and difference above.
template <typename T>
__device__ __forceinline__ T inverse_along(const T* block, int m, int n) {
  const int near = n < 4 ? n : 7 - n;
  const T* row = block + 8 * m;
  const T even = basis<T>(0, near) * row[0] + basis<T>(2, near) * row[2] + basis<T>(4, near) * row[4] +
                 basis<T>(6, near) * row[6];
  const T odd = basis<T>(1, near) * row[1] + basis<T>(3, near) * row[3] + basis<T>(5, near) * row[5] +
                basis<T>(7, near) * row[7];
  return n < 4 ? even + odd : even - odd;
}

// Output m of the inverse transform down column n.
template <typename T>
__device__ __forceinline__ T inverse_down(const T* block, int m, int n) {
  const int near = m < 4 ? m : 7 - m;
  const T even = basis<T>(0, near) * block[n] + basis<T>(2, near) * block[16 + n] + basis<T>(4, near) * block[32 + n] +
                 basis<T>(6, near) * block[48 + n];
  const T odd = basis<T>(1, near) * block[8 + n] + basis<T>(3, near) * block[24 + n] +
                basis<T>(5, near) * block[40 + n] + basis<T>(7, near) * block[56 + n];
  return m < 4 ? even + odd : even - odd;
}

// ---------------------------------------------------------------------------------------
// Differences (docs/math.md:320-371; cpp/src/operators.cpp:49-198), at one sample.

template <typename T>
__device__ __forceinline__ T load(const T* f, int i, int j, int W) {
  return f[static_cast<size_t>(i) * W + j];
}

// delta_x f at (i, j): f[i, j] (not in the last column) - f[i, j - 1] (not in the first).
template <typename T, typename R = T>
__device__ __forceinline__ R backward_across(const T* f, int i, int j, int W) {
  if (W == 1) return R(0);
  if (j == 0) return static_cast<R>(load(f, i, 0, W));
  if (j == W - 1) return -static_cast<R>(load(f, i, j - 1, W));
  return static_cast<R>(load(f, i, j, W)) - static_cast<R>(load(f, i, j - 1, W));
}

template <typename T, typename R = T>
__device__ __forceinline__ R backward_down(const T* f, int i, int j, int H, int W) {
  const bool first = i == 0, last = i == H - 1;
  if (!first && !last) return static_cast<R>(load(f, i, j, W)) - static_cast<R>(load(f, i - 1, j, W));
  if (!last) return static_cast<R>(load(f, i, j, W));
  if (!first) return -static_cast<R>(load(f, i - 1, j, W));
  return R(0);
}

template <typename T, typename R = T>
__device__ __forceinline__ R forward_across(const T* f, int i, int j, int W) {
  return j + 1 < W ? static_cast<R>(load(f, i, j + 1, W)) - static_cast<R>(load(f, i, j, W)) : R(0);
}

template <typename T, typename R = T>
__device__ __forceinline__ R forward_down(const T* f, int i, int j, int H, int W) {
  return i + 1 < H ? static_cast<R>(load(f, i + 1, j, W)) - static_cast<R>(load(f, i, j, W)) : R(0);
}

// div p = delta_y p_y + delta_x p_x.
template <typename T, typename R = T>
__device__ __forceinline__ R divergence_at(const T* px, const T* py, int i, int j, int H, int W) {
  R value = backward_down<T, R>(py, i, j, H, W);
  value += backward_across<T, R>(px, i, j, W);
  return value;
}

template <typename T>
__device__ __forceinline__ T ball_scale(T square, T radius) {
  const T norm = sqrt(square);
  return norm > radius ? radius / norm : T(1);
}

constexpr int kMaxCell = 4;

// ---------------------------------------------------------------------------------------
// The primal step of one component (primal_dual.hpp:525-558; frames.cpp:490-548).
template <typename T, bool TGV>
__global__ __launch_bounds__(256) void primal_kernel(const DeviceComponent* __restrict__ component, Fields<T> F,
                                                     int c, int H, int W, T descent, T step, T rho, T rest,
                                                     int keep) {
  __shared__ T first[4][64];
  __shared__ T second[4][64];
  const DeviceComponent& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.tiles_y * K.tiles_x;
  const int ty = active ? tile / K.tiles_x : 0, tx = active ? tile % K.tiles_x : 0;
  const bool inside = active && ty < K.rows && tx < K.columns;
  const int dv = K.down, dh = K.across, cells = dv * dh;
  const size_t N = static_cast<size_t>(H) * W;
  T* x = F.x + c * N;
  const T* px = F.p + (2 * c) * N;
  const T* py = F.p + (2 * c + 1) * N;
  T xs[kMaxCell * kMaxCell], vs[kMaxCell * kMaxCell];
  // v at each sample of the lane's cell, and their sum: rows of a column first, then
  // the columns (cpp/src/frames.cpp:323-339).
  T sum = T(0);
#pragma unroll
  for (int b = 0; b < kMaxCell; ++b) {
    T column = T(0);
#pragma unroll
    for (int a = 0; a < kMaxCell; ++a) {
      if (active && a < dv && b < dh) {
        const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + b;
        const T here = load(x, i, j, W);
        const T v = here + descent * divergence_at(px, py, i, j, H, W);
        xs[kMaxCell * a + b] = here;
        vs[kMaxCell * a + b] = v;
        column = a == 0 ? v : column + v;
      }
    }
    if (b < dh) sum = b == 0 ? column : sum + column;
  }
  const T mean = cells > 1 ? sum / static_cast<T>(cells) : sum;
  first[wave][lane] = mean;
  __syncthreads();
  second[wave][lane] = forward_down(first[wave], m, n);
  __syncthreads();
  const T e = forward_along(second[wave], m, n);
  const int b = ty * K.columns + tx;
  const T zeta = inside ? prox_of(K, b, lane, e, step) : T(0);
  if (keep && inside) F.zeta[c][64 * static_cast<size_t>(b) + lane] = zeta;
  first[wave][lane] = zeta;
  __syncthreads();
  second[wave][lane] = inverse_along(first[wave], m, n);
  __syncthreads();
  const T back = inverse_down(second[wave], m, n);
  const T change = back - mean;
  T* xbar = F.xbar + c * N;
  T* tx_field = F.tx + c * N;
#pragma unroll
  for (int a = 0; a < kMaxCell; ++a) {
#pragma unroll
    for (int bb = 0; bb < kMaxCell; ++bb) {
      if (!(active && a < dv && bb < dh)) continue;
      const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + bb;
      const T v = vs[kMaxCell * a + bb], old = xs[kMaxCell * a + bb];
      const T tilde = inside ? (cells > 1 ? v + change : back) : v;
      const size_t at = static_cast<size_t>(i) * W + j;
      xbar[at] = T(2) * tilde - old;
      x[at] = rho * tilde + rest * old;
      if (keep) tx_field[at] = tilde;
      if constexpr (TGV) {
        // w~ = w + tau gamma (p + div_2 r), w- = 2 w~ - w, w relaxed.
        const T* rxx = F.r + (3 * c) * N;
        const T* ryy = F.r + (3 * c + 1) * N;
        const T* rxy = F.r + (3 * c + 2) * N;
        T d2x = forward_across(rxx, i, j, W);
        d2x += forward_down(rxy, i, j, H, W);
        T d2y = forward_across(rxy, i, j, W);
        d2y += forward_down(ryy, i, j, H, W);
        T* wx = F.w + (2 * c) * N;
        T* wy = F.w + (2 * c + 1) * N;
        const T wx_old = wx[at], wy_old = wy[at];
        const T twx = wx_old + descent * (px[at] + d2x);
        const T twy = wy_old + descent * (py[at] + d2y);
        F.wbar[(2 * c) * N + at] = T(2) * twx - wx_old;
        F.wbar[(2 * c + 1) * N + at] = T(2) * twy - wy_old;
        wx[at] = rho * twx + rest * wx_old;
        wy[at] = rho * twy + rest * wy_old;
        if (keep) {
          F.tw[(2 * c) * N + at] = twx;
          F.tw[(2 * c + 1) * N + at] = twy;
        }
      }
    }
  }
}

// The dual steps of every channel at one pixel (primal_dual.hpp:572-735).
template <typename T, bool TGV, bool COUPLED>
__global__ __launch_bounds__(256) void dual_kernel(Fields<T> F, int C, int H, int W, T a0, T a1, T a2, T radius,
                                                   T radius0, T rho, T rest, int keep) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W;
  const size_t at = static_cast<size_t>(i) * W + j;
  const T ascents[3] = {a0, a1, a2};
  T step_p[3][2], step_r[3][3];
  T square = T(0), tensor_square = T(0);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= C) break;
    const T a = ascents[c];
    const T* xbar = F.xbar + c * N;
    const T g1 = forward_across(xbar, i, j, W);
    const T g2 = forward_down(xbar, i, j, H, W);
    const T pxv = F.p[(2 * c) * N + at], pyv = F.p[(2 * c + 1) * N + at];
    if constexpr (TGV) {
      const T* bwx = F.wbar + (2 * c) * N;
      const T* bwy = F.wbar + (2 * c + 1) * N;
      step_p[c][0] = pxv + a * (g1 - bwx[at]);
      step_p[c][1] = pyv + a * (g2 - bwy[at]);
      const T exx = backward_across(bwx, i, j, W);
      const T eyy = backward_down(bwy, i, j, H, W);
      T exy = backward_down(bwx, i, j, H, W);
      exy += backward_across(bwy, i, j, W);
      exy = T(0.5) * exy;
      step_r[c][0] = F.r[(3 * c) * N + at] + a * exx;
      step_r[c][1] = F.r[(3 * c + 1) * N + at] + a * eyy;
      step_r[c][2] = F.r[(3 * c + 2) * N + at] + a * exy;
    } else {
      step_p[c][0] = pxv + a * g1;
      step_p[c][1] = pyv + a * g2;
    }
    if constexpr (COUPLED) {
      square += step_p[c][0] * step_p[c][0] + step_p[c][1] * step_p[c][1];
      if constexpr (TGV) {
        tensor_square +=
            step_r[c][0] * step_r[c][0] + step_r[c][1] * step_r[c][1] + T(2) * (step_r[c][2] * step_r[c][2]);
      }
    }
  }
  const T shared_scale = COUPLED ? ball_scale(square, radius) : T(1);
  const T shared_tensor_scale = (COUPLED && TGV) ? ball_scale(tensor_square, radius0) : T(1);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (c >= C) break;
    T scale = shared_scale;
    if constexpr (!COUPLED) {
      scale = ball_scale(step_p[c][0] * step_p[c][0] + step_p[c][1] * step_p[c][1], radius);
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const T projected = step_p[c][e] * scale;
      T* p = F.p + (2 * c + e) * N;
      if (keep) F.tp[(2 * c + e) * N + at] = projected;
      p[at] = rho * projected + rest * p[at];
    }
    if constexpr (TGV) {
      T tensor_scale = shared_tensor_scale;
      if constexpr (!COUPLED) {
        tensor_scale = ball_scale(step_r[c][0] * step_r[c][0] + step_r[c][1] * step_r[c][1] +
                                      T(2) * (step_r[c][2] * step_r[c][2]),
                                  radius0);
      }
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        const T projected = step_r[c][e] * tensor_scale;
        T* r = F.r + (3 * c + e) * N;
        if (keep) F.tr[(3 * c + e) * N + at] = projected;
        r[at] = rho * projected + rest * r[at];
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// The start (primal_dual.hpp:401-432; frames.cpp:569-593): x = the canvas of the
// coefficients given (or the centres) clipped to their intervals, 128 beyond the blocks.
template <typename T>
__global__ __launch_bounds__(256) void start_kernel(const DeviceComponent* __restrict__ component, Fields<T> F, int c,
                                                    int H, int W, const double* __restrict__ given) {
  __shared__ double first[4][64];
  __shared__ double second[4][64];
  const DeviceComponent& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.tiles_y * K.tiles_x;
  const int ty = active ? tile / K.tiles_x : 0, tx = active ? tile % K.tiles_x : 0;
  const bool inside = active && ty < K.rows && tx < K.columns;
  const int b = ty * K.columns + tx;
  double zeta = 0.0;
  if (inside) {
    double lower, upper, centre, own_lower, own_upper;
    ends_of(K, b, lane, lower, upper, centre, own_lower, own_upper);
    const double from = given != nullptr ? given[64 * static_cast<size_t>(b) + lane] : centre;
    zeta = clamp_to(from, lower, upper);
    F.zeta[c][64 * static_cast<size_t>(b) + lane] = static_cast<T>(zeta);
  }
  first[wave][lane] = zeta;
  __syncthreads();
  second[wave][lane] = inverse_along(first[wave], m, n);
  __syncthreads();
  const double back = inside ? inverse_down(second[wave], m, n) : 128.0;
  const size_t N = static_cast<size_t>(H) * W;
  for (int a = 0; a < K.down; ++a) {
    for (int bb = 0; bb < K.across; ++bb) {
      if (!active) continue;
      const int i = (8 * ty + m) * K.down + a, j = (8 * tx + n) * K.across + bb;
      const size_t at = c * N + static_cast<size_t>(i) * W + j;
      F.x[at] = static_cast<T>(back);
      F.tx[at] = static_cast<T>(back);
    }
  }
}

// w = grad x (the default start of TGV's w), or the given fields converted.
template <typename T>
__global__ void gradient_kernel(Fields<T> F, int C, int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  for (int c = 0; c < C; ++c) {
    const T* x = F.x + c * N;
    F.w[(2 * c) * N + at] = forward_across(x, i, j, W);
    F.w[(2 * c + 1) * N + at] = forward_down(x, i, j, H, W);
  }
}

template <typename T>
__global__ void convert_kernel(const double* __restrict__ from, T* __restrict__ to, size_t count) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) to[k] = static_cast<T>(from[k]);
}

// The duals of a start projected onto their balls (operators.cpp:378-425), and the
// tilde point made the start.
template <typename T>
__global__ void project_start_kernel(Fields<T> F, int C, int H, int W, int tgv, int coupled, T radius, T radius0) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  T square = T(0), tensor_square = T(0);
  for (int c = 0; c < C && coupled; ++c) {
    const T a = F.p[(2 * c) * N + at], b = F.p[(2 * c + 1) * N + at];
    square += a * a + b * b;
    if (tgv) {
      const T xx = F.r[(3 * c) * N + at], yy = F.r[(3 * c + 1) * N + at], xy = F.r[(3 * c + 2) * N + at];
      tensor_square += xx * xx + yy * yy + T(2) * (xy * xy);
    }
  }
  for (int c = 0; c < C; ++c) {
    T scale = ball_scale(square, radius), tensor_scale = ball_scale(tensor_square, radius0);
    if (!coupled) {
      const T a = F.p[(2 * c) * N + at], b = F.p[(2 * c + 1) * N + at];
      scale = ball_scale(a * a + b * b, radius);
      if (tgv) {
        const T xx = F.r[(3 * c) * N + at], yy = F.r[(3 * c + 1) * N + at], xy = F.r[(3 * c + 2) * N + at];
        tensor_scale = ball_scale(xx * xx + yy * yy + T(2) * (xy * xy), radius0);
      }
    }
    for (int e = 0; e < 2; ++e) {
      const T value = F.p[(2 * c + e) * N + at] * scale;
      F.p[(2 * c + e) * N + at] = value;
      F.tp[(2 * c + e) * N + at] = value;
    }
    if (tgv) {
      for (int e = 0; e < 2; ++e) F.tw[(2 * c + e) * N + at] = F.w[(2 * c + e) * N + at];
      for (int e = 0; e < 3; ++e) {
        const T value = F.r[(3 * c + e) * N + at] * tensor_scale;
        F.r[(3 * c + e) * N + at] = value;
        F.tr[(3 * c + e) * N + at] = value;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Records, in binary64 (primal_dual.hpp:446-501, 791-895; frames.cpp:616-681).

// A work-group's sum (or largest value) of one value per lane, in a fixed tree.
template <bool LARGEST>
__device__ double group_reduce(double value, double* room) {
  room[threadIdx.x] = value;
  __syncthreads();
  for (int half = 128; half > 0; half >>= 1) {
    if (static_cast<int>(threadIdx.x) < half) {
      const double other = room[threadIdx.x + half];
      room[threadIdx.x] = LARGEST ? fmax(room[threadIdx.x], other) : room[threadIdx.x] + other;
    }
    __syncthreads();
  }
  const double total = room[0];
  __syncthreads();
  return total;
}

// The norms of the tilde point's pixels: the first order (TV's gradient, or TGV's
// grad x~ - w~), TGV's |E w~|_F, the largest norm of div_2 r~, and the partial gap's
// sum over the channels of gamma |p~ + div_2 r~|; each work-group's sums and largest into
// partials[4 group + 0..3].
template <typename T, bool TGV>
__global__ __launch_bounds__(256) void norms_kernel(Fields<T> F, int C, int H, int W, double g0, double g1, double g2,
                                                    int coupled, double* __restrict__ partials) {
  __shared__ double room[256];
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  const bool here = i < H && j < W;
  const size_t N = static_cast<size_t>(H) * W;
  const double gammas[3] = {g0, g1, g2};
  double first = 0.0, second = 0.0, most = 0.0, residual = 0.0;
  if (here) {
    const size_t at = static_cast<size_t>(i) * W + j;
    double squares = 0.0;
    for (int c = 0; c < C; ++c) {
      const T* x = F.tx + c * N;
      double a = forward_across<T, double>(x, i, j, W);
      double b = forward_down<T, double>(x, i, j, H, W);
      if (TGV) {
        a = a + -1.0 * static_cast<double>(F.tw[(2 * c) * N + at]);
        b = b + -1.0 * static_cast<double>(F.tw[(2 * c + 1) * N + at]);
      }
      squares += (coupled ? gammas[c] * gammas[c] : 1.0) * (a * a + b * b);
      if (!coupled) {
        first += gammas[c] * sqrt(squares);
        squares = 0.0;
      }
    }
    if (coupled) first = sqrt(squares);
    if (TGV) {
      squares = 0.0;
      for (int c = 0; c < C; ++c) {
        const T* wx = F.tw + (2 * c) * N;
        const T* wy = F.tw + (2 * c + 1) * N;
        const double xx = backward_across<T, double>(wx, i, j, W);
        const double yy = backward_down<T, double>(wy, i, j, H, W);
        double xy = backward_down<T, double>(wx, i, j, H, W);
        xy += backward_across<T, double>(wy, i, j, W);
        xy = 0.5 * xy;
        squares += (coupled ? gammas[c] * gammas[c] : 1.0) * (xx * xx + yy * yy + 2.0 * (xy * xy));
        if (!coupled) {
          second += gammas[c] * sqrt(squares);
          squares = 0.0;
        }
      }
      if (coupled) second = sqrt(squares);
      squares = 0.0;
      for (int c = 0; c < C; ++c) {
        const T* rxx = F.tr + (3 * c) * N;
        const T* ryy = F.tr + (3 * c + 1) * N;
        const T* rxy = F.tr + (3 * c + 2) * N;
        double dx = forward_across<T, double>(rxx, i, j, W);
        dx += forward_down<T, double>(rxy, i, j, H, W);
        double dy = forward_across<T, double>(rxy, i, j, W);
        dy += forward_down<T, double>(ryy, i, j, H, W);
        squares += 1.0 * (dx * dx + dy * dy);
        if (!coupled) {
          most = fmax(most, sqrt(squares));
          squares = 0.0;
        }
        const double ra = static_cast<double>(F.tp[(2 * c) * N + at]) + 1.0 * dx;
        const double rb = static_cast<double>(F.tp[(2 * c + 1) * N + at]) + 1.0 * dy;
        residual += gammas[c] * sqrt(1.0 * (ra * ra + rb * rb));
      }
      if (coupled) most = sqrt(squares);
    }
  }
  const int group = blockIdx.y * gridDim.x + blockIdx.x;
  const double s_first = group_reduce<false>(first, room);
  const double s_second = group_reduce<false>(second, room);
  const double s_most = group_reduce<true>(most, room);
  const double s_residual = group_reduce<false>(residual, room);
  if (threadIdx.x == 0) {
    partials[4 * group + 0] = s_first;
    partials[4 * group + 1] = s_second;
    partials[4 * group + 2] = s_most;
    partials[4 * group + 3] = s_residual;
  }
}

// G's terms at the kept coefficients of one component, a work-group's sum each.
template <typename T>
__global__ __launch_bounds__(256) void value_kernel(const DeviceComponent* __restrict__ component, const T* zeta,
                                                    double* __restrict__ partials) {
  __shared__ double room[256];
  const DeviceComponent& K = *component;
  const size_t index = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  const size_t count = static_cast<size_t>(K.rows) * K.columns * 64;
  double term = 0.0;
  if (index < count) {
    const int b = static_cast<int>(index / 64), k = static_cast<int>(index % 64);
    term = value_of(K, b, k, static_cast<double>(zeta[index]));
  }
  const double total = group_reduce<false>(term, room);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// div_2 r at one sample, in binary64.
template <typename T>
__device__ void tensor_divergence_at(const T* rxx, const T* ryy, const T* rxy, int i, int j, int H, int W, double& dx,
                                     double& dy) {
  dx = forward_across<T, double>(rxx, i, j, W);
  dx += forward_down<T, double>(rxy, i, j, H, W);
  dy = forward_across<T, double>(rxy, i, j, W);
  dy += forward_down<T, double>(ryy, i, j, H, W);
}

// G_c* at xi = gamma div p~ (from_r = 0) or xi = gamma div(-theta div_2 r~) (from_r = 1),
// with the free samples' terms (frames.cpp:616-681): a work-group's sum into partials.
template <typename T>
__global__ __launch_bounds__(256) void conjugate_kernel(const DeviceComponent* __restrict__ component, Fields<T> F,
                                                        int c, int H, int W, double gamma, int from_r,
                                                        const double* __restrict__ most, double alpha1, double radius,
                                                        double* __restrict__ partials) {
  __shared__ double first[4][64];
  __shared__ double second[4][64];
  __shared__ double room[256];
  const DeviceComponent& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.tiles_y * K.tiles_x;
  const int ty = active ? tile / K.tiles_x : 0, tx = active ? tile % K.tiles_x : 0;
  const bool inside = active && ty < K.rows && tx < K.columns;
  const int dv = K.down, dh = K.across;
  const size_t N = static_cast<size_t>(H) * W;
  double theta = 1.0;
  if (from_r) theta = *most > alpha1 ? alpha1 / *most : 1.0;
  const T* px = F.tp + (2 * c) * N;
  const T* py = F.tp + (2 * c + 1) * N;
  const T* rxx = F.tr + (3 * c) * N;
  const T* ryy = F.tr + (3 * c + 1) * N;
  const T* rxy = F.tr + (3 * c + 2) * N;
  double xis[kMaxCell * kMaxCell];
  double sum = 0.0;
#pragma unroll
  for (int b = 0; b < kMaxCell; ++b) {
    double column = 0.0;
#pragma unroll
    for (int a = 0; a < kMaxCell; ++a) {
      if (active && a < dv && b < dh) {
        const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + b;
        double div = 0.0;
        if (!from_r) {
          div = divergence_at<T, double>(px, py, i, j, H, W);
        } else {
          // g = -theta div_2 r~, then div g = delta_y g_y + delta_x g_x.
          double dx, dy, gy_here = 0.0, gy_up = 0.0, gx_here = 0.0, gx_left = 0.0;
          tensor_divergence_at(rxx, ryy, rxy, i, j, H, W, dx, dy);
          gx_here = -theta * dx;
          gy_here = -theta * dy;
          if (i > 0) {
            tensor_divergence_at(rxx, ryy, rxy, i - 1, j, H, W, dx, dy);
            gy_up = -theta * dy;
          }
          if (j > 0) {
            tensor_divergence_at(rxx, ryy, rxy, i, j - 1, H, W, dx, dy);
            gx_left = -theta * dx;
          }
          const bool top = i == 0, bottom = i == H - 1;
          double down_part;
          if (!top && !bottom) down_part = gy_here - gy_up;
          else if (!bottom) down_part = gy_here;
          else if (!top) down_part = -gy_up;
          else down_part = 0.0;
          double across_part;
          if (W == 1) across_part = 0.0;
          else if (j == 0) across_part = gx_here;
          else if (j == W - 1) across_part = -gx_left;
          else across_part = gx_here - gx_left;
          div = down_part;
          div += across_part;
        }
        const double xi = gamma * div;
        xis[kMaxCell * a + b] = xi;
        column = a == 0 ? xi : column + xi;
      }
    }
    if (b < dh) sum = b == 0 ? column : sum + column;
  }
  first[wave][lane] = sum;
  __syncthreads();
  second[wave][lane] = forward_down(first[wave], m, n);
  __syncthreads();
  const double s = forward_along(second[wave], m, n);
  double term = inside ? conjugate_of(K, ty * K.columns + tx, lane, s) : 0.0;
  if (K.free_samples && active) {
    // <zeta, m> + R |zeta|_1 of the samples: within the blocks R |xi - mean of its cell|,
    // beyond them 128 xi + R |xi| (docs/math.md:911-949).
    const double mean = sum / static_cast<double>(dv * dh);
    for (int a = 0; a < dv; ++a) {
      for (int bb = 0; bb < dh; ++bb) {
        const double xi = xis[kMaxCell * a + bb];
        term += inside ? radius * fabs(xi - mean) : 128.0 * xi + radius * fabs(xi);
      }
    }
  }
  const double total = group_reduce<false>(term, room);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Sums (or the largest of) count partials, one work-group per column of a row-major
// [count][columns] array: each lane takes every 256th in order, then a fixed tree.
__global__ __launch_bounds__(256) void reduce_kernel(const double* __restrict__ partials, int count, int columns,
                                                     unsigned largest_mask, double* __restrict__ out) {
  __shared__ double room[256];
  const int column = blockIdx.x;
  const bool largest = (largest_mask >> column) & 1u;
  double value = 0.0;
  for (int k = threadIdx.x; k < count; k += 256) {
    const double next = partials[static_cast<size_t>(k) * columns + column];
    value = largest ? fmax(value, next) : value + next;
  }
  const double total = largest ? group_reduce<true>(value, room) : group_reduce<false>(value, room);
  if (threadIdx.x == 0) out[column] = total;
}

// The final projection in binary64: the last canvas's coefficients (of the means of its
// cells) clipped to their intervals narrowed by `margin`, and the canvas moved by the
// change, so that the result's block DCT requantizes to the file's levels.
template <typename T>
__global__ __launch_bounds__(256) void final_kernel(const DeviceComponent* __restrict__ component, const T* tx_field,
                                                    int W, double margin, double* __restrict__ canvas,
                                                    double* __restrict__ coefficients) {
  __shared__ double first[4][64];
  __shared__ double second[4][64];
  const DeviceComponent& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.tiles_y * K.tiles_x;
  const int ty = active ? tile / K.tiles_x : 0, tx = active ? tile % K.tiles_x : 0;
  const bool inside = active && ty < K.rows && tx < K.columns;
  const int dv = K.down, dh = K.across, cells = dv * dh;
  double vs[kMaxCell * kMaxCell];
  double sum = 0.0;
#pragma unroll
  for (int b = 0; b < kMaxCell; ++b) {
    double column = 0.0;
#pragma unroll
    for (int a = 0; a < kMaxCell; ++a) {
      if (active && a < dv && b < dh) {
        const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + b;
        const double v = static_cast<double>(load(tx_field, i, j, W));
        vs[kMaxCell * a + b] = v;
        column = a == 0 ? v : column + v;
      }
    }
    if (b < dh) sum = b == 0 ? column : sum + column;
  }
  const double mean = cells > 1 ? sum / static_cast<double>(cells) : sum;
  first[wave][lane] = mean;
  __syncthreads();
  second[wave][lane] = forward_down(first[wave], m, n);
  __syncthreads();
  const double e = forward_along(second[wave], m, n);
  double zeta = 0.0;
  if (inside) {
    double lower, upper, centre, own_lower, own_upper;
    const int b = ty * K.columns + tx;
    ends_of(K, b, lane, lower, upper, centre, own_lower, own_upper);
    zeta = clamp_to(e, lower + margin, upper - margin);
    coefficients[64 * static_cast<size_t>(b) + lane] = zeta;
  }
  first[wave][lane] = zeta;
  __syncthreads();
  second[wave][lane] = inverse_along(first[wave], m, n);
  __syncthreads();
  const double back = inverse_down(second[wave], m, n);
  for (int a = 0; a < dv; ++a) {
    for (int bb = 0; bb < dh; ++bb) {
      if (!active) continue;
      const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + bb;
      const double v = vs[kMaxCell * a + bb];
      canvas[static_cast<size_t>(i) * W + j] = inside ? (cells > 1 ? v + (back - mean) : back) : v;
    }
  }
}

// ---------------------------------------------------------------------------------------
// The host driver.

struct DeviceBuffer {
  void* pointer = nullptr;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) {
    if (bytes > 0) HIP_CHECK(hipMalloc(&pointer, bytes));
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& other) noexcept : pointer(other.pointer) { other.pointer = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
    std::swap(pointer, other.pointer);
    return *this;
  }
  ~DeviceBuffer() {
    if (pointer != nullptr) (void)hipFree(pointer);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(pointer);
  }
};

struct StreamHolder {
  hipStream_t stream = nullptr;
  StreamHolder() { HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking)); }
  ~StreamHolder() {
    if (stream != nullptr) (void)hipStreamDestroy(stream);
  }
};

struct EventPair {
  hipEvent_t begin = nullptr, end = nullptr;
  EventPair() {
    HIP_CHECK(hipEventCreate(&begin));
    HIP_CHECK(hipEventCreate(&end));
  }
  ~EventPair() {
    (void)hipEventDestroy(begin);
    (void)hipEventDestroy(end);
  }
};

// The DCT basis into each device's constant memory, once per device (the solve's device
// is the current one).
void upload_basis() {
  static std::mutex guard;
  static std::set<int> devices;
  int device = 0;
  HIP_CHECK(hipGetDevice(&device));
  const std::lock_guard<std::mutex> lock(guard);
  if (devices.count(device) != 0) return;
  const std::array<double, 64>& b = dct_basis();
  float single[64];
  for (int k = 0; k < 64; ++k) single[k] = static_cast<float>(b[k]);
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(basis64), b.data(), sizeof(double) * 64));
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(basis32), single, sizeof(float) * 64));
  devices.insert(device);
}

template <typename T>
class Solver {
 public:
  Solver(const FrameModel& frame, const Settings& settings, const Plan& plan, bool timing)
      : frame_(frame), settings_(settings), plan_(plan), timing_(timing) {}

  void run(const StartState* start, const Observer& observer, SolveOutput& out) {
    const auto began = std::chrono::steady_clock::now();
    upload_basis();
    C_ = static_cast<int>(frame_.channels());
    H_ = static_cast<int>(frame_.rows);
    W_ = static_cast<int>(frame_.columns);
    N_ = static_cast<size_t>(H_) * W_;
    tgv_ = settings_.method == Method::tgv;
    const bool solving = settings_.method == Method::tv || settings_.method == Method::tgv;
    coupled_ = settings_.coupled && C_ > 1;
    hipStream_t s = stream_.stream;
    allocate();
    upload_components();
    // The start.
    std::vector<DeviceBuffer> given(C_);
    for (int c = 0; c < C_; ++c) {
      const double* coefficients = start != nullptr && static_cast<size_t>(c) < start->coefficients.size()
                                       ? start->coefficients[c]
                                       : nullptr;
      double* device_given = nullptr;
      if (coefficients != nullptr) {
        const size_t bytes = sizeof(double) * 64 * frame_.components[c].blocks();
        given[c] = DeviceBuffer(bytes);
        HIP_CHECK(hipMemcpyAsync(given[c].pointer, coefficients, bytes, hipMemcpyHostToDevice, s));
        device_given = given[c].as<double>();
      }
      start_kernel<T><<<tile_groups(c), 256, 0, s>>>(component(c), fields_, c, H_, W_, device_given);
      HIP_CHECK(hipGetLastError());
    }
    const dim3 pixels((W_ + 63) / 64, (H_ + 3) / 4);
    if (solving) {
      upload_field(start != nullptr ? start->p : nullptr, fields_.p, 2);
      if (tgv_) {
        if (start != nullptr && start->w != nullptr) {
          upload_field(start->w, fields_.w, 2);
        } else {
          gradient_kernel<T><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_);
          HIP_CHECK(hipGetLastError());
        }
        upload_field(start != nullptr ? start->r : nullptr, fields_.r, 3);
      }
      const T radius = static_cast<T>(tgv_ ? settings_.alpha1 : settings_.alpha);
      project_start_kernel<T><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, tgv_, settings_.coupled ? 1 : 0, radius,
                                                     static_cast<T>(settings_.alpha0));
      HIP_CHECK(hipGetLastError());
    }
    out.tau = plan_.tau;
    out.sigma = plan_.sigma;
    uint64_t iteration = 0;
    int32_t stop = 0;
    if (solving) {
      const uint64_t most = plan_.iterations, every = plan_.record_every;
      const double samples = static_cast<double>(C_) * H_ * W_;
      bool stopped = false;
      std::vector<double> canvas_room;
      const auto take = [&](uint64_t at) {
        Record record = this->record(at);
        record.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - began).count();
        const double gap = record.primal - record.dual;
        bool met = false;
        if (plan_.tolerance > 0.0 && gap / samples <= plan_.tolerance) met = true;
        if (plan_.relative_tolerance > 0.0 && gap <= plan_.relative_tolerance * record.primal) met = true;
        if (plan_.partial_tolerance > 0.0 && plan_.has_partial && record.partial / samples <= plan_.partial_tolerance)
          met = true;
        out.records.push_back(record);
        if (met) {
          stop = 1;
          stopped = true;
          return;
        }
        if (observer && at > 0) {
          tilde_canvas(canvas_room);
          if (observer(record, canvas_room.data())) {
            stop = 3;
            stopped = true;
          }
        }
      };
      if (every > 0 || most == 0) take(0);
      while (!stopped && iteration < most) {
        const uint64_t next = iteration + 1;
        const bool keep = (every > 0 && next % every == 0) || next == most;
        if (timing_) {
          // The events of every launch, read once the iterations between two records
          // are done, so that timing adds no wait between launches.
          launches_.emplace_back();
          HIP_CHECK(hipEventRecord(launches_.back().begin, s));
        }
        iterate(keep);
        if (timing_) HIP_CHECK(hipEventRecord(launches_.back().end, s));
        iteration = next;
        if ((every > 0 && iteration % every == 0) || iteration == most) {
          if (timing_) collect_kernel_times(out.times);
          take(iteration);
        }
      }
      if (timing_) collect_kernel_times(out.times);
      if (stop == 0) stop = 2;
    }
    out.iterations = iteration;
    out.stop = solving ? stop : 0;
    finish(out);
  }

 private:
  int tile_groups(int c) const {
    const ComponentModel& k = frame_.components[c];
    const int tiles = static_cast<int>((frame_.rows / (8 * k.down)) * (frame_.columns / (8 * k.across)));
    return (tiles + 3) / 4;
  }

  const DeviceComponent* component(int c) const { return components_.as<DeviceComponent>() + c; }

  void allocate() {
    const size_t C = static_cast<size_t>(C_);
    size_t count = 3 * C * N_ + 2 * 2 * C * N_;  // x, xbar, tx, p, tp
    if (tgv_) count += 3 * 2 * C * N_ + 2 * 3 * C * N_;  // w, wbar, tw, r, tr
    size_t zeta_count = 0;
    for (const ComponentModel& k : frame_.components) zeta_count += 64 * k.blocks();
    state_ = DeviceBuffer(sizeof(T) * (count + zeta_count));
    HIP_CHECK(hipMemsetAsync(state_.pointer, 0, sizeof(T) * (count + zeta_count), stream_.stream));
    T* at = state_.as<T>();
    const auto take = [&at](size_t n) {
      T* here = at;
      at += n;
      return here;
    };
    fields_ = Fields<T>{};
    fields_.x = take(C * N_);
    fields_.xbar = take(C * N_);
    fields_.tx = take(C * N_);
    fields_.p = take(2 * C * N_);
    fields_.tp = take(2 * C * N_);
    if (tgv_) {
      fields_.w = take(2 * C * N_);
      fields_.wbar = take(2 * C * N_);
      fields_.tw = take(2 * C * N_);
      fields_.r = take(3 * C * N_);
      fields_.tr = take(3 * C * N_);
    }
    for (int c = 0; c < C_; ++c) fields_.zeta[c] = take(64 * frame_.components[c].blocks());
    // Partials of the records: the largest count of work-groups any of their kernels has.
    const size_t pixel_groups = static_cast<size_t>((W_ + 63) / 64) * ((H_ + 3) / 4);
    size_t value_groups = 0, conjugate_groups = 0;
    for (int c = 0; c < C_; ++c) {
      value_groups += (64 * frame_.components[c].blocks() + 255) / 256;
      conjugate_groups += static_cast<size_t>(tile_groups(c));
    }
    norm_partials_count_ = pixel_groups;
    value_partials_count_ = value_groups;
    conjugate_partials_count_ = conjugate_groups;
    partials_ = DeviceBuffer(sizeof(double) * (4 * pixel_groups + value_groups + 2 * conjugate_groups + 16));
    record_out_ = DeviceBuffer(sizeof(double) * 16);
  }

  void upload_components() {
    std::vector<DeviceComponent> host(C_);
    levels_.clear();
    flats_.clear();
    for (int c = 0; c < C_; ++c) {
      const ComponentModel& k = frame_.components[c];
      levels_.emplace_back(sizeof(int16_t) * std::max<size_t>(1, k.levels.size()));
      flats_.emplace_back(std::max<size_t>(1, k.flat.size()));
      HIP_CHECK(hipMemcpyAsync(levels_.back().pointer, k.levels.data(), sizeof(int16_t) * k.levels.size(),
                               hipMemcpyHostToDevice, stream_.stream));
      HIP_CHECK(hipMemcpyAsync(flats_.back().pointer, k.flat.data(), k.flat.size(), hipMemcpyHostToDevice,
                               stream_.stream));
      DeviceComponent& d = host[c];
      d.levels = levels_.back().as<int16_t>();
      d.flat = flats_.back().as<uint8_t>();
      d.rows = static_cast<int>(k.rows);
      d.columns = static_cast<int>(k.columns);
      d.down = static_cast<int>(k.down);
      d.across = static_cast<int>(k.across);
      d.tiles_y = static_cast<int>(frame_.rows / (8 * k.down));
      d.tiles_x = static_cast<int>(frame_.columns / (8 * k.across));
      d.mmse = k.mmse ? 1 : 0;
      d.costed = k.costed ? 1 : 0;
      d.free_samples = frame_.free(static_cast<size_t>(c)) ? 1 : 0;
      d.slack = k.slack;
      for (int q = 0; q < 64; ++q) {
        d.steps[q] = k.steps[q];
        d.weights[q] = k.weights[q];
        d.flat_weights[q] = k.flat_weights[q];
        d.shrink[q] = k.shrink[q];
        d.costs[q] = k.costs[q];
      }
      if (k.down > static_cast<uint32_t>(kMaxCell) || k.across > static_cast<uint32_t>(kMaxCell)) {
        throw GpuError("a cell larger than 4 x 4");
      }
    }
    components_ = DeviceBuffer(sizeof(DeviceComponent) * C_);
    HIP_CHECK(hipMemcpyAsync(components_.pointer, host.data(), sizeof(DeviceComponent) * C_, hipMemcpyHostToDevice,
                             stream_.stream));
    HIP_CHECK(hipStreamSynchronize(stream_.stream));
  }

  // A field of the start, converted from binary64, or left at 0.
  void upload_field(const double* given, T* field, int entries) {
    if (given == nullptr) return;
    const size_t count = static_cast<size_t>(entries) * C_ * N_;
    DeviceBuffer staging(sizeof(double) * count);
    HIP_CHECK(hipMemcpyAsync(staging.pointer, given, sizeof(double) * count, hipMemcpyHostToDevice, stream_.stream));
    convert_kernel<T><<<static_cast<unsigned>((count + 255) / 256), 256, 0, stream_.stream>>>(staging.as<double>(),
                                                                                             field, count);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream_.stream));
  }

  void iterate(bool keep) {
    hipStream_t s = stream_.stream;
    const T rho = static_cast<T>(plan_.relaxation), rest = static_cast<T>(1.0 - plan_.relaxation);
    if (timing_) {
      for (int c = 0; c < C_; ++c) primal_events_.emplace_back();
      dual_events_.emplace_back();
    }
    const size_t first_event = timing_ ? primal_events_.size() - static_cast<size_t>(C_) : 0;
    for (int c = 0; c < C_; ++c) {
      const ComponentModel& k = frame_.components[c];
      const T descent = static_cast<T>(plan_.tau * plan_.gamma[c]);
      const T step = static_cast<T>(plan_.tau / static_cast<double>(k.cells()));
      if (timing_) HIP_CHECK(hipEventRecord(primal_events_[first_event + c].begin, s));
      if (tgv_) {
        primal_kernel<T, true><<<tile_groups(c), 256, 0, s>>>(component(c), fields_, c, H_, W_, descent, step, rho,
                                                              rest, keep ? 1 : 0);
      } else {
        primal_kernel<T, false><<<tile_groups(c), 256, 0, s>>>(component(c), fields_, c, H_, W_, descent, step, rho,
                                                               rest, keep ? 1 : 0);
      }
      HIP_CHECK(hipGetLastError());
      if (timing_) HIP_CHECK(hipEventRecord(primal_events_[first_event + c].end, s));
    }
    const dim3 pixels((W_ + 63) / 64, (H_ + 3) / 4);
    T a[3] = {T(0), T(0), T(0)};
    for (int c = 0; c < C_; ++c) a[c] = static_cast<T>(plan_.sigma * plan_.gamma[c]);
    const T radius = static_cast<T>(tgv_ ? settings_.alpha1 : settings_.alpha);
    const T radius0 = static_cast<T>(settings_.alpha0);
    if (timing_) HIP_CHECK(hipEventRecord(dual_events_.back().begin, s));
    if (tgv_ && coupled_) {
      dual_kernel<T, true, true><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, a[0], a[1], a[2], radius, radius0, rho,
                                                        rest, keep ? 1 : 0);
    } else if (tgv_) {
      dual_kernel<T, true, false><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, a[0], a[1], a[2], radius, radius0, rho,
                                                         rest, keep ? 1 : 0);
    } else if (coupled_) {
      dual_kernel<T, false, true><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, a[0], a[1], a[2], radius, radius0, rho,
                                                         rest, keep ? 1 : 0);
    } else {
      dual_kernel<T, false, false><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, a[0], a[1], a[2], radius, radius0,
                                                          rho, rest, keep ? 1 : 0);
    }
    HIP_CHECK(hipGetLastError());
    if (timing_) HIP_CHECK(hipEventRecord(dual_events_.back().end, s));
  }

  // The durations of the launches since the last collection, then their events freed.
  void collect_kernel_times(KernelTimes& times) {
    HIP_CHECK(hipStreamSynchronize(stream_.stream));
    const auto elapsed = [](const EventPair& pair) {
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, pair.begin, pair.end));
      return ms * 1e-3;
    };
    for (const EventPair& pair : primal_events_) {
      times.primal_seconds += elapsed(pair);
      times.primal_launches += 1;
    }
    for (const EventPair& pair : dual_events_) {
      times.dual_seconds += elapsed(pair);
      times.dual_launches += 1;
    }
    for (const EventPair& pair : launches_) {
      times.iterations_seconds += elapsed(pair);
      times.iterations += 1;
    }
    primal_events_.clear();
    dual_events_.clear();
    launches_.clear();
  }

  // The record of the tilde point (primal_dual.hpp:446-501).
  Record record(uint64_t iteration) {
    hipStream_t s = stream_.stream;
    double* partials = partials_.as<double>();
    double* norm_partials = partials;
    double* value_partials = norm_partials + 4 * norm_partials_count_;
    double* conjugate_partials = value_partials + value_partials_count_;
    double* partial_partials = conjugate_partials + conjugate_partials_count_;
    double* results = record_out_.as<double>();
    const dim3 pixels((W_ + 63) / 64, (H_ + 3) / 4);
    const double g[3] = {plan_.gamma[0], C_ > 1 ? plan_.gamma[1] : 0.0, C_ > 2 ? plan_.gamma[2] : 0.0};
    if (tgv_) {
      norms_kernel<T, true><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, g[0], g[1], g[2], settings_.coupled ? 1 : 0,
                                                   norm_partials);
    } else {
      norms_kernel<T, false><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, g[0], g[1], g[2], settings_.coupled ? 1 : 0,
                                                    norm_partials);
    }
    HIP_CHECK(hipGetLastError());
    // results[0..3]: first, second, most, residual.
    reduce_kernel<<<4, 256, 0, s>>>(norm_partials, static_cast<int>(norm_partials_count_), 4, 1u << 2, results);
    HIP_CHECK(hipGetLastError());
    size_t offset = 0;
    for (int c = 0; c < C_; ++c) {
      const size_t groups = (64 * frame_.components[c].blocks() + 255) / 256;
      if (groups > 0) {
        value_kernel<T><<<static_cast<unsigned>(groups), 256, 0, s>>>(component(c), fields_.zeta[c],
                                                                      value_partials + offset);
        HIP_CHECK(hipGetLastError());
      }
      offset += groups;
    }
    reduce_kernel<<<1, 256, 0, s>>>(value_partials, static_cast<int>(value_partials_count_), 1, 0u, results + 4);
    const auto conjugates = [&](int from_r, double* into, double* result) {
      size_t at = 0;
      for (int c = 0; c < C_; ++c) {
        const int groups = tile_groups(c);
        conjugate_kernel<T><<<groups, 256, 0, s>>>(component(c), fields_, c, H_, W_, plan_.gamma[c], from_r,
                                                   results + 2, settings_.alpha1, plan_.free_radius, into + at);
        HIP_CHECK(hipGetLastError());
        at += static_cast<size_t>(groups);
      }
      reduce_kernel<<<1, 256, 0, s>>>(into, static_cast<int>(conjugate_partials_count_), 1, 0u, result);
      HIP_CHECK(hipGetLastError());
    };
    conjugates(tgv_ ? 1 : 0, conjugate_partials, results + 5);
    if (tgv_ && plan_.has_partial) conjugates(0, partial_partials, results + 6);
    double host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIP_CHECK(hipMemcpyAsync(host, results, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    Record record;
    record.iteration = iteration;
    const double data = host[4];
    if (!tgv_) {
      record.primal = settings_.alpha * host[0] + data;
      record.theta = std::nan("");
      record.partial = std::nan("");
    } else {
      record.primal = (settings_.alpha1 * host[0] + settings_.alpha0 * host[1]) + data;
      const double most = host[2];
      record.theta = most > settings_.alpha1 ? settings_.alpha1 / most : 1.0;
      record.partial = plan_.has_partial ? record.primal + host[6] + plan_.partial_radius * host[3] : std::nan("");
    }
    record.dual = -host[5];
    return record;
  }

  void tilde_canvas(std::vector<double>& canvas) {
    std::vector<T> values(static_cast<size_t>(C_) * N_);
    HIP_CHECK(hipMemcpyAsync(values.data(), fields_.tx, sizeof(T) * values.size(), hipMemcpyDeviceToHost,
                             stream_.stream));
    HIP_CHECK(hipStreamSynchronize(stream_.stream));
    canvas.assign(values.begin(), values.end());
  }

  void finish(SolveOutput& out) {
    hipStream_t s = stream_.stream;
    DeviceBuffer canvas(sizeof(double) * C_ * N_);
    std::vector<DeviceBuffer> coefficients;
    // Narrowed by 2^-30 of a unit: beyond what any binary64 block DCT of the result rounds
    // by, so that the coefficients requantize to their levels whatever the ties' rule.
    const double margin = std::ldexp(1.0, -30);
    out.coefficients.assign(C_, {});
    for (int c = 0; c < C_; ++c) {
      coefficients.emplace_back(sizeof(double) * std::max<size_t>(1, 64 * frame_.components[c].blocks()));
      final_kernel<T><<<tile_groups(c), 256, 0, s>>>(component(c), fields_.tx + c * N_, W_, margin,
                                                     canvas.as<double>() + c * N_, coefficients.back().as<double>());
      HIP_CHECK(hipGetLastError());
    }
    out.canvas.resize(static_cast<size_t>(C_) * N_);
    HIP_CHECK(hipMemcpyAsync(out.canvas.data(), canvas.pointer, sizeof(double) * out.canvas.size(),
                             hipMemcpyDeviceToHost, s));
    for (int c = 0; c < C_; ++c) {
      out.coefficients[c].resize(64 * frame_.components[c].blocks());
      HIP_CHECK(hipMemcpyAsync(out.coefficients[c].data(), coefficients[c].pointer,
                               sizeof(double) * out.coefficients[c].size(), hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }

  const FrameModel& frame_;
  const Settings& settings_;
  const Plan& plan_;
  bool timing_;
  int C_ = 0, H_ = 0, W_ = 0;
  size_t N_ = 0;
  bool tgv_ = false, coupled_ = false;
  StreamHolder stream_;
  DeviceBuffer state_, components_, partials_, record_out_;
  std::vector<DeviceBuffer> levels_, flats_;
  Fields<T> fields_{};
  size_t norm_partials_count_ = 0, value_partials_count_ = 0, conjugate_partials_count_ = 0;
  std::deque<EventPair> primal_events_, dual_events_, launches_;
};

}  // namespace

bool solve_on_gpu(const FrameModel& frame, const Settings& settings, const Plan& plan, const StartState* start,
                  const Observer& observer, bool timing, SolveOutput& out, std::string& message, int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    if (settings.precision == Precision::dual) {
      Solver<double> solver(frame, settings, plan, timing);
      solver.run(start, observer, out);
    } else {
      Solver<float> solver(frame, settings, plan, timing);
      solver.run(start, observer, out);
    }
    return true;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

}  // namespace unround
