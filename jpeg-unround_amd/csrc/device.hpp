// Device-side pieces shared by the kernels: the DCT basis in constant memory, a
// component's constants, the ends, proximal map, conjugate and value of a coefficient,
// the 8-point transforms, the finite differences at a sample, and the storage of the dual
// fields (binary32, binary64 or bfloat16).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "planes.hpp"

namespace unround {
namespace device {

extern __constant__ double basis64[64];
extern __constant__ float basis32[64];

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

// a b + c rounded once. The kernels' arithmetic on the iteration's path writes its fused
// multiply-adds out and lets the compiler fuse nothing else (`fp contract(off)`), so that
// every copy of a computation (a kernel's own, a fused kernel's, a batch's) rounds alike.
template <typename T>
__device__ __forceinline__ T fused(T a, T b, T c) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_fmaf(a, b, c);
  } else {
    return __builtin_fma(a, b, c);
  }
}

// A bfloat16 in memory: the high half of a binary32.
struct bf16 {
  uint16_t bits;
};

// Values of the storage type S as the compute type, and back, rounding to nearest even.
__device__ __forceinline__ float widen(float v) { return v; }
__device__ __forceinline__ double widen(double v) { return v; }
__device__ __forceinline__ float widen(bf16 v) { return __uint_as_float(static_cast<uint32_t>(v.bits) << 16); }
template <typename S>
__device__ __forceinline__ S narrow(float v);
template <>
__device__ __forceinline__ float narrow<float>(float v) {
  return v;
}
template <>
__device__ __forceinline__ bf16 narrow<bf16>(float v) {
  const uint32_t bits = __float_as_uint(v);
  const uint32_t rounding = 0x7fffu + ((bits >> 16) & 1u);
  return bf16{static_cast<uint16_t>((bits + rounding) >> 16)};
}
template <typename S>
__device__ __forceinline__ S narrow(double v);
template <>
__device__ __forceinline__ double narrow<double>(double v) {
  return v;
}

// What the kernels know of a component: its levels and flat marks, its blocks and cells,
// and its per-frequency constants (cpp/include/unround/model.hpp:78-164).
struct Component {
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

// A component's per-frequency constants copied into LDS by a work-group, so that lanes
// that each take other frequencies read them without a trip to memory; Staged reads them
// there, with the component's scalars, as a Component is read.
struct StagedArrays {
  double steps[64];
  double weights[64];
  double flat_weights[64];
  double shrink[64];
  double costs[64];
};
static_assert(offsetof(Component, costs) == offsetof(Component, steps) + 4 * 64 * sizeof(double));
static_assert(offsetof(StagedArrays, costs) == offsetof(StagedArrays, steps) + 4 * 64 * sizeof(double));

struct Staged {
  const double* steps;
  const double* weights;
  const double* flat_weights;
  const double* shrink;
  const double* costs;
  double slack;
  int mmse, costed;
};

// The work-group's 256 threads copy the constants of K into `into` (two loads each at
// most, with no branch); a barrier must follow before any thread reads them. In two
// halves, so that the loads can be issued before others and the copy made after them:
// stage_load() loads a thread's two constants, stage_store() puts them in LDS.
struct StageLoaded {
  double a, b;
};
__device__ __forceinline__ StageLoaded stage_load(const Component& K) {
  static_assert(5 * 64 <= 2 * 256);
  const double* from = K.steps;
  const int e = threadIdx.x;
  const int other = e + 256 < 5 * 64 ? e + 256 : e;
  return StageLoaded{from[e], from[other]};
}
__device__ __forceinline__ Staged stage_store(const Component& K, StagedArrays& into, const StageLoaded& loaded) {
  double* to = into.steps;
  const int e = threadIdx.x;
  const bool second = e + 256 < 5 * 64;
  const int other = second ? e + 256 : e;
  to[e] = loaded.a;
  to[other] = second ? loaded.b : loaded.a;
  return Staged{into.steps, into.weights, into.flat_weights, into.shrink, into.costs, K.slack, K.mmse, K.costed};
}
__device__ __forceinline__ Staged stage(const Component& K, StagedArrays& into) {
  return stage_store(K, into, stage_load(K));
}

// Orders a wavefront's own LDS stores before its lanes' loads of them, with no other
// wavefront taking part: a wave barrier between wavefront-scope release and acquire fences
// (a wavefront's LDS operations are performed in order, so no wait is needed; the fences
// keep the compiler from moving the accesses across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A pointer into device memory seen as one into the global address space, so that loads
// through it are global loads rather than generic ones.
template <typename X>
__device__ __forceinline__ const __attribute__((address_space(1))) X* global_view(const X* p) {
  return (const __attribute__((address_space(1))) X*)(p);
}

// The ends, centre and file's own ends of coefficient k of level q, in binary64, each
// step rounded apart as the reference rounds it (cpp/src/model.cpp:44-48, 242-262), from
// the frequency's step and MMSE shrinkage.
__device__ __forceinline__ void ends_from(int16_t q, int k, double step, double shrink, double slack, bool mmse,
                                          double& lower, double& upper, double& centre, double& own_lower,
                                          double& own_upper) {
#pragma clang fp contract(off)
  const double level = static_cast<double>(q);
  const double shift = k == 0 ? 1024.0 : 0.0;
  lower = ((level - 0.5) - slack) * step + shift;
  upper = ((level + 0.5) + slack) * step + shift;
  own_lower = (level - 0.5) * step + shift;
  own_upper = (level + 0.5) * step + shift;
  const double size = (fabs(level) - shrink) * step;
  const double shrunk = q == 0 ? 0.0 : (q < 0 ? -size : size);
  centre = (mmse && k > 0) ? shrunk : level * step + shift;
}

template <typename Constants>
__device__ inline void ends_of(const Constants& K, int16_t q, int k, double& lower, double& upper, double& centre,
                               double& own_lower, double& own_upper) {
  ends_from(q, k, K.steps[k], K.shrink[k], K.slack, K.mmse != 0, lower, upper, centre, own_lower, own_upper);
}

template <typename T>
__device__ __forceinline__ T clamp_to(T value, T low, T high) {
  const T up = value < low ? low : value;
  return high < up ? high : up;
}

// The proximal map of one coefficient of level q (docs/math.md:457-491;
// cpp/src/model.cpp:62-75): the quadratic's minimizer, clipped to the interval. The
// frequency's constants are read first and every case is computed and selected, so that
// a row of coefficients runs without branches. `within`: whether the quadratic's
// minimizer (with no slack cost) lay inside the interval. PLAIN: for a component with no
// slack cost and midpoint centres (the defaults), which the caller has checked; the same
// values with the other cases left out.
template <bool PLAIN = false, typename Constants, typename T>
__device__ __forceinline__ T prox_of(const Constants& K, int16_t q, bool flat, int k, T e, T step, bool& within) {
#pragma clang fp contract(off)
  const double frequency_step = K.steps[k];
  const double weight_of = (flat ? K.flat_weights : K.weights)[k];
  if constexpr (PLAIN) {
    // ends_from's lower, upper and midpoint centre.
    const double level = static_cast<double>(q);
    const double shift = k == 0 ? 1024.0 : 0.0;
    const double lower = ((level - 0.5) - K.slack) * frequency_step + shift;
    const double upper = ((level + 0.5) + K.slack) * frequency_step + shift;
    const double centre = level * frequency_step + shift;
    const T weight = static_cast<T>(weight_of);
    const T t = step * weight;
    const T r = T(1) / (T(1) + t);
    const T z = fused(t, static_cast<T>(centre), e) * r;
    within = z >= static_cast<T>(lower) && z <= static_cast<T>(upper);
    return clamp_to(z, static_cast<T>(lower), static_cast<T>(upper));
  }
  const double shrink = K.shrink[k];
  const double cost = K.costs[k];
  double lower, upper, centre, own_lower, own_upper;
  ends_from(q, k, frequency_step, shrink, K.slack, K.mmse != 0, lower, upper, centre, own_lower, own_upper);
  const T weight = static_cast<T>(weight_of);
  const T t = step * weight;
  const T r = T(1) / (T(1) + t);
  const T z = fused(t, static_cast<T>(centre), e) * r;
  const T h = step * static_cast<T>(cost) * r;
  const T high_own = static_cast<T>(own_upper), low_own = static_cast<T>(own_lower);
  const T above = high_own > z - h ? high_own : z - h;
  const T below = low_own < z + h ? low_own : z + h;
  const T inside = z < low_own ? below : z;
  const T costed = z > high_own ? above : inside;
  within = z >= static_cast<T>(lower) && z <= static_cast<T>(upper);
  return clamp_to(K.costed ? costed : z, static_cast<T>(lower), static_cast<T>(upper));
}
template <bool PLAIN = false, typename Constants, typename T>
__device__ __forceinline__ T prox_of(const Constants& K, int16_t q, bool flat, int k, T e, T step) {
  bool within;
  return prox_of<PLAIN>(K, q, flat, k, e, step, within);
}

// g*(s) of one coefficient, in binary64 (cpp/src/model.cpp:80-109).
__device__ inline double conjugate_of(const Component& K, int16_t q, bool flat, int k, double s) {
  double low, high, centre, own_low, own_high;
  ends_of(K, q, k, low, high, centre, own_low, own_high);
  const double m = flat ? K.flat_weights[k] : K.weights[k];
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
  const double flat_end = s > 0.0 ? upper_end : lower_end;
  const double best = m > 0.0 ? weighted : flat_end;
  const double away = best - centre;
  const double beyond = fmax(fmax(best - own_high, own_low - best), 0.0);
  return s * best - 0.5 * m * (away * away) - lambda * beyond;
}

// G's term of one coefficient (cpp/src/model.cpp:411-436).
__device__ inline double value_of(const Component& K, int16_t q, bool flat, int k, double c) {
  double low, high, centre, own_low, own_high;
  ends_of(K, q, k, low, high, centre, own_low, own_high);
  const double m = flat ? K.flat_weights[k] : K.weights[k];
  const double away = c - centre;
  double term = 0.5 * m * (away * away);
  if (K.costed) term = term + K.costs[k] * fmax(fmax(c - own_high, own_low - c), 0.0);
  return term;
}

// ---------------------------------------------------------------------------------------
// b0 v0 + b1 v1 + b2 v2 + b3 v3, fused from the first term on (docs/math.md:108-122).
template <typename T>
__device__ __forceinline__ T sum_of_four(T b0, T v0, T b1, T v1, T b2, T v2, T b3, T v3) {
#pragma clang fp contract(off)
  return fused(b3, v3, fused(b2, v2, fused(b1, v1, b0 * v0)));
}

// ---------------------------------------------------------------------------------------
// 8-point transforms by their even and odd halves (docs/math.md:108-122): each output a
// sum of four products from its first term on.

// y = C x of eight values in registers.
template <typename T>
__device__ __forceinline__ void forward8(const T (&x)[8], T (&y)[8]) {
#pragma clang fp contract(off)
  T s[4], d[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    s[j] = x[j] + x[7 - j];
    d[j] = x[j] - x[7 - j];
  }
#pragma unroll
  for (int k = 0; k < 8; k += 2) {
    y[k] = sum_of_four(basis<T>(k, 0), s[0], basis<T>(k, 1), s[1], basis<T>(k, 2), s[2], basis<T>(k, 3), s[3]);
    y[k + 1] = sum_of_four(basis<T>(k + 1, 0), d[0], basis<T>(k + 1, 1), d[1], basis<T>(k + 1, 2), d[2],
                           basis<T>(k + 1, 3), d[3]);
  }
}

// x = C^T y of eight values in registers: even and odd parts, their sum and difference.
template <typename T>
__device__ __forceinline__ void inverse8(const T (&y)[8], T (&x)[8]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const T even = sum_of_four(basis<T>(0, n), y[0], basis<T>(2, n), y[2], basis<T>(4, n), y[4], basis<T>(6, n), y[6]);
    const T odd = sum_of_four(basis<T>(1, n), y[1], basis<T>(3, n), y[3], basis<T>(5, n), y[5], basis<T>(7, n), y[7]);
    x[n] = even + odd;
    x[7 - n] = even - odd;
  }
}

// One lane per entry (m, n) of a block in LDS (block[8 row + column]).
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
// Differences (docs/math.md:320-371; cpp/src/operators.cpp:49-198) at one sample, of a
// field stored as S and taken as R.

template <typename R, typename S>
__device__ __forceinline__ R at(const S* f, int i, int j, int W) {
  return static_cast<R>(widen(f[static_cast<size_t>(i) * W + j]));
}

// delta_x f at (i, j): f[i, j] (not in the last column) - f[i, j - 1] (not in the first).
template <typename R, typename S>
__device__ __forceinline__ R backward_across(const S* f, int i, int j, int W) {
  if (W == 1) return R(0);
  if (j == 0) return at<R>(f, i, 0, W);
  if (j == W - 1) return -at<R>(f, i, j - 1, W);
  return at<R>(f, i, j, W) - at<R>(f, i, j - 1, W);
}

template <typename R, typename S>
__device__ __forceinline__ R backward_down(const S* f, int i, int j, int H, int W) {
  const bool first = i == 0, last = i == H - 1;
  if (!first && !last) return at<R>(f, i, j, W) - at<R>(f, i - 1, j, W);
  if (!last) return at<R>(f, i, j, W);
  if (!first) return -at<R>(f, i - 1, j, W);
  return R(0);
}

template <typename R, typename S>
__device__ __forceinline__ R forward_across(const S* f, int i, int j, int W) {
  return j + 1 < W ? at<R>(f, i, j + 1, W) - at<R>(f, i, j, W) : R(0);
}

template <typename R, typename S>
__device__ __forceinline__ R forward_down(const S* f, int i, int j, int H, int W) {
  return i + 1 < H ? at<R>(f, i + 1, j, W) - at<R>(f, i, j, W) : R(0);
}

// div p = delta_y p_y + delta_x p_x.
template <typename R, typename S>
__device__ __forceinline__ R divergence_at(const S* px, const S* py, int i, int j, int H, int W) {
  R value = backward_down<R>(py, i, j, H, W);
  value += backward_across<R>(px, i, j, W);
  return value;
}

// div_2 r = (d_x r_xx + d_y r_xy, d_x r_xy + d_y r_yy) at one sample.
template <typename R, typename S>
__device__ __forceinline__ void tensor_divergence_at(const S* rxx, const S* ryy, const S* rxy, int i, int j, int H,
                                                     int W, R& dx, R& dy) {
  dx = forward_across<R>(rxx, i, j, W);
  dx += forward_down<R>(rxy, i, j, H, W);
  dy = forward_across<R>(rxy, i, j, W);
  dy += forward_down<R>(ryy, i, j, H, W);
}

// A cell's shape (samples down, across) as a type, for choosing a kernel's instance.
template <int DOWN, int ACROSS>
struct Shape {
  static constexpr int first = DOWN, second = ACROSS;
};

// The same differences with every sample loaded first, at positions clamped into the
// canvas, and the case chosen after: no load sits under a branch, so that all of a lane's
// loads are in flight together. The values are those of the helpers above.
template <typename R, typename S>
__device__ __forceinline__ R clamped_at(const S* f, int i, int j, int H, int W) {
  return at<R>(f, min(max(i, 0), H - 1), min(max(j, 0), W - 1), W);
}

template <typename R, typename S>
__device__ __forceinline__ R backward_across_loaded(const S* f, int i, int j, int H, int W) {
  const R here = clamped_at<R>(f, i, j, H, W), left = clamped_at<R>(f, i, j - 1, H, W);
  R value = here - left;
  value = j == W - 1 ? -left : value;
  value = j == 0 ? here : value;
  return W == 1 ? R(0) : value;
}

template <typename R, typename S>
__device__ __forceinline__ R backward_down_loaded(const S* f, int i, int j, int H, int W) {
  const R here = clamped_at<R>(f, i, j, H, W), up = clamped_at<R>(f, i - 1, j, H, W);
  const bool first = i == 0, last = i == H - 1;
  R value = here - up;
  value = first && !last ? here : value;
  value = !first && last ? -up : value;
  return first && last ? R(0) : value;
}

template <typename R, typename S>
__device__ __forceinline__ R forward_across_loaded(const S* f, int i, int j, int H, int W) {
  const R here = clamped_at<R>(f, i, j, H, W), right = clamped_at<R>(f, i, j + 1, H, W);
  return j + 1 < W ? right - here : R(0);
}

template <typename R, typename S>
__device__ __forceinline__ R forward_down_loaded(const S* f, int i, int j, int H, int W) {
  const R here = clamped_at<R>(f, i, j, H, W), below = clamped_at<R>(f, i + 1, j, H, W);
  return i + 1 < H ? below - here : R(0);
}

template <typename R, typename S>
__device__ __forceinline__ R divergence_loaded(const S* px, const S* py, int i, int j, int H, int W) {
  R value = backward_down_loaded<R>(py, i, j, H, W);
  value += backward_across_loaded<R>(px, i, j, H, W);
  return value;
}

template <typename R, typename S>
__device__ __forceinline__ void tensor_divergence_loaded(const S* rxx, const S* ryy, const S* rxy, int i, int j, int H,
                                                         int W, R& dx, R& dy) {
  dx = forward_across_loaded<R>(rxx, i, j, H, W);
  dx += forward_down_loaded<R>(rxy, i, j, H, W);
  dy = forward_across_loaded<R>(rxy, i, j, H, W);
  dy += forward_down_loaded<R>(ryy, i, j, H, W);
}

// radius / |p| beyond the ball, 1 within (cpp/include/unround/operators.hpp:112-115).
template <typename T>
__device__ __forceinline__ T ball_scale(T square, T radius) {
  const T norm = sqrt(square);
  return norm > radius ? radius / norm : T(1);
}

// The hardware hands work-group g to XCD g mod 8 (each XCD with its own L2). The order in
// which a kernel takes its tiles, ordered(g), gives each XCD one contiguous run of them
// (XCD x the x-th eighth), so that tiles next to each other, which read each other's edge
// samples, share an L2. A bijection of [0, total).
__device__ __forceinline__ int ordered(int g, int total) {
  constexpr int kXcds = 8;
  const int xcd = g % kXcds, slot = g / kXcds;
  const int q = total / kXcds, r = total % kXcds;
  return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// ordered() over a grid of frames (blockIdx.y) of gridDim.x tiles each: the hardware counts
// work-groups x first, so each XCD takes one contiguous run of (frame, tile) pairs whatever
// the tiles a frame.
struct FrameTile {
  int frame, tile;
};
__device__ __forceinline__ FrameTile ordered_frames() {
  const int tiles = static_cast<int>(gridDim.x);
  const int t = ordered(static_cast<int>(blockIdx.y) * tiles + static_cast<int>(blockIdx.x), tiles * static_cast<int>(gridDim.y));
  return FrameTile{t / tiles, t % tiles};
}

// ---------------------------------------------------------------------------------------
// Planes of one kind of field: `count` planes of N samples each from `base`, a sample
// addressed by its index in its plane.
//
// BufferPlanes reads and writes through one buffer resource: the plane's byte offset is a
// scalar and the sample's a 32-bit lane offset, so that a row of every plane of the kind
// shares one address register (64-bit addresses per plane cost two registers each). The
// planes must span at most 2^31 bytes (planes::fits, which the host checks before it
// chooses an instance), so that kOutside lies beyond the resource. WidePlanes takes 64-bit
// addresses for larger spans.
constexpr int kRawBufferWord3 = 0x00020000;  // a raw (untyped) buffer resource on gfx9

// Cache policies of the buffer accesses (the instructions' cache-policy bits).
constexpr int kTemporal = 0, kStreaming = 2;

template <typename X, int POLICY, int NEXT>
struct BufferPlanesOf {
  // A byte offset beyond the resource (its planes span at most 2^31 bytes): a load there
  // returns zero and a store there is dropped, by the hardware's range check.
  static constexpr uint32_t kOutside = planes::kOutsideOffset;
  // The accesses' cache policy where they ask for none, and the policy of the stores that
  // the very next launch reads.
  static constexpr int kPolicy = POLICY, kNext = NEXT;
  __amdgpu_buffer_rsrc_t resource;
  uint32_t plane_bytes;
  __device__ __forceinline__ BufferPlanesOf(X* base, int count, size_t N)
      : resource(__builtin_amdgcn_make_buffer_rsrc(
            const_cast<void*>(static_cast<const void*>(base)), 0,
            static_cast<int>(static_cast<uint32_t>(static_cast<size_t>(count) * N * sizeof(X))), kRawBufferWord3)),
        plane_bytes(static_cast<uint32_t>(N * sizeof(X))) {}
  // The sample, or zero where `valid` is false (spans of at most 2 GiB), with no branch.
  template <int P = kPolicy>
  __device__ __forceinline__ X load_if(bool valid, int plane, uint32_t index) const {
    return load_at<P>(plane, valid ? index * static_cast<uint32_t>(sizeof(X)) : kOutside);
  }
  // The store where `valid`, dropped elsewhere (spans of at most 2 GiB), with no branch.
  template <int P = kPolicy>
  __device__ __forceinline__ void store_if(bool valid, int plane, uint32_t index, X value) const {
    store_at<P>(plane, valid ? index * static_cast<uint32_t>(sizeof(X)) : kOutside, value);
  }
  template <int P = kPolicy>
  __device__ __forceinline__ X load(int plane, uint32_t index) const {
    return load_at<P>(plane, index * static_cast<uint32_t>(sizeof(X)));
  }
  template <int P = kPolicy>
  __device__ __forceinline__ void store(int plane, uint32_t index, X value) const {
    store_at<P>(plane, index * static_cast<uint32_t>(sizeof(X)), value);
  }
  // Samples index and index + 1 (index even, so that the pair is aligned), zeros where
  // `valid` is false: one load of twice the width.
  template <int P = kPolicy>
  __device__ __forceinline__ void load_pair_if(bool valid, int plane, uint32_t index, X& first, X& second) const {
    const int offset = static_cast<int>(valid ? index * static_cast<uint32_t>(sizeof(X)) : kOutside);
    const int scalar = static_cast<int>(static_cast<uint32_t>(plane) * plane_bytes);
    if constexpr (sizeof(X) == 2) {
      const unsigned int both = __builtin_bit_cast(unsigned int, __builtin_amdgcn_raw_buffer_load_b32(resource, offset,
                                                                                                      scalar, P));
      first = __builtin_bit_cast(X, static_cast<unsigned short>(both & 0xFFFFu));
      second = __builtin_bit_cast(X, static_cast<unsigned short>(both >> 16));
    } else if constexpr (sizeof(X) == 4) {
      using Pair = unsigned int __attribute__((ext_vector_type(2)));
      const Pair both =
          __builtin_bit_cast(Pair, __builtin_amdgcn_raw_buffer_load_b64(resource, offset, scalar, P));
      // The words go through plain variables: a bit cast of a vector element (`both.y`)
      // is lowered by clang as a load from the vector's first element.
      const unsigned int low = both[0], high = both[1];
      first = __builtin_bit_cast(X, low);
      second = __builtin_bit_cast(X, high);
    } else {
      using Quad = unsigned int __attribute__((ext_vector_type(4)));
      using Pair = unsigned int __attribute__((ext_vector_type(2)));
      const Quad both =
          __builtin_bit_cast(Quad, __builtin_amdgcn_raw_buffer_load_b128(resource, offset, scalar, P));
      first = __builtin_bit_cast(X, Pair{both.x, both.y});
      second = __builtin_bit_cast(X, Pair{both.z, both.w});
    }
  }
  // The stores of index and index + 1 (index even) where `valid`, dropped elsewhere.
  template <int P = kPolicy>
  __device__ __forceinline__ void store_pair_if(bool valid, int plane, uint32_t index, X first, X second) const {
    const int offset = static_cast<int>(valid ? index * static_cast<uint32_t>(sizeof(X)) : kOutside);
    const int scalar = static_cast<int>(static_cast<uint32_t>(plane) * plane_bytes);
    if constexpr (sizeof(X) == 2) {
      const unsigned int both = static_cast<unsigned int>(__builtin_bit_cast(unsigned short, first)) |
                                (static_cast<unsigned int>(__builtin_bit_cast(unsigned short, second)) << 16);
      __builtin_amdgcn_raw_buffer_store_b32(both, resource, offset, scalar, P);
    } else if constexpr (sizeof(X) == 4) {
      using Pair = unsigned int __attribute__((ext_vector_type(2)));
      const Pair both{__builtin_bit_cast(unsigned int, first), __builtin_bit_cast(unsigned int, second)};
      __builtin_amdgcn_raw_buffer_store_b64(both, resource, offset, scalar, P);
    } else {
      using Quad = unsigned int __attribute__((ext_vector_type(4)));
      using Pair = unsigned int __attribute__((ext_vector_type(2)));
      const Pair a = __builtin_bit_cast(Pair, first), b = __builtin_bit_cast(Pair, second);
      const Quad both{a.x, a.y, b.x, b.y};
      __builtin_amdgcn_raw_buffer_store_b128(both, resource, offset, scalar, P);
    }
  }
  template <int P = kPolicy>
  __device__ __forceinline__ X load_at(int plane, uint32_t bytes) const {
    const int offset = static_cast<int>(bytes);
    const int scalar = static_cast<int>(static_cast<uint32_t>(plane) * plane_bytes);
    if constexpr (sizeof(X) == 2) {
      return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b16(resource, offset, scalar, P));
    } else if constexpr (sizeof(X) == 4) {
      return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b32(resource, offset, scalar, P));
    } else {
      return __builtin_bit_cast(X, __builtin_amdgcn_raw_buffer_load_b64(resource, offset, scalar, P));
    }
  }
  template <int P = kPolicy>
  __device__ __forceinline__ void store_at(int plane, uint32_t bytes, X value) const {
    const int offset = static_cast<int>(bytes);
    const int scalar = static_cast<int>(static_cast<uint32_t>(plane) * plane_bytes);
    if constexpr (sizeof(X) == 2) {
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, value), resource, offset, scalar,
                                            P);
    } else if constexpr (sizeof(X) == 4) {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, value), resource, offset, scalar, P);
    } else {
      using Pair = unsigned int __attribute__((ext_vector_type(2)));
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(Pair, value), resource, offset, scalar, P);
    }
  }
};

// One picture's fields: non-temporal (kStreaming), for a field read or written once in a
// launch and not again until the next iteration, a canvas' worth of traffic later; a store
// whose samples the very next launch reads (x- of the prox for the dual step, p and w of
// the dual step) asks for kTemporal and stays in the caches for it. This took the dual
// kernel from 161 to 147 us on 2048x2048 RGB 4:2:0 (207 to 196 us on 4096x4096 grey).
template <typename X>
using BufferPlanes = BufferPlanesOf<X, kStreaming, kTemporal>;
// The batch's fields, whose canvases outgrow the caches many times over. Its dual step's
// accesses take the default policy: its short runs lean on the caches for the samples
// their neighbours read too (the streaming policy measured 5 % slower there). The stores
// the next launch reads stream: they cannot stay in the caches until then (64 pictures of
// 96 to 1920 a side, 142 M samples, 300 iterations: 0.857 to 0.818 s).
template <typename X>
using CachedBufferPlanes = BufferPlanesOf<X, kTemporal, kStreaming>;
// The batch's prox: streaming throughout, x- too (the same batch: 0.819 to 0.816 s).
template <typename X>
using StreamedBufferPlanes = BufferPlanesOf<X, kStreaming, kStreaming>;

template <typename X>
struct WidePlanes {
  static constexpr int kPolicy = kTemporal, kNext = kTemporal;  // one policy: the default
  X* base;
  size_t N;
  __device__ __forceinline__ WidePlanes(X* base_, int, size_t N_) : base(base_), N(N_) {}
  template <int P = kStreaming>
  __device__ __forceinline__ X load(int plane, uint32_t index) const {
    return base[plane * N + index];
  }
  template <int P = kStreaming>
  __device__ __forceinline__ void store(int plane, uint32_t index, X value) const {
    base[plane * N + index] = value;
  }
  template <int P = kStreaming>
  __device__ __forceinline__ X load_if(bool valid, int plane, uint32_t index) const {
    return valid ? base[plane * N + index] : X{};
  }
  template <int P = kStreaming>
  __device__ __forceinline__ void store_if(bool valid, int plane, uint32_t index, X value) const {
    if (valid) base[plane * N + index] = value;
  }
  template <int P = kStreaming>
  __device__ __forceinline__ void load_pair_if(bool valid, int plane, uint32_t index, X& first, X& second) const {
    first = load_if(valid, plane, index);
    second = load_if(valid, plane, index + 1);
  }
  template <int P = kStreaming>
  __device__ __forceinline__ void store_pair_if(bool valid, int plane, uint32_t index, X first, X second) const {
    store_if(valid, plane, index, first);
    store_if(valid, plane, index + 1, second);
  }
};

}  // namespace device
}  // namespace unround
