// HIP kernels of the TV/TGV primal-dual iteration for gfx950, and the host driver.
//
// Layout in HBM: every field is planes of the canvas, row-major, pitch = canvas width:
// x[C][H][W], p[C][2][H][W] (x then y), w[C][2][H][W], r[C][3][H][W] (xx, yy, xy). x and
// w are stored as the iterate's type T (binary32 or binary64); p and r as S (T, or
// bfloat16 in the mixed precision). Levels stay int16 (rows x columns x 64 per
// component): the kernels make the ends, centres and weights of each coefficient from
// its level and 64 per-frequency constants, so the data term costs 2 bytes a coefficient.
//
// One iteration (docs/math.md:599-611) is two kernels:
//  * prox_kernel, per component: a wavefront takes a strip of 8 blocks (64 columns of the
//    component) and 8 rows, one lane per column. Each lane makes v = x + tau gamma div p
//    down its column (the means of its cells where a sample stands for several), the DCT
//    down the column in registers; the lanes swap through LDS so that each holds a row of
//    a block, which it transforms, maps (the proximal map of each coefficient) and
//    transforms back; they swap again and transform back down the columns. It writes
//    x- = 2 x~ - x and relaxes x in place (each lane reads x only at its own samples).
//    It reads x, p and the levels, and writes x- and x.
//  * dual_kernel, one lane per pixel and every channel: TGV's w~ = w + tau gamma (p +
//    div_2 r) at the pixel and at its neighbours left and above (for E w-), the steps of p
//    and r from the differences of x- and w-, their projections onto the balls (coupled
//    over the channels or apart), and the relaxation of w, p and r into the other of two
//    buffers (the neighbours read the buffers as they were).
// Records reduce in binary64 on the device, each work-group's terms in a fixed tree and
// the work-groups' partial sums in a fixed order, so that a run repeats to the last bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "planes.hpp"
#include "solver.hpp"

namespace unround {
namespace device {
__constant__ double basis64[64];
__constant__ float basis32[64];
}  // namespace device

namespace {

using device::bf16;
using device::Component;

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

// The fields of the method on the device, [channel][entry][H][W]: the state (w, p and r
// in two buffers each, the iteration reading one and writing the other), the
// extrapolation x-, the tilde point kept for records, and the coefficients zeta.
template <typename T, typename S>
struct Fields {
  T* x;
  T* xbar;
  T* xbar_next;  // the other buffer of x-, which the fused sweep writes
  T* tx;
  S* p;
  S* p_next;
  S* tp;
  T* w;
  T* w_next;
  T* tw;
  S* r;
  S* r_next;
  S* tr;
  T* zeta[3];
};

// The steps of an iteration.
template <typename T>
struct Steps {
  T descent[3];  // tau gamma_c
  T ascent[3];   // sigma gamma_c
  T radius;      // alpha (TV) or alpha1
  T radius0;     // alpha0
  T rho;
  T rest;  // 1 - rho
};

constexpr int kMaxCell = 4;

#include "iterate.hpp"
#include "sweep.hpp"

// ---------------------------------------------------------------------------------------
// The start (primal_dual.hpp:401-432; frames.cpp:569-593): x = the canvas of the
// coefficients given (or the centres) clipped to their intervals, 128 beyond the blocks.
template <typename T, typename S>
__global__ __launch_bounds__(256) void start_kernel(const Component* __restrict__ component, Fields<T, S> F, int c,
                                                    int H, int W, const double* __restrict__ given) {
  __shared__ double first[4][64];
  __shared__ double second[4][64];
  const Component& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.tiles_y * K.tiles_x;
  const int ty = active ? tile / K.tiles_x : 0, tx = active ? tile % K.tiles_x : 0;
  const bool inside = active && ty < K.rows && tx < K.columns;
  const size_t block = static_cast<size_t>(ty) * K.columns + tx;
  double zeta = 0.0;
  if (inside) {
    double lower, upper, centre, own_lower, own_upper;
    device::ends_of(K, K.levels[64 * block + lane], lane, lower, upper, centre, own_lower, own_upper);
    const double from = given != nullptr ? given[64 * block + lane] : centre;
    zeta = device::clamp_to(from, lower, upper);
    F.zeta[c][64 * block + lane] = static_cast<T>(zeta);
  }
  first[wave][lane] = zeta;
  __syncthreads();
  second[wave][lane] = device::inverse_along(first[wave], m, n);
  __syncthreads();
  const double back = inside ? device::inverse_down(second[wave], m, n) : 128.0;
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

// w = grad x, the default start of TGV's w.
template <typename T, typename S>
__global__ void gradient_kernel(Fields<T, S> F, int C, int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  for (int c = 0; c < C; ++c) {
    const T* x = F.x + c * N;
    F.w[(2 * c) * N + at] = device::forward_across<T>(x, i, j, W);
    F.w[(2 * c + 1) * N + at] = device::forward_down<T>(x, i, j, H, W);
  }
}

template <typename S>
__global__ void convert_kernel(const double* __restrict__ from, S* __restrict__ to, size_t count) {
  using Wide = decltype(device::widen(S{}));
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) to[k] = device::narrow<S>(static_cast<Wide>(from[k]));
}

// The duals of a start projected onto their balls (operators.cpp:378-425), and the tilde
// point made the start.
template <typename T, typename S>
__global__ void project_start_kernel(Fields<T, S> F, int C, int H, int W, int tgv, int coupled, T radius, T radius0) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const auto p_of = [&](int c, int e) { return static_cast<T>(device::widen(F.p[(2 * c + e) * N + at])); };
  const auto r_of = [&](int c, int e) { return static_cast<T>(device::widen(F.r[(3 * c + e) * N + at])); };
  T square = T(0), tensor_square = T(0);
  for (int c = 0; c < C && coupled; ++c) {
    const T a = p_of(c, 0), b = p_of(c, 1);
    square += a * a + b * b;
    if (tgv) tensor_square += r_of(c, 0) * r_of(c, 0) + r_of(c, 1) * r_of(c, 1) + T(2) * (r_of(c, 2) * r_of(c, 2));
  }
  for (int c = 0; c < C; ++c) {
    T scale = device::ball_scale(square, radius), tensor_scale = device::ball_scale(tensor_square, radius0);
    if (!coupled) {
      const T a = p_of(c, 0), b = p_of(c, 1);
      scale = device::ball_scale(a * a + b * b, radius);
      if (tgv) {
        tensor_scale = device::ball_scale(
            r_of(c, 0) * r_of(c, 0) + r_of(c, 1) * r_of(c, 1) + T(2) * (r_of(c, 2) * r_of(c, 2)), radius0);
      }
    }
    T projected_p[2];
    for (int e = 0; e < 2; ++e) projected_p[e] = p_of(c, e) * scale;
    for (int e = 0; e < 2; ++e) {
      const S value = device::narrow<S>(projected_p[e]);
      F.p[(2 * c + e) * N + at] = value;
      F.tp[(2 * c + e) * N + at] = value;
    }
    if (tgv) {
      for (int e = 0; e < 2; ++e) F.tw[(2 * c + e) * N + at] = F.w[(2 * c + e) * N + at];
      T projected[3];
      for (int e = 0; e < 3; ++e) projected[e] = r_of(c, e) * tensor_scale;
      for (int e = 0; e < 3; ++e) {
        const S value = device::narrow<S>(projected[e]);
        F.r[(3 * c + e) * N + at] = value;
        F.tr[(3 * c + e) * N + at] = value;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// Records, in binary64 (primal_dual.hpp:446-501, 791-895; frames.cpp:616-681).

// A work-group's sum (or largest value) of one value per lane, in a fixed order, for every
// lane: within each wavefront a butterfly (the pairs are the same in every lane, and a + b
// is b + a, so all lanes end with the same bits), then the wavefronts' results in order.
template <bool LARGEST>
__device__ double group_reduce(double value, double* room) {
#pragma unroll
  for (int mask = 32; mask > 0; mask >>= 1) {
    const double other = __shfl_xor(value, mask, 64);
    value = LARGEST ? fmax(value, other) : value + other;
  }
  const int waves = static_cast<int>(blockDim.x) >> 6;
  if ((threadIdx.x & 63) == 0) room[threadIdx.x >> 6] = value;
  __syncthreads();
  double total = room[0];
  for (int w = 1; w < waves; ++w) total = LARGEST ? fmax(total, room[w]) : total + room[w];
  __syncthreads();
  return total;
}

// The norms at one pixel of the tilde point (every channel): the first order (TV's
// gradient, or TGV's grad x~ - w~), TGV's |E w~|_F, the largest norm of div_2 r~ and the
// partial gap's sum over the channels of gamma |p~ + div_2 r~| (where PARTIAL: a solve with
// no partial gap leaves it 0 and reads no p~). Every load is made first.
template <typename T, typename S, bool TGV, int C, bool PARTIAL>
__device__ __forceinline__ void pixel_norms(const Fields<T, S>& F, int H, int W, int i, int j, const double* gammas,
                                            int coupled, double& first, double& second, double& most,
                                            double& residual) {
  const size_t N = static_cast<size_t>(H) * W;
  const size_t at = static_cast<size_t>(min(i, H - 1)) * W + min(j, W - 1);
  double a[C], b[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const T* x = F.tx + c * N;
    a[c] = device::forward_across_loaded<double>(x, i, j, H, W);
    b[c] = device::forward_down_loaded<double>(x, i, j, H, W);
    if (TGV) {
      a[c] = a[c] + -1.0 * static_cast<double>(F.tw[(2 * c) * N + at]);
      b[c] = b[c] + -1.0 * static_cast<double>(F.tw[(2 * c + 1) * N + at]);
    }
  }
  double squares = 0.0;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    squares += (coupled ? gammas[c] * gammas[c] : 1.0) * (a[c] * a[c] + b[c] * b[c]);
    if (!coupled) {
      first += gammas[c] * sqrt(squares);
      squares = 0.0;
    }
  }
  if (coupled) first = sqrt(squares);
  if constexpr (TGV) {
    double xx[C], yy[C], xy[C], dx[C], dy[C], pa[C], pb[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const T* wx = F.tw + (2 * c) * N;
      const T* wy = F.tw + (2 * c + 1) * N;
      xx[c] = device::backward_across_loaded<double>(wx, i, j, H, W);
      yy[c] = device::backward_down_loaded<double>(wy, i, j, H, W);
      xy[c] = device::backward_down_loaded<double>(wx, i, j, H, W);
      xy[c] += device::backward_across_loaded<double>(wy, i, j, H, W);
      xy[c] = 0.5 * xy[c];
      device::tensor_divergence_loaded<double>(F.tr + (3 * c) * N, F.tr + (3 * c + 1) * N, F.tr + (3 * c + 2) * N, i,
                                               j, H, W, dx[c], dy[c]);
      if (PARTIAL) {
        pa[c] = static_cast<double>(device::widen(F.tp[(2 * c) * N + at]));
        pb[c] = static_cast<double>(device::widen(F.tp[(2 * c + 1) * N + at]));
      }
    }
    squares = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      squares += (coupled ? gammas[c] * gammas[c] : 1.0) * (xx[c] * xx[c] + yy[c] * yy[c] + 2.0 * (xy[c] * xy[c]));
      if (!coupled) {
        second += gammas[c] * sqrt(squares);
        squares = 0.0;
      }
    }
    if (coupled) second = sqrt(squares);
    squares = 0.0;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      squares += 1.0 * (dx[c] * dx[c] + dy[c] * dy[c]);
      if (!coupled) {
        most = fmax(most, sqrt(squares));
        squares = 0.0;
      }
      if (PARTIAL) {
        const double ra = pa[c] + 1.0 * dx[c];
        const double rb = pb[c] + 1.0 * dy[c];
        residual += gammas[c] * sqrt(1.0 * (ra * ra + rb * rb));
      }
    }
    if (coupled) most = sqrt(squares);
  }
}

// The norms of the tilde point's pixels, each work-group's sums and largest into
// partials[4 group + 0..3] (first order, TGV's second order, the largest |div_2 r~|, the
// partial-gap residual, 0 unless `partial`).
//
// Work-group `group` takes 64 columns and kNormRows rows (a wavefront every fourth row of
// them, a lane's pixels summed down its column first), groups numbered across, then down,
// with `across` groups to a row of them.
constexpr int kNormRows = 8;

template <typename T, typename S, bool TGV>
__device__ __forceinline__ void norms_tile(const Fields<T, S>& F, int C, int H, int W, double g0, double g1,
                                           double g2, int coupled, int partial, double* __restrict__ partials,
                                           int across, int group) {
  __shared__ double room[256];
  const int j = (group % across) * 64 + (threadIdx.x & 63);
  const int top = (group / across) * kNormRows + (threadIdx.x >> 6);
  const double gammas[3] = {g0, g1, g2};
  double first = 0.0, second = 0.0, most = 0.0, residual = 0.0;
  const auto add = [&](auto channels, auto with_partial) {
    constexpr int kC = decltype(channels)::value;
    constexpr bool kPartial = decltype(with_partial)::value;
#pragma unroll
    for (int pass = 0; pass < kNormRows / 4; ++pass) {
      const int i = top + 4 * pass;
      double f = 0.0, g = 0.0, m = 0.0, r = 0.0;
      pixel_norms<T, S, TGV, kC, kPartial>(F, H, W, i, j, gammas, coupled, f, g, m, r);
      const bool here = i < H && j < W;
      first += here ? f : 0.0;
      second += here ? g : 0.0;
      most = fmax(most, here ? m : 0.0);
      residual += here ? r : 0.0;
    }
  };
  using One = std::integral_constant<int, 1>;
  using Three = std::integral_constant<int, 3>;
  using Partial = std::integral_constant<bool, TGV>;
  using NoPartial = std::integral_constant<bool, false>;
  if (C == 1) {
    partial ? add(One{}, Partial{}) : add(One{}, NoPartial{});
  } else {
    partial ? add(Three{}, Partial{}) : add(Three{}, NoPartial{});
  }
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

template <typename T, typename S, bool TGV>
__global__ __launch_bounds__(256) void norms_kernel(Fields<T, S> F, int C, int H, int W, double g0, double g1,
                                                    double g2, int coupled, int partial,
                                                    double* __restrict__ partials) {
  norms_tile<T, S, TGV>(F, C, H, W, g0, g1, g2, coupled, partial, partials, gridDim.x,
                        blockIdx.y * gridDim.x + blockIdx.x);
}

// G's terms at the kept coefficients of one component, a work-group's sum each: a
// work-group takes kValueSpan coefficients, each lane every 256th of them.
constexpr int kValueSpan = 4096;
inline size_t value_groups_of(size_t blocks) { return (64 * blocks + kValueSpan - 1) / kValueSpan; }

template <typename T>
__device__ __forceinline__ void value_tile(const Component& K, const T* zeta, double* __restrict__ partials,
                                           int group) {
  __shared__ double room[256];
  const size_t count = static_cast<size_t>(K.rows) * K.columns * 64;
  double term = 0.0;
#pragma unroll 4
  for (int run = 0; run < kValueSpan / 256; ++run) {
    const size_t index = static_cast<size_t>(group) * kValueSpan + static_cast<size_t>(run) * 256 + threadIdx.x;
    if (index < count) {
      const size_t block = index / 64;
      const int k = static_cast<int>(index % 64);
      term += device::value_of(K, K.levels[index], K.flat[block] != 0, k, static_cast<double>(zeta[index]));
    }
  }
  const double total = group_reduce<false>(term, room);
  if (threadIdx.x == 0) partials[group] = total;
}

template <typename T>
__global__ __launch_bounds__(256) void value_kernel(const Component* __restrict__ component, const T* zeta,
                                                    double* __restrict__ partials) {
  value_tile<T>(*component, zeta, partials, blockIdx.x);
}

// div p~ (from_r = 0) or div(-theta div_2 r~) (from_r = 1) of channel c at (i, j), in
// binary64: the divergence that the dual values take their xi from. Every load is made
// first (at clamped positions), the cases chosen after.
// div g at (i, j) for g = -theta div_2 r~, from div_2 r~ at (i, j) (dx, dy), its y part a
// row up and its x part a column left (each taken as 0 beyond the canvas).
__device__ __forceinline__ double g_divergence(double theta, double dx, double dy, double up_y, double left_x, int i,
                                               int j, int H, int W) {
  const double gx_here = -theta * dx, gy_here = -theta * dy;
  const double gy_up = i > 0 ? -theta * up_y : 0.0;
  const double gx_left = j > 0 ? -theta * left_x : 0.0;
  const bool top = i == 0, bottom = i == H - 1;
  const double down_part = (!top && !bottom) ? gy_here - gy_up : (!bottom ? gy_here : (!top ? -gy_up : 0.0));
  const double across_part = W == 1 ? 0.0 : (j == 0 ? gx_here : (j == W - 1 ? -gx_left : gx_here - gx_left));
  double div = down_part;
  div += across_part;
  return div;
}

template <typename T, typename S>
__device__ __forceinline__ double dual_divergence_at(const Fields<T, S>& F, int c, int i, int j, int H, int W,
                                                     int from_r, double theta) {
  const size_t N = static_cast<size_t>(H) * W;
  if (!from_r) return device::divergence_loaded<double>(F.tp + (2 * c) * N, F.tp + (2 * c + 1) * N, i, j, H, W);
  const S* rxx = F.tr + (3 * c) * N;
  const S* ryy = F.tr + (3 * c + 1) * N;
  const S* rxy = F.tr + (3 * c + 2) * N;
  // g = -theta div_2 r~, then div g = delta_y g_y + delta_x g_x.
  double dx, dy, up_x, up_y, left_x, left_y;
  device::tensor_divergence_loaded<double>(rxx, ryy, rxy, i, j, H, W, dx, dy);
  device::tensor_divergence_loaded<double>(rxx, ryy, rxy, i - 1, j, H, W, up_x, up_y);
  device::tensor_divergence_loaded<double>(rxx, ryy, rxy, i, j - 1, H, W, left_x, left_y);
  (void)up_x;
  (void)left_y;
  return g_divergence(theta, dx, dy, up_y, left_x, i, j, H, W);
}

// A lane's G* terms of one strip (and the free samples' terms of its cells), laid out as
// the prox's: a wavefront takes one row of blocks x 8 blocks across, a lane per cell
// column, so that the loads of each canvas row are coalesced stretches. The lane's xi down
// its column, the cell sums, the DCT down the column in registers, an LDS transpose, the
// DCT along the rows, the conjugate at each of the lane's 8 coefficients (row n of block
// b). DV and DH are the cell's shape, or 0 for the shape at run time (any up to kMaxCell).
template <typename T, typename S, int DV, int DH>
__device__ __forceinline__ double conjugate_terms(const Component& K, const Fields<T, S>& F, int c, int H, int W,
                                                  double gamma, int from_r, double theta, double radius, int g,
                                                  double (*room)[65]) {
  constexpr int kDown = DV > 0 ? DV : kMaxCell, kAcross = DH > 0 ? DH : kMaxCell;
  const int dv = DV > 0 ? DV : K.down, dh = DH > 0 ? DH : K.across;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int strips = (K.tiles_x + 7) >> 3;
  const int ty = (g / strips) * 4 + wave;
  const int b = lane >> 3, n = lane & 7;
  const int tx = (g % strips) * 8 + b;
  const bool live = (ty < K.tiles_y) & (tx < K.tiles_x);
  const bool inside = (ty < K.rows) & (tx < K.columns);
  const int first_column = (8 * tx + n) * dh;
  const int top = 8 * ty * dv;
  // The free samples' terms (frames.cpp:616-681; docs/math.md:911-949), <zeta, m> +
  // R |zeta|_1: within the blocks R |xi - mean of its cell|, beyond them 128 xi + R |xi|.
  //
  // With r~ (TGV) and a cell shape known at compile time, div_2 r~ is taken once a sample:
  // the row above's y parts carried down the lane's columns, the x part of the column on
  // the left from the lane on the left (lane 0 loads its own), the same values the
  // samples' own loads give.
  constexpr bool kCarry = DV > 0 && DH > 0;
  const size_t N = static_cast<size_t>(H) * W;
  const S* rxx = F.tr + (3 * c) * N;
  const S* ryy = F.tr + (3 * c + 1) * N;
  const S* rxy = F.tr + (3 * c + 2) * N;
  double up_y[kAcross];
  if (kCarry && from_r) {
#pragma unroll
    for (int bb = 0; bb < kAcross; ++bb) {
      double unused;
      device::tensor_divergence_loaded<double>(rxx, ryy, rxy, top - 1, first_column + bb, H, W, unused, up_y[bb]);
    }
  }
  double column[8], free_terms = 0.0;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    double xis[kDown * kAcross];
    if (kCarry && from_r) {
#pragma unroll
      for (int a = 0; a < kDown; ++a) {
        const int i = top + m * dv + a;
        double dx[kAcross], dy[kAcross], own_left, unused;
#pragma unroll
        for (int bb = 0; bb < kAcross; ++bb) {
          device::tensor_divergence_loaded<double>(rxx, ryy, rxy, i, first_column + bb, H, W, dx[bb], dy[bb]);
        }
        device::tensor_divergence_loaded<double>(rxx, ryy, rxy, i, first_column - 1, H, W, own_left, unused);
        const double from_left = __shfl_up(dx[kAcross - 1], 1, 64);
        const double left = (threadIdx.x & 63) == 0 ? own_left : from_left;
#pragma unroll
        for (int bb = 0; bb < kAcross; ++bb) {
          const double left_x = bb == 0 ? left : dx[bb - 1];
          xis[kAcross * a + bb] =
              gamma * g_divergence(theta, dx[bb], dy[bb], up_y[bb], left_x, i, first_column + bb, H, W);
          up_y[bb] = dy[bb];
        }
      }
    } else {
#pragma unroll
      for (int bb = 0; bb < kAcross; ++bb) {
#pragma unroll
        for (int a = 0; a < kDown; ++a) {
          if (a < dv && bb < dh) {
            const int i = top + m * dv + a, j = first_column + bb;
            xis[kAcross * a + bb] = gamma * dual_divergence_at(F, c, i, j, H, W, from_r, theta);
          }
        }
      }
    }
    double sum = 0.0;
#pragma unroll
    for (int bb = 0; bb < kAcross; ++bb) {
      double part = 0.0;
#pragma unroll
      for (int a = 0; a < kDown; ++a) {
        if (a < dv && bb < dh) part = a == 0 ? xis[kAcross * a + bb] : part + xis[kAcross * a + bb];
      }
      if (bb < dh) sum = bb == 0 ? part : sum + part;
    }
    column[m] = sum;
    if (K.free_samples) {
      const double mean = sum / static_cast<double>(dv * dh);
#pragma unroll
      for (int a = 0; a < kDown; ++a) {
#pragma unroll
        for (int bb = 0; bb < kAcross; ++bb) {
          if (a < dv && bb < dh) {
            const double xi = xis[kAcross * a + bb];
            free_terms += inside ? radius * fabs(xi - mean) : 128.0 * xi + radius * fabs(xi);
          }
        }
      }
    }
  }
  double y[8];
  device::forward8(column, y);
#pragma unroll
  for (int k = 0; k < 8; ++k) room[k][lane] = y[k];
  device::wave_sync();  // the wavefront's own room
  double row[8], coefficient[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) row[u] = room[n][8 * b + u];
  device::forward8(row, coefficient);
  const LevelRow own = level_row(K, ty, tx, n, inside);
  double term = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const double t = device::conjugate_of(K, own.q[u], own.flat, 8 * n + u, coefficient[u]);
    term += inside ? t : 0.0;
  }
  term += free_terms;
  return live ? term : 0.0;
}

// G_c* at xi = gamma div p~ (from_r = 0) or xi = gamma div(-theta div_2 r~) (from_r = 1),
// with the free samples' terms (frames.cpp:616-681): work-group `group`'s sum into
// partials, its tiles those of the prox's work-group `group`. ANY: the cell's shape is
// taken at run time (any up to kMaxCell); else it is one of 1 x 1, 2 x 2, 1 x 2 and 2 x 1,
// whose instance holds a third fewer registers (101 VGPRs against 150, with no scalar
// registers spilled).
template <typename T, typename S, bool ANY = false>
__device__ __forceinline__ void conjugate_tile(const Component& K, const Fields<T, S>& F, int c, int H, int W,
                                               double gamma, int from_r, const double* __restrict__ most,
                                               double alpha1, double radius, double* __restrict__ partials,
                                               int group) {
  __shared__ double room[4][8][65];
  __shared__ double sums[256];
  const int wave = threadIdx.x >> 6;
  double theta = 1.0;
  if (from_r) theta = *most > alpha1 ? alpha1 / *most : 1.0;
  // The cell's shape is the component's: the same in every wavefront of the work-group.
  const auto terms = [&](auto shape) {
    constexpr int kDV = decltype(shape)::first, kDH = decltype(shape)::second;
    return conjugate_terms<T, S, kDV, kDH>(K, F, c, H, W, gamma, from_r, theta, radius, group, room[wave]);
  };
  double term = 0.0;
  if constexpr (ANY) {
    term = terms(device::Shape<0, 0>{});
  } else if (K.down == 1 && K.across == 1) {
    term = terms(device::Shape<1, 1>{});
  } else if (K.down == 2 && K.across == 2) {
    term = terms(device::Shape<2, 2>{});
  } else if (K.down == 1 && K.across == 2) {
    term = terms(device::Shape<1, 2>{});
  } else {
    term = terms(device::Shape<2, 1>{});
  }
  const double total = group_reduce<false>(term, sums);
  if (threadIdx.x == 0) partials[group] = total;
}

template <typename T, typename S, bool ANY>
__global__ __launch_bounds__(256) void conjugate_kernel(const Component* __restrict__ component, Fields<T, S> F, int c,
                                                        int H, int W, double gamma, int from_r,
                                                        const double* __restrict__ most, double alpha1, double radius,
                                                        double* __restrict__ partials) {
  conjugate_tile<T, S, ANY>(*component, F, c, H, W, gamma, from_r, most, alpha1, radius, partials, blockIdx.x);
}

// Sums (or the largest of) count partials, one work-group per column of a row-major
// [count][columns] array: each lane takes every 256th in order, then a fixed tree.
__device__ __forceinline__ void reduce_tile(const double* __restrict__ partials, int count, int columns,
                                            unsigned largest_mask, double* __restrict__ out, int column) {
  __shared__ double room[256];
  const bool largest = (largest_mask >> column) & 1u;
  double value = 0.0;
  for (int k = threadIdx.x; k < count; k += 256) {
    const double next = partials[static_cast<size_t>(k) * columns + column];
    value = largest ? fmax(value, next) : value + next;
  }
  const double total = largest ? group_reduce<true>(value, room) : group_reduce<false>(value, room);
  if (threadIdx.x == 0) out[column] = total;
}

__global__ __launch_bounds__(256) void reduce_kernel(const double* __restrict__ partials, int count, int columns,
                                                     unsigned largest_mask, double* __restrict__ out) {
  reduce_tile(partials, count, columns, largest_mask, out, blockIdx.x);
}

// The final projection in binary64: the last canvas's coefficients (of the means of its
// cells) clipped to their intervals narrowed by `margin`, and the canvas moved by the
// change, so that the result's block DCT requantizes to the file's levels.
template <typename T>
__global__ __launch_bounds__(256) void final_kernel(const Component* __restrict__ component, const T* tx_field, int W,
                                                    double margin, double* __restrict__ canvas,
                                                    double* __restrict__ coefficients, double* __restrict__ excess) {
  __shared__ double first[4][64];
  __shared__ double second[4][64];
  const Component& K = *component;
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
        const double v = static_cast<double>(tx_field[static_cast<size_t>(i) * W + j]);
        vs[kMaxCell * a + b] = v;
        column = a == 0 ? v : column + v;
      }
    }
    if (b < dh) sum = b == 0 ? column : sum + column;
  }
  const double mean = cells > 1 ? sum / static_cast<double>(cells) : sum;
  first[wave][lane] = mean;
  __syncthreads();
  second[wave][lane] = device::forward_down(first[wave], m, n);
  __syncthreads();
  const double e = device::forward_along(second[wave], m, n);
  double zeta = 0.0;
  if (inside) {
    double lower, upper, centre, own_lower, own_upper;
    const size_t block = static_cast<size_t>(ty) * K.columns + tx;
    device::ends_of(K, K.levels[64 * block + lane], lane, lower, upper, centre, own_lower, own_upper);
    zeta = device::clamp_to(e, lower + margin, upper - margin);
    coefficients[64 * block + lane] = zeta;
    // How far the canvas's coefficient lies outside its interval, in steps (0 within it).
    if (excess != nullptr) excess[64 * block + lane] = (fmax(lower - e, 0.0) + fmax(e - upper, 0.0)) / K.steps[lane];
  }
  first[wave][lane] = zeta;
  __syncthreads();
  second[wave][lane] = device::inverse_along(first[wave], m, n);
  __syncthreads();
  const double back = device::inverse_down(second[wave], m, n);
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

// Device memory: its own (hipMalloc'd, freed with it), or a stretch of an Arena that it
// only points into.
struct DeviceBuffer {
  void* pointer = nullptr;
  bool owned = false;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes) {
    if (bytes > 0) {
      HIP_CHECK(hipMalloc(&pointer, bytes));
      owned = true;
    }
  }
  static DeviceBuffer within(void* stretch) {
    DeviceBuffer buffer;
    buffer.pointer = stretch;
    return buffer;
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& other) noexcept : pointer(other.pointer), owned(other.owned) {
    other.pointer = nullptr;
    other.owned = false;
  }
  DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
    std::swap(pointer, other.pointer);
    std::swap(owned, other.owned);
    return *this;
  }
  ~DeviceBuffer() {
    if (owned && pointer != nullptr) (void)hipFree(pointer);
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(pointer);
  }
};

// A fixed region of device memory handed out in order, each stretch 256-byte aligned: the
// buffers of a solve that may not allocate. Nothing is given back; the region is the
// caller's to reuse once the solve is done.
constexpr size_t kArenaAlign = 256;
inline size_t arena_stretch(size_t bytes) { return bytes == 0 ? 0 : (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign; }

struct Arena {
  char* base = nullptr;
  size_t capacity = 0;
  size_t used = 0;
  void* take(size_t bytes) {
    if (bytes == 0) return nullptr;
    const size_t stretch = arena_stretch(bytes);
    if (stretch > capacity - used) throw GpuError("the solve needs more device memory than its region holds");
    void* at = base + used;
    used += stretch;
    return at;
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
  uint64_t iterations = 0;  // the iterations between the two events
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
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(device::basis64), b.data(), sizeof(double) * 64));
  HIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(device::basis32), single, sizeof(float) * 64));
  devices.insert(device);
}

template <typename T, typename S>
class Solver {
 public:
  // On its own stream, or on `stream` (a batch's) where one is given.
  Solver(const FrameModel& frame, const Settings& settings, const Plan& plan, Timing timing,
         hipStream_t stream = nullptr)
      : frame_(frame), settings_(settings), plan_(plan), timing_(timing), stream_override_(stream) {
    if (stream_override_ == nullptr) own_stream_.emplace();
  }
  virtual ~Solver() = default;

  void run(const StartState* start, const Observer& observer, SolveOutput& out) {
    const auto began = std::chrono::steady_clock::now();
    const bool solving = prepare(start);
    hipStream_t s = stream();
    out.tau = plan_.tau;
    out.sigma = plan_.sigma;
    uint64_t iteration = 0;
    int32_t stop = 0;
    if (solving) {
      choose_dual_rows();
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
        // The observer sees every record after the first, the one that meets a tolerance
        // too, which then stops the solver as converged (rust/unround/src/pdhg.rs:645-665).
        bool observed = false;
        if (observer && at > 0) {
          tilde_canvas(canvas_room);
          observed = observer(record, canvas_room.data());
        }
        if (met || observed) {
          stop = met ? 1 : 3;
          stopped = true;
        }
      };
      if (every > 0 || most == 0) take(0);
      bool segment_open = false;
      const uint64_t untimed = timing_ == Timing::none ? 0 : untimed_iterations();
      while (!stopped && iteration < most) {
        const uint64_t next = iteration + 1;
        const bool keep = (every > 0 && next % every == 0) || next == most;
        // The first `untimed` iterations run untimed, straight before the timed ones.
        timed_ = timing_ != Timing::none && iteration >= untimed;
        // The events are read once the iterations between two records are done, so that
        // timing adds no wait between launches.
        if (timed_ && (timing_ == Timing::kernels || !segment_open)) {
          launches_.emplace_back();
          HIP_CHECK(hipEventRecord(launches_.back().begin, s));
          segment_open = true;
        }
        // The dual step of iteration `next` takes the proximal step of the one after it into
        // the same launch where the records do not read between them.
        const uint64_t after = next + 1;
        const bool keep_after = (every > 0 && after % every == 0) || after == most;
        iterate(keep, next < most && !keep, keep_after);
        if (timed_) launches_.back().iterations += 1;
        if (timed_ && timing_ == Timing::kernels) {
          HIP_CHECK(hipEventRecord(launches_.back().end, s));
          segment_open = false;
        }
        iteration = next;
        if ((every > 0 && iteration % every == 0) || iteration == most) {
          if (timing_ == Timing::iterations && segment_open) HIP_CHECK(hipEventRecord(launches_.back().end, s));
          segment_open = false;
          if (timing_ != Timing::none) collect_kernel_times(out.times);
          take(iteration);
        }
      }
      if (timing_ == Timing::iterations && segment_open) HIP_CHECK(hipEventRecord(launches_.back().end, s));
      if (timing_ != Timing::none) collect_kernel_times(out.times);
      if (stop == 0) stop = 2;
    }
    out.iterations = iteration;
    out.stop = solving ? stop : 0;
    primal_done_ = false;
    finish(out);
  }

  // The frame's sizes and the method's kind.
  void shape() {
    C_ = static_cast<int>(frame_.channels());
    H_ = static_cast<int>(frame_.rows);
    W_ = static_cast<int>(frame_.columns);
    N_ = static_cast<size_t>(H_) * W_;
    tgv_ = settings_.method == Method::tgv;
    coupled_ = settings_.coupled && C_ > 1;
    if (C_ != 1 && C_ != 3) throw GpuError("a frame of 1 or 3 channels");
  }

  // The device memory a solve from the default start takes, stretch by stretch as
  // allocate(), upload_components() and project_tilde() take it (each rounded up to an
  // Arena's alignment); after shape().
  size_t device_bytes() const {
    const size_t C = static_cast<size_t>(C_);
    size_t t_count = 4 * C * N_, s_count = 3 * 2 * C * N_;
    if (tgv_) {
      t_count += 3 * 2 * C * N_;
      s_count += 3 * 3 * C * N_;
    }
    size_t zeta_count = 0, value_groups = 0, conjugate_groups = 0;
    data_sizes(zeta_count, value_groups, conjugate_groups);
    const size_t norm_groups = static_cast<size_t>((W_ + 63) / 64) * ((H_ + kNormRows - 1) / kNormRows);
    size_t total = arena_stretch(sizeof(T) * (t_count + zeta_count) + sizeof(S) * s_count + 256);
    total += arena_stretch(sizeof(double) * (4 * norm_groups + value_groups + 2 * conjugate_groups + 16));
    total += arena_stretch(sizeof(double) * 16);
    for (int c = 0; c < C_; ++c) {
      const ComponentModel& k = frame_.components[c];
      total += arena_stretch(sizeof(int16_t) * std::max<size_t>(1, k.levels.size()));
      total += arena_stretch(std::max<size_t>(1, k.flat.size()));
      total += arena_stretch(sizeof(double) * std::max<size_t>(1, 64 * k.blocks()));
    }
    total += arena_stretch(sizeof(Component) * C);
    total += arena_stretch(sizeof(double) * C * N_);
    return total;
  }

  // Buffers from `arena` rather than hipMalloc from now on (null: hipMalloc again).
  void use_arena(Arena* arena) { arena_ = arena; }
  // The final projection's planes (the canvas cut to the picture, channel by channel,
  // binary64) copied into host memory at `planes` instead of into the output's canvas and
  // coefficients.
  void planes_into(double* planes) { planes_into_ = planes; }

  // Everything before the iterations: the device, the fields, the components and the start
  // (x from the coefficients given or the centres, w = grad x unless given, p and r given
  // or 0, projected onto their balls). Returns whether the method iterates at all.
  bool prepare(const StartState* start) {
    upload_basis();
    shape();
    const bool solving = settings_.method == Method::tv || settings_.method == Method::tgv;
    dual_rows_ = dual_rows_for(H_, W_);
    fusable_ = fusable();
    if (const char* rows = std::getenv("UNROUND_DUAL_ROWS")) dual_rows_ = std::max(1, std::atoi(rows));
    hipStream_t s = stream();
    allocate();
    start_primal(start);
    const dim3 pixels((W_ + 63) / 64, (H_ + 3) / 4);
    if (solving) {
      upload_field(start != nullptr ? start->p : nullptr, fields_.p, 2);
      if (tgv_) {
        if (start != nullptr && start->w != nullptr) {
          upload_field(start->w, fields_.w, 2);
        } else {
          gradient_kernel<T, S><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_);
          HIP_CHECK(hipGetLastError());
        }
        upload_field(start != nullptr ? start->r : nullptr, fields_.r, 3);
      }
      const T radius = static_cast<T>(tgv_ ? settings_.alpha1 : settings_.alpha);
      project_start_kernel<T, S><<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, tgv_, settings_.coupled ? 1 : 0, radius,
                                                        static_cast<T>(settings_.alpha0));
      HIP_CHECK(hipGetLastError());
    }
    return solving;
  }

 protected:
  template <typename, typename>
  friend class Batch;
  friend class Unrolled;

  // What depends on the data term: a JPEG frame's here, a picture of rounded samples' in
  // SamplesSolver.

  // The coefficients kept (zeta), and the work-groups of the kernels of G and G*.
  virtual void data_sizes(size_t& zeta, size_t& value_groups, size_t& conjugate_groups) const {
    zeta = value_groups = conjugate_groups = 0;
    for (int c = 0; c < C_; ++c) {
      zeta += 64 * frame_.components[c].blocks();
      value_groups += value_groups_of(frame_.components[c].blocks());
      conjugate_groups += static_cast<size_t>(prox_groups(c));
    }
  }

  // The data term onto the device, and x and x~ of the start: the canvas of the
  // coefficients given (or the centres) clipped to their intervals.
  virtual void start_primal(const StartState* start) {
    hipStream_t s = stream();
    upload_components();
    std::vector<DeviceBuffer> given(C_);
    for (int c = 0; c < C_; ++c) {
      const double* coefficients = start != nullptr && static_cast<size_t>(c) < start->coefficients.size()
                                       ? start->coefficients[c]
                                       : nullptr;
      double* device_given = nullptr;
      if (coefficients != nullptr) {
        const size_t bytes = sizeof(double) * 64 * frame_.components[c].blocks();
        given[c] = buffer(bytes);
        HIP_CHECK(hipMemcpyAsync(given[c].pointer, coefficients, bytes, hipMemcpyHostToDevice, s));
        device_given = given[c].as<double>();
      }
      start_kernel<T, S><<<(tiles(c) + 3) / 4, 256, 0, s>>>(component(c), fields_, c, H_, W_, device_given);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }

  int tiles(int c) const {
    const ComponentModel& k = frame_.components[c];
    return static_cast<int>((frame_.rows / (8 * k.down)) * (frame_.columns / (8 * k.across)));
  }

  // The prox kernel's work-groups: strips of 8 tiles across, 4 rows of tiles each.
  int prox_groups(int c) const {
    const ComponentModel& k = frame_.components[c];
    const int tiles_y = static_cast<int>(frame_.rows / (8 * k.down));
    const int tiles_x = static_cast<int>(frame_.columns / (8 * k.across));
    return ((tiles_x + 7) / 8) * ((tiles_y + 3) / 4);
  }

  const Component* component(int c) const { return components_.as<Component>() + c; }

  DeviceBuffer buffer(size_t bytes) {
    return arena_ != nullptr ? DeviceBuffer::within(arena_->take(bytes)) : DeviceBuffer(bytes);
  }

  void allocate() {
    const size_t C = static_cast<size_t>(C_);
    size_t t_count = 4 * C * N_;      // x, xbar, xbar_next, tx
    size_t s_count = 3 * 2 * C * N_;  // p, p_next, tp
    if (tgv_) {
      t_count += 3 * 2 * C * N_;  // w, w_next, tw
      s_count += 3 * 3 * C * N_;  // r, r_next, tr
    }
    size_t zeta_count = 0, value_groups = 0, conjugate_groups = 0;
    data_sizes(zeta_count, value_groups, conjugate_groups);
    const size_t t_bytes = sizeof(T) * (t_count + zeta_count);
    const size_t s_bytes = sizeof(S) * s_count;
    state_ = buffer(t_bytes + s_bytes + 256);
    HIP_CHECK(hipMemsetAsync(state_.pointer, 0, t_bytes + s_bytes + 256, stream()));
    T* at = state_.as<T>();
    const auto take = [&at](size_t n) {
      T* here = at;
      at += n;
      return here;
    };
    fields_ = Fields<T, S>{};
    fields_.x = take(C * N_);
    fields_.xbar = take(C * N_);
    fields_.xbar_next = take(C * N_);
    fields_.tx = take(C * N_);
    if (tgv_) {
      fields_.w = take(2 * C * N_);
      fields_.w_next = take(2 * C * N_);
      fields_.tw = take(2 * C * N_);
    }
    for (int c = 0; c < C_; ++c) fields_.zeta[c] = take(64 * frame_.components[c].blocks());
    S* dual = reinterpret_cast<S*>((reinterpret_cast<uintptr_t>(at) + 255) & ~static_cast<uintptr_t>(255));
    const auto take_dual = [&dual](size_t n) {
      S* here = dual;
      dual += n;
      return here;
    };
    fields_.p = take_dual(2 * C * N_);
    fields_.p_next = take_dual(2 * C * N_);
    fields_.tp = take_dual(2 * C * N_);
    if (tgv_) {
      fields_.r = take_dual(3 * C * N_);
      fields_.r_next = take_dual(3 * C * N_);
      fields_.tr = take_dual(3 * C * N_);
    }
    // Partials of the records: as many as the work-groups of their kernels.
    const size_t norm_groups = static_cast<size_t>((W_ + 63) / 64) * ((H_ + kNormRows - 1) / kNormRows);
    norm_partials_count_ = norm_groups;
    value_partials_count_ = value_groups;
    conjugate_partials_count_ = conjugate_groups;
    partials_ = buffer(sizeof(double) * (4 * norm_groups + value_groups + 2 * conjugate_groups + 16));
    record_out_ = buffer(sizeof(double) * 16);
  }

  void upload_components() {
    std::vector<Component> host(C_);
    levels_.clear();
    flats_.clear();
    for (int c = 0; c < C_; ++c) {
      const ComponentModel& k = frame_.components[c];
      if (k.down > static_cast<uint32_t>(kMaxCell) || k.across > static_cast<uint32_t>(kMaxCell)) {
        throw GpuError("a cell larger than 4 x 4");
      }
      levels_.push_back(buffer(sizeof(int16_t) * std::max<size_t>(1, k.levels.size())));
      flats_.push_back(buffer(std::max<size_t>(1, k.flat.size())));
      HIP_CHECK(hipMemcpyAsync(levels_.back().pointer, k.levels.data(), sizeof(int16_t) * k.levels.size(),
                               hipMemcpyHostToDevice, stream()));
      HIP_CHECK(hipMemcpyAsync(flats_.back().pointer, k.flat.data(), k.flat.size(), hipMemcpyHostToDevice,
                               stream()));
      Component& d = host[c];
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
    }
    components_ = buffer(sizeof(Component) * C_);
    HIP_CHECK(hipMemcpyAsync(components_.pointer, host.data(), sizeof(Component) * C_, hipMemcpyHostToDevice,
                             stream()));
    HIP_CHECK(hipStreamSynchronize(stream()));
  }

  // A field of the start, converted from binary64, or left at 0.
  template <typename F>
  void upload_field(const double* given, F* field, int entries) {
    if (given == nullptr) return;
    const size_t count = static_cast<size_t>(entries) * C_ * N_;
    DeviceBuffer staging = buffer(sizeof(double) * count);
    HIP_CHECK(hipMemcpyAsync(staging.pointer, given, sizeof(double) * count, hipMemcpyHostToDevice, stream()));
    convert_kernel<F><<<static_cast<unsigned>((count + 255) / 256), 256, 0, stream()>>>(staging.as<double>(),
                                                                                             field, count);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(stream()));
  }

  // Rows a wavefront of the dual kernel walks down, before a single solve measures its own
  // (choose_dual_rows), and where nothing is measured (the sweep, the unrolled solve's
  // pictures): long enough to spread its first rows' extra loads thin, short enough for
  // about 2048 wavefronts (two per SIMD); a power of two from 8 to 64. With BF16 duals a row
  // moves fewer bytes for the same arithmetic, and runs of at most 16 rows (more wavefronts
  // to hide it) measured best: 2048 x 2048 RGB 139 -> 117 us, 4096 x 4096 grey 181 -> 163 us.
  static int dual_rows_for(int H, int W) {
    const long strips = (W + kStripColumns - 1) / kStripColumns;
    const int most = sizeof(S) < sizeof(T) ? 16 : 64;
    int rows = 8;
    while (rows < most && strips * ((H + 2 * rows - 1) / (2 * rows)) >= 2048) rows *= 2;
    return rows;
  }

  // The dual step addresses each kind of field through one buffer resource where every
  // kind spans at most 2^31 bytes (planes::dual_fits), else with 64-bit addresses.
  bool narrow_spans() const { return planes::dual_buffers(static_cast<size_t>(C_), N_, sizeof(T), sizeof(S)); }
  // The prox step's choice, per component (x: one plane of T, p: two of S).
  bool prox_buffers() const { return planes::prox_buffers(N_, sizeof(T), sizeof(S)); }

  template <int C, bool TGV, bool COUPLED>
  void launch_dual(const Steps<T>& steps, int keep) {
    const int strips = (W_ + kStripColumns - 1) / kStripColumns;
    const int across = (strips + 3) / 4, down = (H_ + dual_rows_ - 1) / dual_rows_;
    const dim3 grid(across * down);
    if (narrow_spans() && keep == 0) {
      dual_kernel<T, S, C, TGV, COUPLED, device::BufferPlanes, false>
          <<<grid, 256, 0, stream()>>>(fields_, H_, W_, dual_rows_, across, steps, 0);
    } else if (narrow_spans()) {
      dual_kernel<T, S, C, TGV, COUPLED, device::BufferPlanes>
          <<<grid, 256, 0, stream()>>>(fields_, H_, W_, dual_rows_, across, steps, keep);
    } else {
      dual_kernel<T, S, C, TGV, COUPLED, device::WidePlanes>
          <<<grid, 256, 0, stream()>>>(fields_, H_, W_, dual_rows_, across, steps, keep);
    }
  }

  // Whether component c and the one after it share a launch: the same cell shape and the
  // same tiles (the chroma of a frame).
  bool shares_launch(int c) const {
    if (c + 1 >= C_) return false;
    const ComponentModel& a = frame_.components[c];
    const ComponentModel& b = frame_.components[c + 1];
    return a.down == b.down && a.across == b.across && tiles(c) == tiles(c + 1) && prox_groups(c) == prox_groups(c + 1);
  }

  // The proximal step of every component: components of one cell shape and size in one
  // launch, each launch timed on its own where timing asks.
  virtual void primal_step(bool keep) {
    hipStream_t s = stream();
    const T rho = static_cast<T>(plan_.relaxation), rest = static_cast<T>(1.0 - plan_.relaxation);
    const bool per_kernel = timing_kernels();
    for (int c = 0; c < C_;) {
      int count = 1;
      while (shares_launch(c + count - 1)) ++count;
      const ComponentModel& k = frame_.components[c];
      ProxSteps<T> steps{};
      for (int m = 0; m < count; ++m) {
        steps.descent[m] = static_cast<T>(plan_.tau * plan_.gamma[c + m]);
        steps.step[m] = static_cast<T>(plan_.tau / static_cast<double>(frame_.components[c + m].cells()));
      }
      if (per_kernel) {
        primal_events_.emplace_back();
        HIP_CHECK(hipEventRecord(primal_events_.back().begin, s));
      }
      const int groups = prox_groups(c), live = keep ? 1 : 0;
      const dim3 grid(static_cast<unsigned>(groups), static_cast<unsigned>(count));
      const auto prox = [&](auto kernel) {
        kernel<<<grid, 256, 0, s>>>(component(c), fields_, c, H_, W_, steps, rho, rest, live);
      };
      // Through buffer resources where a component's planes span at most 2^31 bytes, else
      // with 64-bit addresses.
      const bool buffers = prox_buffers();
      // The two common cell shapes with the defaults' data term (no slack cost, midpoint
      // centres) take the kernel that leaves the other cases out.
      bool plain = true;
      for (int m = 0; m < count; ++m) plain = plain && !frame_.components[c + m].costed && !frame_.components[c + m].mmse;
      if (k.down == 1 && k.across == 1 && buffers && plain) {
        prox(prox_kernel<T, S, 1, 1, device::BufferPlanes, true>);
      } else if (k.down == 2 && k.across == 2 && buffers && plain) {
        prox(prox_kernel<T, S, 2, 2, device::BufferPlanes, true>);
      } else if (k.down == 1 && k.across == 1) {
        buffers ? prox(prox_kernel<T, S, 1, 1, device::BufferPlanes>) : prox(prox_kernel<T, S, 1, 1, device::WidePlanes>);
      } else if (k.down == 2 && k.across == 2) {
        buffers ? prox(prox_kernel<T, S, 2, 2, device::BufferPlanes>) : prox(prox_kernel<T, S, 2, 2, device::WidePlanes>);
      } else if (k.down == 1 && k.across == 2) {
        buffers ? prox(prox_kernel<T, S, 1, 2, device::BufferPlanes>) : prox(prox_kernel<T, S, 1, 2, device::WidePlanes>);
      } else if (k.down == 2 && k.across == 1) {
        buffers ? prox(prox_kernel<T, S, 2, 1, device::BufferPlanes>) : prox(prox_kernel<T, S, 2, 1, device::WidePlanes>);
      } else {
        for (int m = 0; m < count; ++m) {
          prox_any_cell_kernel<T, S><<<groups, 256, 0, s>>>(component(c + m), fields_, c + m, H_, W_,
                                                            steps.descent[m], steps.step[m], rho, rest, live);
        }
      }
      HIP_CHECK(hipGetLastError());
      if (per_kernel) HIP_CHECK(hipEventRecord(primal_events_.back().end, s));
      c += count;
    }
  }

  // Whether the fused sweep takes this frame: asked for (UNROUND_SWEEP=1), and a layout it
  // has: one component, or three whose first has cells of one sample and the other two
  // cells of 1 x 1 or 2 x 2 (4:4:4, 4:2:0); binary32 iterates; buffer-resource planes. It
  // makes the same values as the two kernels, but measured slower (its tiles' registers
  // leave 2 or 3 wavefronts a SIMD), so the two kernels are the default.
  virtual bool fusable() const {
    if (sizeof(T) != sizeof(float)) return false;
    const char* sweep = std::getenv("UNROUND_SWEEP");
    if (sweep == nullptr || std::strcmp(sweep, "1") != 0) return false;
    if (!(narrow_spans() && prox_buffers())) return false;
    if (C_ == 1) return frame_.components[0].down == 1 && frame_.components[0].across == 1;
    const ComponentModel &y = frame_.components[0], &b = frame_.components[1], &r = frame_.components[2];
    if (y.down != 1 || y.across != 1 || b.down != r.down || b.across != r.across) return false;
    return (b.down == 1 && b.across == 1) || (b.down == 2 && b.across == 2);
  }

  // The dual step's scalars.
  Steps<T> dual_steps() const {
    Steps<T> steps{};
    for (int c = 0; c < C_; ++c) {
      steps.descent[c] = static_cast<T>(plan_.tau * plan_.gamma[c]);
      steps.ascent[c] = static_cast<T>(plan_.sigma * plan_.gamma[c]);
    }
    steps.radius = static_cast<T>(tgv_ ? settings_.alpha1 : settings_.alpha);
    steps.radius0 = static_cast<T>(settings_.alpha0);
    steps.rho = static_cast<T>(plan_.relaxation);
    steps.rest = static_cast<T>(1.0 - plan_.relaxation);
    return steps;
  }

  // The dual step of every channel, in the kernel of the frame's kind.
  void dual_step(const Steps<T>& steps, int keep) {
    if (C_ == 1) {
      tgv_ ? launch_dual<1, true, false>(steps, keep) : launch_dual<1, false, false>(steps, keep);
    } else if (coupled_) {
      tgv_ ? launch_dual<3, true, true>(steps, keep) : launch_dual<3, false, true>(steps, keep);
    } else {
      tgv_ ? launch_dual<3, true, false>(steps, keep) : launch_dual<3, false, false>(steps, keep);
    }
    HIP_CHECK(hipGetLastError());
  }

  // The dual step's rows per wavefront for the frame's shape and kernel, measured once a
  // process: 8, 16, 32 and 64 rows, each the mean of three launches after an untimed one
  // (the dual step only reads the state and writes the other buffers, which the first
  // iteration's dual step writes again, so trying it changes nothing). The measured best
  // differs from shape to shape and by kernel (2048 x 2048 RGB: 16 or 32 rows within 2 %;
  // 4096 x 4096 grey binary32: 64, with 32 the slowest; BF16 duals: 16). The environment's
  // UNROUND_DUAL_ROWS and the fused sweep keep their own.
  void choose_dual_rows() {
    if (std::getenv("UNROUND_DUAL_ROWS") != nullptr || fusable_) return;
    struct Chosen {
      std::mutex lock;
      std::map<std::vector<int>, int> rows;
    };
    static Chosen chosen;
    const std::vector<int> key{C_, H_, W_, static_cast<int>(sizeof(T)), static_cast<int>(sizeof(S)), tgv_ ? 1 : 0,
                               coupled_ ? 1 : 0, narrow_spans() ? 1 : 0};
    {
      std::lock_guard<std::mutex> hold(chosen.lock);
      const auto found = chosen.rows.find(key);
      if (found != chosen.rows.end()) {
        dual_rows_ = found->second;
        return;
      }
    }
    hipStream_t s = stream();
    const Steps<T> steps = dual_steps();
    // Each choice timed in two passes, in opposite orders, the smaller time kept: the first
    // launches of a solve run slower than the rest.
    constexpr int kChoices[4] = {8, 16, 32, 64};
    float least[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int pass = 0; pass < 2; ++pass) {
      for (int n = 0; n < 4; ++n) {
        const int k = pass == 0 ? n : 3 - n;
        dual_rows_ = kChoices[k];
        dual_step(steps, 0);
        EventPair timing;
        HIP_CHECK(hipEventRecord(timing.begin, s));
        for (int launch = 0; launch < 3; ++launch) dual_step(steps, 0);
        HIP_CHECK(hipEventRecord(timing.end, s));
        HIP_CHECK(hipEventSynchronize(timing.end));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, timing.begin, timing.end));
        least[k] = pass == 0 ? ms : std::min(least[k], ms);
      }
    }
    int best = 0;
    for (int k = 1; k < 4; ++k) best = least[k] < least[best] ? k : best;
    dual_rows_ = kChoices[best];
    std::lock_guard<std::mutex> hold(chosen.lock);
    chosen.rows[key] = dual_rows_;
  }

  // One iteration: its proximal step (unless the last launch made it), then its dual step,
  // alone or fused with the proximal step of the next iteration (`fuse`, the next keeping
  // its tilde point where keep_after).
  void iterate(bool keep, bool fuse, bool keep_after) {
    hipStream_t s = stream();
    const bool per_kernel = timing_kernels();
    if (!primal_done_) primal_step(keep);
    fuse = fuse && fusable_;
    if (per_kernel) dual_events_.emplace_back();
    const Steps<T> steps = dual_steps();
    // p~ is read by the records of TV and by the partial gap's only.
    const int k = keep ? kKeepTilde | (!tgv_ || plan_.has_partial ? kKeepDualP : 0) : 0;
    if (per_kernel) HIP_CHECK(hipEventRecord(dual_events_.back().begin, s));
    if (fuse) {
      if constexpr (sizeof(T) == sizeof(float)) {
        const int kept = keep_after ? 1 : 0;
        if (C_ == 1) {
          tgv_ ? launch_sweep<1, true, false>(steps, k, kept) : launch_sweep<1, false, false>(steps, k, kept);
        } else if (coupled_) {
          tgv_ ? launch_sweep<3, true, true>(steps, k, kept) : launch_sweep<3, false, true>(steps, k, kept);
        } else {
          tgv_ ? launch_sweep<3, true, false>(steps, k, kept) : launch_sweep<3, false, false>(steps, k, kept);
        }
      }
    } else {
      dual_step(steps, k);
    }
    HIP_CHECK(hipGetLastError());
    if (per_kernel) HIP_CHECK(hipEventRecord(dual_events_.back().end, s));
    std::swap(fields_.p, fields_.p_next);
    std::swap(fields_.w, fields_.w_next);
    std::swap(fields_.r, fields_.r_next);
    if (fuse) std::swap(fields_.xbar, fields_.xbar_next);
    primal_done_ = fuse;
  }

  // The fused sweep: the dual step of this iteration and the proximal step of the next
  // (keep_prox: that one's tilde point kept), in runs of whole bands.
  template <int C, bool TGV, bool COUPLED>
  void launch_sweep(const Steps<T>& steps, int keep_dual, int keep_prox) {
    ProxSteps<T> prox{};
    for (int c = 0; c < C_; ++c) {
      prox.descent[c] = static_cast<T>(plan_.tau * plan_.gamma[c]);
      prox.step[c] = static_cast<T>(plan_.tau / static_cast<double>(frame_.components[c].cells()));
    }
    const bool halves = C == 3 && frame_.components[1].down == 2;
    const int band = halves ? 16 : 8;
    const int rows = std::max(dual_rows_, band) / band * band;
    const int strips = (W_ + kStripColumns - 1) / kStripColumns;
    const int across = (strips + 3) / 4, down = (H_ + rows - 1) / rows;
    const dim3 grid(across * down);
    const auto go = [&](auto kernel) {
      kernel<<<grid, 256, 0, stream()>>>(component(0), fields_, H_, W_, rows, across, steps, prox, keep_dual,
                                         keep_prox);
    };
    if constexpr (C == 1) {
      go(sweep_kernel<T, S, 1, TGV, COUPLED, 1, 1, device::BufferPlanes>);
    } else if (halves) {
      go(sweep_kernel<T, S, 3, TGV, COUPLED, 2, 2, device::BufferPlanes>);
    } else {
      go(sweep_kernel<T, S, 3, TGV, COUPLED, 1, 1, device::BufferPlanes>);
    }
  }

  // The durations of the launches since the last collection, then their events freed.
  void collect_kernel_times(KernelTimes& times) {
    HIP_CHECK(hipStreamSynchronize(stream()));
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
      times.iterations += pair.iterations;
    }
    times.dual_rows = dual_rows_;
    primal_events_.clear();
    dual_events_.clear();
    launches_.clear();
  }

  // The record of the tilde point (primal_dual.hpp:446-501).
  Record record(uint64_t iteration) {
    hipStream_t s = stream();
    double* partials = partials_.as<double>();
    double* norm_partials = partials;
    double* value_partials = norm_partials + 4 * norm_partials_count_;
    double* conjugate_partials = value_partials + value_partials_count_;
    double* partial_partials = conjugate_partials + conjugate_partials_count_;
    double* results = record_out_.as<double>();
    const dim3 pixels((W_ + 63) / 64, (H_ + kNormRows - 1) / kNormRows);
    const double g[3] = {plan_.gamma[0], C_ > 1 ? plan_.gamma[1] : 0.0, C_ > 2 ? plan_.gamma[2] : 0.0};
    const int coupled = settings_.coupled ? 1 : 0;
    const int partial = plan_.has_partial ? 1 : 0;
    if (tgv_) {
      norms_kernel<T, S, true>
          <<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, g[0], g[1], g[2], coupled, partial, norm_partials);
    } else {
      norms_kernel<T, S, false>
          <<<pixels, 256, 0, s>>>(fields_, C_, H_, W_, g[0], g[1], g[2], coupled, partial, norm_partials);
    }
    HIP_CHECK(hipGetLastError());
    // results[0..3]: first, second, most, residual.
    reduce_kernel<<<4, 256, 0, s>>>(norm_partials, static_cast<int>(norm_partials_count_), 4, 1u << 2, results);
    HIP_CHECK(hipGetLastError());
    data_records(value_partials, conjugate_partials, partial_partials, results);
    double host[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIP_CHECK(hipMemcpyAsync(host, results, sizeof(double) * 7, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    Record record;
    record.iteration = iteration;
    const double data = data_value(host[4]);
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

  // The data term's parts of a record: G's sum into results[4], G* at the dual's xi (TV:
  // gamma div p~; TGV: the feasible dual's, from the largest norm of div_2 r~ in results[2])
  // into results[5] and, for TGV's partial gap, G* at gamma div p~ into results[6].
  virtual void data_records(double* value_partials, double* conjugate_partials, double* partial_partials,
                            double* results) {
    hipStream_t s = stream();
    size_t offset = 0;
    for (int c = 0; c < C_; ++c) {
      const size_t groups = value_groups_of(frame_.components[c].blocks());
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
        const int groups = prox_groups(c);
        const ComponentModel& k = frame_.components[c];
        const bool fixed = (k.down == 1 || k.down == 2) && (k.across == 1 || k.across == 2);
        (fixed ? conjugate_kernel<T, S, false> : conjugate_kernel<T, S, true>)<<<groups, 256, 0, s>>>(
            component(c), fields_, c, H_, W_, plan_.gamma[c], from_r, results + 2, settings_.alpha1, plan_.free_radius,
            into + at);
        HIP_CHECK(hipGetLastError());
        at += static_cast<size_t>(groups);
      }
      reduce_kernel<<<1, 256, 0, s>>>(into, static_cast<int>(conjugate_partials_count_), 1, 0u, result);
      HIP_CHECK(hipGetLastError());
    };
    conjugates(tgv_ ? 1 : 0, conjugate_partials, results + 5);
    if (tgv_ && plan_.has_partial) conjugates(0, partial_partials, results + 6);
  }

  // G of the sum its kernels reduced.
  virtual double data_value(double sum) const { return sum; }

  void tilde_canvas(std::vector<double>& canvas) {
    std::vector<T> values(static_cast<size_t>(C_) * N_);
    HIP_CHECK(hipMemcpyAsync(values.data(), fields_.tx, sizeof(T) * values.size(), hipMemcpyDeviceToHost,
                             stream()));
    HIP_CHECK(hipStreamSynchronize(stream()));
    canvas.assign(values.begin(), values.end());
  }

  // The final projection onto the data term's intervals, in binary64, into out.
  virtual void finish(SolveOutput& out) {
    // Narrowed by 2^-30 of a unit: beyond what any binary64 block DCT of the result rounds
    // by, so that the coefficients requantize to their levels whatever the ties' rule.
    project_tilde(out, std::ldexp(1.0, -30), nullptr);
  }

 public:
  // A canvas (C x H x W binary64) projected onto the intervals (docs/math.md 12;
  // rust/unround/src/frames.rs:408-432): each block's coefficients, of its cells' means,
  // clipped to their intervals and the cells moved by the change; and how far each
  // coefficient lay outside its interval, in steps (frames.rs:492-504). Of a Solver of
  // binary64 iterates.
  void project(const std::vector<double>& canvas, SolveOutput& out, std::vector<std::vector<double>>& excess) {
    static_assert(sizeof(T) == sizeof(double), "the projection is taken in binary64");
    C_ = static_cast<int>(frame_.channels());
    H_ = static_cast<int>(frame_.rows);
    W_ = static_cast<int>(frame_.columns);
    N_ = static_cast<size_t>(H_) * W_;
    tgv_ = false;
    if (canvas.size() != static_cast<size_t>(C_) * N_) throw GpuError("a canvas of the frame's size");
    allocate();
    upload_components();
    HIP_CHECK(hipMemcpyAsync(fields_.tx, canvas.data(), sizeof(double) * canvas.size(), hipMemcpyHostToDevice,
                             stream()));
    project_tilde(out, 0.0, &excess);
    out.iterations = 0;
    out.stop = 0;
  }

 protected:
  // The final projection of x~ into out (and, where asked, the excess of each coefficient).
  void project_tilde(SolveOutput& out, double margin, std::vector<std::vector<double>>* excess) {
    hipStream_t s = stream();
    DeviceBuffer canvas = buffer(sizeof(double) * C_ * N_);
    std::vector<DeviceBuffer> coefficients, excesses;
    out.coefficients.assign(planes_into_ != nullptr ? 0 : C_, {});
    for (int c = 0; c < C_; ++c) {
      const size_t count = std::max<size_t>(1, 64 * frame_.components[c].blocks());
      coefficients.push_back(buffer(sizeof(double) * count));
      excesses.push_back(buffer(excess != nullptr ? sizeof(double) * count : 0));
      final_kernel<T><<<(tiles(c) + 3) / 4, 256, 0, s>>>(component(c), fields_.tx + c * N_, W_, margin,
                                                         canvas.as<double>() + c * N_,
                                                         coefficients.back().as<double>(),
                                                         excesses.back().as<double>());
      HIP_CHECK(hipGetLastError());
    }
    if (planes_into_ != nullptr) {
      // The picture's rows of each channel, straight into the caller's planes.
      const size_t h = frame_.height, w = frame_.width;
      for (int c = 0; c < C_; ++c) {
        HIP_CHECK(hipMemcpy2DAsync(planes_into_ + c * h * w, sizeof(double) * w, canvas.as<double>() + c * N_,
                                   sizeof(double) * W_, sizeof(double) * w, h, hipMemcpyDeviceToHost, s));
      }
      HIP_CHECK(hipStreamSynchronize(s));
      return;
    }
    if (excess != nullptr) {
      excess->assign(C_, {});
      for (int c = 0; c < C_; ++c) {
        (*excess)[c].resize(64 * frame_.components[c].blocks());
        HIP_CHECK(hipMemcpyAsync((*excess)[c].data(), excesses[c].pointer, sizeof(double) * (*excess)[c].size(),
                                 hipMemcpyDeviceToHost, s));
      }
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
  Timing timing_;
  bool timed_ = false;  // whether the launches of the iteration under way are timed
  bool timing_kernels() const { return timing_ == Timing::kernels && timed_; }
  // Where timing is asked for: how many iterations run untimed first, directly before
  // the timed ones, so that those start at the GPU's running clocks
  // (UNROUND_TIMING_WARMUP; none by default).
  static uint64_t untimed_iterations() {
    const char* count = std::getenv("UNROUND_TIMING_WARMUP");
    return count == nullptr ? 0 : static_cast<uint64_t>(std::max(0LL, std::atoll(count)));
  }
  int C_ = 0, H_ = 0, W_ = 0;
  size_t N_ = 0;
  bool tgv_ = false, coupled_ = false;
  hipStream_t stream_override_ = nullptr;
  std::optional<StreamHolder> own_stream_;
  hipStream_t stream() const { return stream_override_ != nullptr ? stream_override_ : own_stream_->stream; }
  DeviceBuffer state_, components_, partials_, record_out_;
  std::vector<DeviceBuffer> levels_, flats_;
  Fields<T, S> fields_{};
  size_t norm_partials_count_ = 0, value_partials_count_ = 0, conjugate_partials_count_ = 0;
  std::deque<EventPair> primal_events_, dual_events_, launches_;
  // Rows a wavefront of the dual kernel walks down.
  int dual_rows_ = 32;
  Arena* arena_ = nullptr;
  double* planes_into_ = nullptr;
  // Whether the last launch made the next iteration's proximal step (the fused sweep), and
  // whether this frame takes the sweep at all.
  bool primal_done_ = false;
  bool fusable_ = false;
};

#include "batch.hpp"
#include "samples.hpp"
#include "unrolled.hpp"

// ---------------------------------------------------------------------------------------
// The region behind the solve that allocates nothing (solve_in_scratch): every GPU the
// library runs on holds its own copy, from the code object's load to the end of the
// process. In host code `scratch` names a host-side stand-in, not the device's copy: its
// device address is taken once per device with hipGetSymbolAddress, with that device
// current, and only that address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct ScratchOfDevice {
  std::mutex use;  // one solve at a time holds the region
  unsigned char* address = nullptr;
};

// The current device's region: its address and lock, looked up once per device.
ScratchOfDevice& scratch_of_current_device() {
  static std::mutex guard;
  static std::map<int, std::unique_ptr<ScratchOfDevice>> devices;
  int device = 0;
  HIP_CHECK(hipGetDevice(&device));
  const std::lock_guard<std::mutex> lock(guard);
  std::unique_ptr<ScratchOfDevice>& slot = devices[device];
  if (slot == nullptr) {
    auto made = std::make_unique<ScratchOfDevice>();
    void* address = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
    if (address == nullptr) throw GpuError("the scratch region has no device address");
    made->address = static_cast<unsigned char*>(address);
    slot = std::move(made);
  }
  return *slot;
}

template <typename T, typename S>
void solve_in_region(const FrameModel& frame, const Settings& settings, const Plan& plan, double* planes,
                     uint64_t& iterations, int32_t& stop, std::string& message, int32_t& status, bool& done) {
  // The solve's own stream would be an allocation of the runtime's: it takes the calling
  // thread's default stream.
  Solver<T, S> solver(frame, settings, plan, Timing::none, hipStreamPerThread);
  solver.shape();
  const size_t needed = solver.device_bytes();
  if (needed > kScratchBytes) {
    status = 1;
    message = "this input's solve needs " + std::to_string(needed) + " bytes of device memory, more than the " +
              std::to_string(kScratchBytes) + " of the region that unround_solve_scratch holds";
    done = false;
    return;
  }
  ScratchOfDevice& region = scratch_of_current_device();
  const std::lock_guard<std::mutex> hold(region.use);
  Arena arena{reinterpret_cast<char*>(region.address), kScratchBytes, 0};
  solver.use_arena(&arena);
  solver.planes_into(planes);
  SolveOutput out;
  solver.run(nullptr, Observer{}, out);
  iterations = out.iterations;
  stop = out.stop;
  done = true;
}

class UnrolledOnGpu final : public UnrolledSolve {
 public:
  UnrolledOnGpu(const std::vector<const FrameModel*>& frames, const std::vector<const Plan*>& plans,
                const Settings& settings, int iterations, int every, Timing timing)
      : solve_(frames, plans, settings, iterations, every, timing == Timing::kernels) {}
  int frames() const override { return solve_.frames(); }
  int height() const override { return solve_.height(); }
  int width() const override { return solve_.width(); }
  void constants(double* out) const override {
    const Plan& plan = solve_.plan();
    out[0] = plan.tau;
    out[1] = plan.sigma;
    out[2] = plan.gamma[0];
    out[3] = plan.relaxation;
  }
  void start_state(double* out) override { solve_.start_state(out); }
  void forward(const float* map, double alpha0, double alpha1, float* canvas, void* stream) override {
    solve_.forward(map, alpha0, alpha1, canvas, static_cast<hipStream_t>(stream));
  }
  void backward(const float* grad_canvas, float* grad_map, double* grad_alphas, void* stream) override {
    solve_.backward(grad_canvas, grad_map, grad_alphas, static_cast<hipStream_t>(stream));
  }
  bool kernel_times(double* out) override { return solve_.kernel_times(out); }

 private:
  Unrolled solve_;
};

}  // namespace

std::unique_ptr<UnrolledSolve> make_unrolled(const std::vector<const FrameModel*>& frames,
                                             const std::vector<const Plan*>& plans, const Settings& settings,
                                             int iterations, int every, Timing timing, std::string& message,
                                             int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return nullptr;
    }
    return std::make_unique<UnrolledOnGpu>(frames, plans, settings, iterations, every, timing);
  } catch (const Refused& refusal) {
    status = 3;
    message = refusal.what();
    return nullptr;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return nullptr;
  }
}

bool solve_on_gpu(const FrameModel& frame, const Settings& settings, const Plan& plan, const StartState* start,
                  const Observer& observer, Timing timing, SolveOutput& out, std::string& message, int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    if (settings.precision == Precision::dual) {
      Solver<double, double> solver(frame, settings, plan, timing);
      solver.run(start, observer, out);
    } else if (settings.precision == Precision::mixed) {
      Solver<float, bf16> solver(frame, settings, plan, timing);
      solver.run(start, observer, out);
    } else {
      Solver<float, float> solver(frame, settings, plan, timing);
      solver.run(start, observer, out);
    }
    return true;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

bool project_on_gpu(const FrameModel& frame, const Settings& settings, const std::vector<double>& canvas,
                    SolveOutput& out, std::vector<std::vector<double>>& excess, std::string& message,
                    int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    upload_basis();
    const Plan plan;
    Solver<double, double> solver(frame, settings, plan, Timing::none);
    solver.project(canvas, out, excess);
    return true;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

bool solve_samples_on_gpu(const SamplesModel& model, const Settings& settings, const Plan& plan,
                          const Observer& observer, Timing timing, SolveOutput& out, std::string& message,
                          int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    if (settings.precision == Precision::dual) {
      SamplesSolver<double, double> solver(model, settings, plan, timing);
      solver.run(nullptr, observer, out);
    } else if (settings.precision == Precision::mixed) {
      SamplesSolver<float, bf16> solver(model, settings, plan, timing);
      solver.run(nullptr, observer, out);
    } else {
      SamplesSolver<float, float> solver(model, settings, plan, timing);
      solver.run(nullptr, observer, out);
    }
    return true;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

bool solve_in_scratch(const FrameModel& frame, const Settings& settings, const Plan& plan, double* planes,
                      uint64_t& iterations, int32_t& stop, std::string& message, int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    upload_basis();
    bool done = false;
    if (settings.precision == Precision::dual) {
      solve_in_region<double, double>(frame, settings, plan, planes, iterations, stop, message, status, done);
    } else if (settings.precision == Precision::mixed) {
      solve_in_region<float, bf16>(frame, settings, plan, planes, iterations, stop, message, status, done);
    } else {
      solve_in_region<float, float>(frame, settings, plan, planes, iterations, stop, message, status, done);
    }
    return done;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

bool solve_batch_on_gpu(const std::vector<const FrameModel*>& frames, const std::vector<const Plan*>& plans,
                        const Settings& settings, std::vector<SolveOutput>& outs, std::vector<double>& latencies,
                        std::string& message, int32_t& status) {
  try {
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
      status = 6;
      message = "no HIP device: this library computes on an AMD GPU";
      return false;
    }
    if (settings.precision == Precision::dual) {
      Batch<double, double> batch(frames, plans, settings);
      batch.run(outs, latencies);
    } else if (settings.precision == Precision::mixed) {
      Batch<float, bf16> batch(frames, plans, settings);
      batch.run(outs, latencies);
    } else {
      Batch<float, float> batch(frames, plans, settings);
      batch.run(outs, latencies);
    }
    return true;
  } catch (const GpuError& error) {
    status = 6;
    message = std::string("a GPU error: ") + error.what();
    return false;
  }
}

}  // namespace unround
