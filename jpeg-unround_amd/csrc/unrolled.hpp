// Unrolled TGV iterations with their adjoint, for gradients of a loss on the last canvas
// with respect to alpha0, alpha1 and a per-pixel weight m(i, j) of the first-order term
// (included by solver.hip, inside its anonymous namespace, after Solver).
//
// The method's iteration on grey frames whose sides are multiples of 8 (one sample a
// cell), in binary32, with p's ball of radius alpha1 m(i, j) at pixel (i, j) (the
// reference has only a radius per channel, docs/math.md:544-566):
//
//   v = x + tau g div p; zeta = clip((DCT v + t c) / (1 + t), lower, upper); x~ = IDCT zeta
//   x- = 2 x~ - x;  x+ = rho x~ + (1 - rho) x
//   w~ = w + tau g (p + div_2 r);  w- = 2 w~ - w;  w+ = rho w~ + (1 - rho) w
//   p^ = p + sigma g (grad x- - w-);  p~ = P_{alpha1 m}(p^);  p+ = rho p~ + (1 - rho) p
//   r^ = r + sigma g E w-;  r~ = P_{alpha0}(r^) (|r|^2 = xx^2 + yy^2 + 2 xy^2);  r+ likewise
//
// The adjoint of one iteration runs in five kernels, each the transpose of a piece of the
// forward, recomputing what it needs from the iteration's input state and x-:
//   1. the balls' Jacobians (p^, r^ and w- recomputed), the radius and alpha0 sensitivities;
//   2. grad x- (the transpose of grad) and grad w- (of -1 and of E), grad w~, grad w, the
//      part of grad p known there;
//   3. grad r (div_2's transpose);
//   4. the data term's block: grad x~ to grad zeta by the DCT, the clip's slope, back by
//      the inverse DCT to grad v; grad x;
//   5. grad p's last part (div's transpose of tau g grad v).
// The forward keeps the state every `every` iterations; the backward recomputes each run of
// iterations from its checkpoint and walks it in reverse.

// What the unrolled solve does not take (a status 3 for the caller, not a GPU failure).
struct Refused : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The differences of one axis and their transposes, at index k of n: the forward difference
// F f(k) = f(k+1) - f(k) (0 at the end), the backward difference B of div and E
// (f(0) at the start, -f(n-2) at the end, 0 where n = 1), and their transposes from the
// neighbouring values of g.
__device__ __forceinline__ float forward_transpose(float g_before, float g_here, int k, int n) {
  return (k >= 1 ? g_before : 0.0f) - (k + 1 < n ? g_here : 0.0f);
}
__device__ __forceinline__ float backward_transpose(float g_here, float g_after, int k, int n) {
  return (n > 1 && k <= n - 2) ? g_here - g_after : 0.0f;
}

// The state of one frame: x, w (2), p (2), r (3) planes of N samples.
struct UnrolledState {
  float* x;
  float* w;
  float* p;
  float* r;
};

// The iteration's scalars.
struct UnrolledSteps {
  float descent;  // tau g
  float ascent;   // sigma g
  float alpha0;
  float rho;
};

// Per frame: the iteration's scalars and the prox's coefficient step (tau: one sample a cell).
struct UnrolledFrame {
  UnrolledSteps k;
  float step;
};

// Where every frame's planes lie for one launch (the launch's frame is blockIdx.z, or
// blockIdx.y for the block kernels): frame f's at each base plus f times its stride, a
// state being 8 planes of N samples [x, w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy].
struct UnrolledFrames {
  float* state;        // [B][8][N]: the iteration's input state
  float* next;         // [B][8][N]: the forward's output state
  float* bars;         // [B][N]: x-
  float* after;        // [B][8][N]: the gradients after the iteration
  float* before;       // [B][8][N]: and before it
  float* work;         // [B][9][N]: the planes between the adjoint's kernels
  float* sensitivity;  // [2][B][N]: the radius's, then alpha0's
  const UnrolledFrame* frames;
  const Component* components;
  int count;
};

__device__ __forceinline__ float* frame_state(float* base, int f, size_t N) {
  return base + static_cast<size_t>(f) * 8 * N;
}

// w- at (i, j) from w, p and r (r at (i, j + 1) and (i + 1, j) too); w~ into `tilde`.
__device__ __forceinline__ void w_bar_at(const float* w, const float* p, const float* r, size_t N, int i, int j, int H,
                                         int W, float descent, float tilde[2], float bar[2]) {
  float dx, dy;
  device::tensor_divergence_at<float>(r, r + N, r + 2 * N, i, j, H, W, dx, dy);
  const size_t at = static_cast<size_t>(i) * W + j;
  const float wx = w[at], wy = w[N + at];
  tilde[0] = wx + descent * (p[at] + dx);
  tilde[1] = wy + descent * (p[N + at] + dy);
  bar[0] = 2.0f * tilde[0] - wx;
  bar[1] = 2.0f * tilde[1] - wy;
}

// p^ and r^ at (i, j) from the state and x-.
__device__ __forceinline__ void ascents_at(const UnrolledState& s, const float* xbar, size_t N, int i, int j, int H,
                                           int W, const UnrolledSteps& k, float p_hat[2], float r_hat[3],
                                           float w_tilde[2]) {
  float bar[2], left[2] = {0.0f, 0.0f}, up[2] = {0.0f, 0.0f}, unused[2];
  w_bar_at(s.w, s.p, s.r, N, i, j, H, W, k.descent, w_tilde, bar);
  if (j > 0) w_bar_at(s.w, s.p, s.r, N, i, j - 1, H, W, k.descent, unused, left);
  if (i > 0) w_bar_at(s.w, s.p, s.r, N, i - 1, j, H, W, k.descent, unused, up);
  const size_t at = static_cast<size_t>(i) * W + j;
  const float g1 = device::forward_across<float>(xbar, i, j, W);
  const float g2 = device::forward_down<float>(xbar, i, j, H, W);
  p_hat[0] = s.p[at] + k.ascent * (g1 - bar[0]);
  p_hat[1] = s.p[N + at] + k.ascent * (g2 - bar[1]);
  const float across_x = backward_across_of(bar[0], left[0], j, W);
  const float down_y = backward_down_of(bar[1], up[1], i, H);
  const float xy = 0.5f * (backward_down_of(bar[0], up[0], i, H) + backward_across_of(bar[1], left[1], j, W));
  r_hat[0] = s.r[at] + k.ascent * across_x;
  r_hat[1] = s.r[N + at] + k.ascent * down_y;
  r_hat[2] = s.r[2 * N + at] + k.ascent * xy;
}

// The forward dual step of every frame (frame blockIdx.y) by the frames' dual step, p's
// ball of radius alpha1 m(i, j) at (i, j): w, p and r of the state into the next state.
__global__ __launch_bounds__(256) void unrolled_dual_kernel(UnrolledFrames u, const float* __restrict__ radius, int H,
                                                            int W, int rows, int groups_across) {
  const int f = static_cast<int>(blockIdx.y);
  const size_t N = static_cast<size_t>(H) * W;
  float* const now = frame_state(u.state, f, N);
  float* const later = frame_state(u.next, f, N);
  Fields<float, float> F{};
  F.xbar = u.bars + f * N;
  F.w = now + N;
  F.w_next = F.tw = later + N;
  F.p = now + 3 * N;
  F.p_next = F.tp = later + 3 * N;
  F.r = now + 5 * N;
  F.r_next = F.tr = later + 5 * N;
  const UnrolledSteps k = u.frames[f].k;
  Steps<float> steps{};
  steps.descent[0] = k.descent;
  steps.ascent[0] = k.ascent;
  steps.radius0 = k.alpha0;
  steps.rho = k.rho;
  steps.rest = 1.0f - k.rho;
  dual_tile<float, float, 1, true, false, device::BufferPlanes, true>(F, H, W, rows, groups_across, steps, 0,
                                                                      device::ordered(blockIdx.x, gridDim.x), radius);
}

// The adjoint's gradients of one frame: of the state (x, w, p, r), and the planes between
// its kernels.
struct UnrolledGrads {
  float* x;
  float* w;
  float* p;
  float* r;
};
struct UnrolledWork {
  float* p_hat;    // 2 planes
  float* r_hat;    // 3 planes
  float* xbar;     // 1
  float* w_tilde;  // 2
  float* v;        // 1: tau g grad v
  float* radius;   // += the radius's sensitivity
  float* alpha0;   // += alpha0's
};

// Frame f's views for the adjoint's kernels.
struct UnrolledBack {
  UnrolledState s;
  const float* xbar;
  UnrolledGrads after, before;
  UnrolledWork work;
  UnrolledSteps k;
};
__device__ __forceinline__ UnrolledBack back_of(const UnrolledFrames& u, int f, size_t N) {
  float* const kept = frame_state(u.state, f, N);
  float* const a = frame_state(u.after, f, N);
  float* const b = frame_state(u.before, f, N);
  float* const w = u.work + static_cast<size_t>(f) * 9 * N;
  UnrolledBack v;
  v.s = UnrolledState{kept, kept + N, kept + 3 * N, kept + 5 * N};
  v.xbar = u.bars + f * N;
  v.after = UnrolledGrads{a, a + N, a + 3 * N, a + 5 * N};
  v.before = UnrolledGrads{b, b + N, b + 3 * N, b + 5 * N};
  v.work = UnrolledWork{w,         w + 2 * N, w + 5 * N, w + 6 * N, w + 8 * N, u.sensitivity + f * N,
                        u.sensitivity + (static_cast<size_t>(u.count) + f) * N};
  v.k = u.frames[f].k;
  return v;
}

// 1. The balls' Jacobians: grad p^ and grad r^ from the gradients after the iteration
// (`after`), and the radius and alpha0 sensitivities added up.
__global__ __launch_bounds__(256) void unrolled_balls_back_kernel(UnrolledFrames u, const float* __restrict__ radius,
                                                                  int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const UnrolledBack v = back_of(u, static_cast<int>(blockIdx.z), N);
  const UnrolledState& s = v.s;
  const float* xbar = v.xbar;
  const UnrolledGrads& after = v.after;
  const UnrolledWork& work = v.work;
  const UnrolledSteps& k = v.k;
  float p_hat[2], r_hat[3], w_tilde[2];
  ascents_at(s, xbar, N, i, j, H, W, k, p_hat, r_hat, w_tilde);
  // p~ = P(p^): where outside the ball, p~ = R p^ / n.
  const float gp[2] = {k.rho * after.p[at], k.rho * after.p[N + at]};
  const float R = radius[at];
  const float n = sqrt(p_hat[0] * p_hat[0] + p_hat[1] * p_hat[1]);
  if (n > R) {
    const float dot = p_hat[0] * gp[0] + p_hat[1] * gp[1];
    for (int e = 0; e < 2; ++e) work.p_hat[e * N + at] = (R / n) * (gp[e] - p_hat[e] * (dot / (n * n)));
    work.radius[at] += dot / n;
  } else {
    for (int e = 0; e < 2; ++e) work.p_hat[e * N + at] = gp[e];
  }
  // r~ = P(r^) in the norm with xy counted twice.
  const float gr[3] = {k.rho * after.r[at], k.rho * after.r[N + at], k.rho * after.r[2 * N + at]};
  const float nr = sqrt(r_hat[0] * r_hat[0] + r_hat[1] * r_hat[1] + 2.0f * (r_hat[2] * r_hat[2]));
  if (nr > k.alpha0) {
    const float dot = r_hat[0] * gr[0] + r_hat[1] * gr[1] + r_hat[2] * gr[2];
    const float weight[3] = {1.0f, 1.0f, 2.0f};
    for (int e = 0; e < 3; ++e) {
      work.r_hat[e * N + at] = (k.alpha0 / nr) * (gr[e] - weight[e] * r_hat[e] * (dot / (nr * nr)));
    }
    work.alpha0[at] += dot / nr;
  } else {
    for (int e = 0; e < 3; ++e) work.r_hat[e * N + at] = gr[e];
  }
}

// 2. grad x- (sigma g grad^T grad p^), grad w- (-sigma g grad p^ + sigma g E^T grad r^),
// then grad w~, grad w, and grad p less div's part.
__global__ __launch_bounds__(256) void unrolled_bars_back_kernel(UnrolledFrames u, int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const UnrolledBack v = back_of(u, static_cast<int>(blockIdx.z), N);
  const UnrolledGrads& after = v.after;
  const UnrolledGrads& before = v.before;
  const UnrolledWork& work = v.work;
  const UnrolledSteps& k = v.k;
  const auto g = [&](const float* plane, int di, int dj) {
    const int a = i + di, b = j + dj;
    return (a >= 0 && a < H && b >= 0 && b < W) ? plane[static_cast<size_t>(a) * W + b] : 0.0f;
  };
  const float* px = work.p_hat;
  const float* py = work.p_hat + N;
  const float* rxx = work.r_hat;
  const float* ryy = work.r_hat + N;
  const float* rxy = work.r_hat + 2 * N;
  // p^ = ... + sigma g (F_x x-, F_y x-).
  work.xbar[at] =
      k.ascent * (forward_transpose(g(px, 0, -1), g(px, 0, 0), j, W) + forward_transpose(g(py, -1, 0), g(py, 0, 0), i, H));
  // r^ = ... + sigma g (B_x w-_x, B_y w-_y, (B_y w-_x + B_x w-_y) / 2).
  const float e_x = backward_transpose(g(rxx, 0, 0), g(rxx, 0, 1), j, W) +
                    0.5f * backward_transpose(g(rxy, 0, 0), g(rxy, 1, 0), i, H);
  const float e_y = backward_transpose(g(ryy, 0, 0), g(ryy, 1, 0), i, H) +
                    0.5f * backward_transpose(g(rxy, 0, 0), g(rxy, 0, 1), j, W);
  const float bar[2] = {-k.ascent * px[at] + k.ascent * e_x, -k.ascent * py[at] + k.ascent * e_y};
  const float rest = 1.0f - k.rho;
  for (int e = 0; e < 2; ++e) {
    const float tilde = k.rho * after.w[e * N + at] + 2.0f * bar[e];
    work.w_tilde[e * N + at] = tilde;
    before.w[e * N + at] = rest * after.w[e * N + at] - bar[e] + tilde;
    before.p[e * N + at] = rest * after.p[e * N + at] + work.p_hat[e * N + at] + k.descent * tilde;
  }
}

// 3. grad r: (1 - rho) grad r+ + grad r^ + div_2^T (tau g grad w~).
__global__ __launch_bounds__(256) void unrolled_r_back_kernel(UnrolledFrames u, int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const UnrolledBack v = back_of(u, static_cast<int>(blockIdx.z), N);
  const UnrolledGrads& after = v.after;
  const UnrolledGrads& before = v.before;
  const UnrolledWork& work = v.work;
  const UnrolledSteps& k = v.k;
  const auto g = [&](const float* plane, int di, int dj) {
    const int a = i + di, b = j + dj;
    return (a >= 0 && a < H && b >= 0 && b < W) ? k.descent * plane[static_cast<size_t>(a) * W + b] : 0.0f;
  };
  const float* dx = work.w_tilde;
  const float* dy = work.w_tilde + N;
  // div_2 r = (F_x r_xx + F_y r_xy, F_x r_xy + F_y r_yy).
  const float xx = forward_transpose(g(dx, 0, -1), g(dx, 0, 0), j, W);
  const float yy = forward_transpose(g(dy, -1, 0), g(dy, 0, 0), i, H);
  const float xy = forward_transpose(g(dx, -1, 0), g(dx, 0, 0), i, H) + forward_transpose(g(dy, 0, -1), g(dy, 0, 0), j, W);
  const float rest = 1.0f - k.rho;
  const float parts[3] = {xx, yy, xy};
  for (int e = 0; e < 3; ++e) before.r[e * N + at] = rest * after.r[e * N + at] + work.r_hat[e * N + at] + parts[e];
}

// 4. The data term's block: one wavefront per block, a lane per sample (and per
// coefficient). grad x~ = 2 grad x- + rho grad x+; grad zeta = DCT grad x~; through the
// clip's slope; grad v = IDCT; grad x = grad v - grad x- + (1 - rho) grad x+.
__global__ __launch_bounds__(256) void unrolled_prox_back_kernel(UnrolledFrames u, int H, int W) {
  __shared__ float first[4][64];
  __shared__ float second[4][64];
  const int f = static_cast<int>(blockIdx.y);
  const UnrolledBack views = back_of(u, f, static_cast<size_t>(H) * W);
  const UnrolledState& s = views.s;
  const UnrolledGrads& after = views.after;
  const UnrolledGrads& before = views.before;
  const UnrolledWork& work = views.work;
  const UnrolledSteps& k = views.k;
  const float step = u.frames[f].step;
  const Component& K = u.components[f];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane >> 3, n = lane & 7;
  const int tile = blockIdx.x * 4 + wave;
  const bool active = tile < K.rows * K.columns;
  const int ty = active ? tile / K.columns : 0, tx = active ? tile % K.columns : 0;
  const size_t N = static_cast<size_t>(H) * W;
  const int i = 8 * ty + m, j = 8 * tx + n;
  const size_t at = static_cast<size_t>(i) * W + j;
  // The forward's coefficient, its clip and slope.
  const float v = active ? s.x[at] + k.descent * device::divergence_at<float>(s.p, s.p + N, i, j, H, W) : 0.0f;
  first[wave][lane] = v;
  __syncthreads();
  second[wave][lane] = device::forward_down(first[wave], m, n);
  __syncthreads();
  const float e = device::forward_along(second[wave], m, n);
  float slope = 0.0f;
  if (active) {
    const size_t block = static_cast<size_t>(ty) * K.columns + tx;
    const int16_t q = K.levels[64 * block + lane];
    const bool flat = K.flat[block] != 0;
    double lower, upper, centre, own_lower, own_upper;
    device::ends_of(K, q, lane, lower, upper, centre, own_lower, own_upper);
    const float t = step * static_cast<float>(flat ? K.flat_weights[lane] : K.weights[lane]);
    const float r = 1.0f / (1.0f + t);
    const float z = (e + t * static_cast<float>(centre)) * r;
    slope = (z >= static_cast<float>(lower) && z <= static_cast<float>(upper)) ? r : 0.0f;
  }
  __syncthreads();
  // grad x~ to grad zeta (the inverse DCT's transpose is the DCT).
  const float grad_tilde = active ? 2.0f * work.xbar[at] + k.rho * after.x[at] : 0.0f;
  first[wave][lane] = grad_tilde;
  __syncthreads();
  second[wave][lane] = device::forward_down(first[wave], m, n);
  __syncthreads();
  const float grad_e = slope * device::forward_along(second[wave], m, n);
  __syncthreads();
  // grad e to grad v (the DCT's transpose is the inverse DCT).
  first[wave][lane] = grad_e;
  __syncthreads();
  second[wave][lane] = device::inverse_along(first[wave], m, n);
  __syncthreads();
  const float grad_v = device::inverse_down(second[wave], m, n);
  if (!active) return;
  before.x[at] = grad_v - work.xbar[at] + (1.0f - k.rho) * after.x[at];
  work.v[at] = k.descent * grad_v;
}

// 5. grad p += div^T (tau g grad v): div p = B_x p_x + B_y p_y.
__global__ __launch_bounds__(256) void unrolled_div_back_kernel(UnrolledFrames u, int H, int W) {
  const int j = blockIdx.x * 64 + (threadIdx.x & 63);
  const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= H || j >= W) return;
  const size_t N = static_cast<size_t>(H) * W, at = static_cast<size_t>(i) * W + j;
  const UnrolledBack v = back_of(u, static_cast<int>(blockIdx.z), N);
  const UnrolledGrads& before = v.before;
  const UnrolledWork& work = v.work;
  const float here = work.v[at];
  const float right = j + 1 < W ? work.v[at + 1] : 0.0f;
  const float below = i + 1 < H ? work.v[at + W] : 0.0f;
  before.p[at] += backward_transpose(here, right, j, W);
  before.p[N + at] += backward_transpose(here, below, i, H);
}

// The forward prox of every frame (frame blockIdx.y): x of the state relaxed into the next
// state, x- into `bars`; no tilde copies.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kProxWaves<float, 1, 1>))) void
unrolled_prox_kernel(UnrolledFrames u, int H, int W) {
  const int f = static_cast<int>(blockIdx.y);
  const size_t N = static_cast<size_t>(H) * W;
  const UnrolledFrame scalars = u.frames[f];
  float* const now = frame_state(u.state, f, N);
  Fields<float, float> F{};
  F.x = now;
  F.xbar = u.bars + f * N;
  F.tx = F.xbar;  // not written: no tilde copies here
  F.p = now + 3 * N;
  prox_tile<float, float, 1, 1, device::BufferPlanes>(u.components + f, F, 0, nullptr, frame_state(u.next, f, N), H,
                                                      W, scalars.k.descent, scalars.step, scalars.k.rho,
                                                      1.0f - scalars.k.rho, 0, device::ordered(blockIdx.x, gridDim.x));
}

__global__ void scale_kernel(const float* __restrict__ map, float factor, float* __restrict__ out, size_t count) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) out[k] = factor * map[k];
}

__global__ void fill_kernel(float* __restrict__ out, float value, size_t count) {
  const size_t k = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k < count) out[k] = value;
}

// The loss gradients of the parameters from the sensitivities: per block of samples, the
// sum of alpha0's, the sum of m x the radius's (over all frames); and the map's gradient
// alpha1 x the frames' sum of the radius's. Partial sums in binary64, summed in order on
// the host.
__global__ __launch_bounds__(256) void unrolled_parameters_kernel(const float* __restrict__ alpha0_sens,
                                                                  const float* __restrict__ radius_sens,
                                                                  const float* __restrict__ map, float alpha1,
                                                                  int frames, size_t N, float* __restrict__ grad_map,
                                                                  double* __restrict__ partials) {
  __shared__ double room[256];
  const size_t k = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  double a0 = 0.0, a1 = 0.0;
  if (k < N) {
    double radius = 0.0;
    for (int f = 0; f < frames; ++f) {
      a0 += static_cast<double>(alpha0_sens[f * N + k]);
      radius += static_cast<double>(radius_sens[f * N + k]);
    }
    a1 = radius * static_cast<double>(map[k]);
    grad_map[k] = static_cast<float>(static_cast<double>(alpha1) * radius);
  }
  const double s0 = group_reduce<false>(a0, room);
  const double s1 = group_reduce<false>(a1, room);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s0;
    partials[2 * blockIdx.x + 1] = s1;
  }
}

class Unrolled {
 public:
  Unrolled(const std::vector<const FrameModel*>& frames, const std::vector<const Plan*>& plans,
           const Settings& settings, int iterations, int every)
      : frames_(frames), plans_(plans), settings_(settings), iterations_(iterations), every_(std::max(1, every)) {
    if (frames_.empty()) throw Refused("an unrolled solve of no frame");
    H_ = static_cast<int>(frames_[0]->rows);
    W_ = static_cast<int>(frames_[0]->columns);
    N_ = static_cast<size_t>(H_) * W_;
    for (const FrameModel* f : frames_) {
      if (f->channels() != 1 || f->components[0].down != 1 || f->components[0].across != 1)
        throw Refused("the unrolled solve takes grey frames");
      if (static_cast<int>(f->rows) != H_ || static_cast<int>(f->columns) != W_)
        throw Refused("the unrolled solve takes frames of one size");
      if (f->rows % 8 != 0 || f->columns % 8 != 0 || f->components[0].rows * 8 != f->rows ||
          f->components[0].columns * 8 != f->columns)
        throw Refused("the unrolled solve takes frames whose blocks cover them exactly");
      if (f->components[0].costed) throw Refused("the unrolled solve takes no slack cost");
      // Its kernels address a frame's planes through buffer resources only.
      const size_t samples = static_cast<size_t>(f->rows) * f->columns;
      if (!planes::dual_fits(1, samples, sizeof(float), sizeof(float)) ||
          !planes::prox_fits(samples, sizeof(float), sizeof(float)))
        throw Refused("the unrolled solve takes frames of at most 2^31 / 12 samples");
    }
    if (settings_.method != Method::tgv) throw Refused("the unrolled solve is TGV's");
    const int B = static_cast<int>(frames_.size());
    hipStream_t s = stream_.stream;
    for (int f = 0; f < B; ++f) {
      solvers_.push_back(std::make_unique<Solver<float, float>>(*frames_[f], settings_, *plans_[f], Timing::none, s));
      solvers_.back()->prepare(nullptr);
    }
    // Every frame's state side by side (two of them, the iteration's input and output);
    // x-; the start, kept; checkpoints; a run's states and x-; gradients; work planes. The
    // frames' components side by side, their level tables staying their solvers'.
    const size_t state = 8 * N_;
    states_[0] = DeviceBuffer(sizeof(float) * B * state);
    states_[1] = DeviceBuffer(sizeof(float) * B * state);
    bars_ = DeviceBuffer(sizeof(float) * B * N_);
    start_ = DeviceBuffer(sizeof(float) * B * state);
    checkpoints_ = DeviceBuffer(sizeof(float) * B * state * (iterations_ / every_ + 1));
    run_states_ = DeviceBuffer(sizeof(float) * B * state * every_);
    run_bars_ = DeviceBuffer(sizeof(float) * B * N_ * every_);
    grads_ = DeviceBuffer(sizeof(float) * B * state * 2);
    work_ = DeviceBuffer(sizeof(float) * B * N_ * 11);
    radius_ = DeviceBuffer(sizeof(float) * N_);
    // The dual step's rows per wavefront: about 8192 wavefronts over all the frames (one
    // channel a row: four times the wavefronts of three channels' 2048 for the bytes in
    // flight), unless the environment says.
    const long strips = (W_ + kStripColumns - 1) / kStripColumns;
    while (dual_rows_ < 64 && B * strips * ((H_ + 2 * dual_rows_ - 1) / (2 * dual_rows_)) >= 8192) dual_rows_ *= 2;
    if (const char* rows = std::getenv("UNROUND_DUAL_ROWS")) dual_rows_ = std::max(1, std::atoi(rows));
    components_ = DeviceBuffer(sizeof(Component) * B);
    scalars_ = DeviceBuffer(sizeof(UnrolledFrame) * B);
    for (int f = 0; f < B; ++f) {
      HIP_CHECK(hipMemcpyAsync(components_.as<Component>() + f, solvers_[f]->component(0), sizeof(Component),
                               hipMemcpyDeviceToDevice, s));
      keep_start(f, start_.as<float>() + f * state);
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }

  int frames() const { return static_cast<int>(frames_.size()); }
  int height() const { return H_; }
  int width() const { return W_; }
  const Plan& plan() const { return *plans_[0]; }

  // The start state [frame][x, w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy][H][W] into host memory.
  void start_state(double* out) {
    std::vector<float> values(static_cast<size_t>(frames()) * 8 * N_);
    HIP_CHECK(hipMemcpy(values.data(), start_.pointer, sizeof(float) * values.size(), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < values.size(); ++k) out[k] = values[k];
  }

  // The forward: `iterations` iterations from the start with radius alpha1 m(i, j) for p and
  // alpha0 for r; the last x of every frame into `canvas` (device, [frame][H][W]).
  void forward(const float* map, double alpha0, double alpha1, float* canvas, hipStream_t caller) {
    hipStream_t s = stream_.stream;
    join(caller, s);
    map_ = map;
    alpha0_ = static_cast<float>(alpha0);
    alpha1_ = static_cast<float>(alpha1);
    scale_kernel<<<static_cast<unsigned>((N_ + 255) / 256), 256, 0, s>>>(map, alpha1_, radius_.as<float>(), N_);
    HIP_CHECK(hipGetLastError());
    upload_scalars();
    const int B = frames();
    const size_t all = static_cast<size_t>(B) * 8 * N_;
    int live = 0;
    copy(states_[live].as<float>(), start_.as<float>(), all);
    for (int it = 0; it < iterations_; ++it) {
      if (it % every_ == 0) copy(checkpoint(it / every_), states_[live].as<float>(), all);
      step(states_[live].as<float>(), states_[1 - live].as<float>(), bars_.as<float>());
      live = 1 - live;
    }
    // x of every frame: the first of its 8 planes.
    HIP_CHECK(hipMemcpy2DAsync(canvas, sizeof(float) * N_, states_[live].pointer, sizeof(float) * 8 * N_,
                               sizeof(float) * N_, static_cast<size_t>(B), hipMemcpyDeviceToDevice, s));
    forwarded_ = true;
    join(s, caller);
  }

  // The backward of the last forward: grad_canvas (device, [frame][H][W]) to the map's
  // gradient (device, [H][W]) and alpha0's and alpha1's (host).
  void backward(const float* grad_canvas, float* grad_map, double* grad_alphas, hipStream_t caller) {
    if (!forwarded_) throw GpuError("a backward before any forward");
    hipStream_t s = stream_.stream;
    join(caller, s);
    const int B = frames();
    const size_t state = 8 * N_, all = static_cast<size_t>(B) * state;
    float* g = grads_.as<float>();
    // The gradients after the last iteration: the canvas's for x, 0 for the rest.
    fill_kernel<<<static_cast<unsigned>((B * state + 255) / 256), 256, 0, s>>>(g, 0.0f, B * state);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy2DAsync(g, sizeof(float) * state, grad_canvas, sizeof(float) * N_, sizeof(float) * N_,
                               static_cast<size_t>(B), hipMemcpyDeviceToDevice, s));
    float* sensitivities = work_.as<float>() + B * N_ * 9;  // radius [B][N], then alpha0 [B][N]
    fill_kernel<<<static_cast<unsigned>((2 * B * N_ + 255) / 256), 256, 0, s>>>(sensitivities, 0.0f, 2 * B * N_);
    HIP_CHECK(hipGetLastError());
    int parity = 0;  // which half of grads_ holds the gradients after the current iteration
    const int runs = (iterations_ + every_ - 1) / every_;
    for (int run = runs - 1; run >= 0; --run) {
      const int first = run * every_, last = std::min(iterations_, first + every_);
      // Recompute the run, each iteration's input state into its own slot (the last one's
      // output, not needed, into a spare state) and its x- beside.
      float* const kept = run_states_.as<float>();
      copy(kept, checkpoint(run), all);
      for (int it = first; it < last; ++it) {
        float* const next = it + 1 < last ? kept + static_cast<size_t>(it + 1 - first) * all : states_[0].as<float>();
        step(kept + static_cast<size_t>(it - first) * all, next,
             run_bars_.as<float>() + static_cast<size_t>(it - first) * B * N_);
      }
      for (int it = last - 1; it >= first; --it) {
        UnrolledFrames u = frames_view();
        u.state = run_states_.as<float>() + static_cast<size_t>(it - first) * all;
        u.bars = run_bars_.as<float>() + static_cast<size_t>(it - first) * B * N_;
        u.after = g + static_cast<size_t>(parity) * all;
        u.before = g + static_cast<size_t>(1 - parity) * all;
        u.work = work_.as<float>();
        u.sensitivity = sensitivities;
        step_back(u);
        parity = 1 - parity;
      }
    }
    DeviceBuffer partials(sizeof(double) * 2 * ((N_ + 255) / 256));
    const unsigned groups = static_cast<unsigned>((N_ + 255) / 256);
    unrolled_parameters_kernel<<<groups, 256, 0, s>>>(sensitivities + B * N_, sensitivities, map_, alpha1_, B, N_,
                                                      grad_map, partials.as<double>());
    HIP_CHECK(hipGetLastError());
    std::vector<double> host(2 * groups);
    HIP_CHECK(hipMemcpyAsync(host.data(), partials.pointer, sizeof(double) * host.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    double a0 = 0.0, a1 = 0.0;
    for (unsigned k = 0; k < groups; ++k) {
      a0 += host[2 * k];
      a1 += host[2 * k + 1];
    }
    grad_alphas[0] = a0;
    grad_alphas[1] = a1;
    join(s, caller);
  }

 private:
  // Work on `to` waits for what `from` has queued.
  static void join(hipStream_t from, hipStream_t to) {
    if (from == to) return;
    hipEvent_t event = nullptr;
    HIP_CHECK(hipEventCreateWithFlags(&event, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(event, from));
    HIP_CHECK(hipStreamWaitEvent(to, event, 0));
    HIP_CHECK(hipEventDestroy(event));
  }

  // The states of every frame at checkpoint `index`.
  float* checkpoint(int index) { return checkpoints_.as<float>() + static_cast<size_t>(index) * frames() * 8 * N_; }

  void copy(float* to, const float* from, size_t count) {
    HIP_CHECK(hipMemcpyAsync(to, from, sizeof(float) * count, hipMemcpyDeviceToDevice, stream_.stream));
  }

  // The frames' scalars (alpha0 is the forward's) onto the device.
  void upload_scalars() {
    scalars_host_.assign(frames(), UnrolledFrame{});
    for (int f = 0; f < frames(); ++f) {
      const Plan& plan = *plans_[f];
      UnrolledFrame& k = scalars_host_[f];
      k.k.descent = static_cast<float>(plan.tau * plan.gamma[0]);
      k.k.ascent = static_cast<float>(plan.sigma * plan.gamma[0]);
      k.k.alpha0 = alpha0_;
      k.k.rho = static_cast<float>(plan.relaxation);
      k.step = static_cast<float>(plan.tau);  // one sample a cell
    }
    HIP_CHECK(hipMemcpyAsync(scalars_.pointer, scalars_host_.data(), sizeof(UnrolledFrame) * scalars_host_.size(),
                             hipMemcpyHostToDevice, stream_.stream));
  }

  UnrolledFrames frames_view() {
    UnrolledFrames u{};
    u.frames = scalars_.as<UnrolledFrame>();
    u.components = components_.as<Component>();
    u.count = frames();
    return u;
  }

  // The start state of frame f, from its solver's fields, to 8 planes.
  void keep_start(int f, float* to) {
    const Fields<float, float>& F = solvers_[f]->fields_;
    copy(to, F.x, N_);
    copy(to + N_, F.w, 2 * N_);
    copy(to + 3 * N_, F.p, 2 * N_);
    copy(to + 5 * N_, F.r, 3 * N_);
  }

  // One forward iteration of every frame from the states at `now` into those at `next`:
  // the prox (x relaxed, x- into `bars`), then the dual step with the radius per pixel.
  void step(float* now, float* next, float* bars) {
    hipStream_t s = stream_.stream;
    UnrolledFrames u = frames_view();
    u.state = now;
    u.next = next;
    u.bars = bars;
    const unsigned B = static_cast<unsigned>(frames());
    unrolled_prox_kernel<<<dim3(static_cast<unsigned>(solvers_[0]->prox_groups(0)), B), 256, 0, s>>>(u, H_, W_);
    HIP_CHECK(hipGetLastError());
    const int strips = (W_ + kStripColumns - 1) / kStripColumns, across = (strips + 3) / 4;
    const int groups = across * ((H_ + dual_rows_ - 1) / dual_rows_);
    unrolled_dual_kernel<<<dim3(static_cast<unsigned>(groups), B), 256, 0, s>>>(u, radius_.as<float>(), H_, W_,
                                                                                dual_rows_, across);
    HIP_CHECK(hipGetLastError());
  }

  // The adjoint of one iteration of every frame, from its input state and x- (u.state,
  // u.bars), the gradients after it (u.after) into those before it (u.before).
  void step_back(const UnrolledFrames& u) {
    hipStream_t s = stream_.stream;
    const unsigned B = static_cast<unsigned>(frames());
    const dim3 pixels((W_ + 63) / 64, (H_ + 3) / 4, B);
    unrolled_balls_back_kernel<<<pixels, 256, 0, s>>>(u, radius_.as<float>(), H_, W_);
    HIP_CHECK(hipGetLastError());
    unrolled_bars_back_kernel<<<pixels, 256, 0, s>>>(u, H_, W_);
    HIP_CHECK(hipGetLastError());
    unrolled_r_back_kernel<<<pixels, 256, 0, s>>>(u, H_, W_);
    HIP_CHECK(hipGetLastError());
    const unsigned blocks = static_cast<unsigned>(frames_[0]->components[0].blocks());
    unrolled_prox_back_kernel<<<dim3((blocks + 3) / 4, B), 256, 0, s>>>(u, H_, W_);
    HIP_CHECK(hipGetLastError());
    unrolled_div_back_kernel<<<pixels, 256, 0, s>>>(u, H_, W_);
    HIP_CHECK(hipGetLastError());
  }

  const std::vector<const FrameModel*> frames_;
  const std::vector<const Plan*> plans_;
  const Settings settings_;
  const int iterations_, every_;
  int H_ = 0, W_ = 0;
  size_t N_ = 0;
  StreamHolder stream_;
  std::vector<std::unique_ptr<Solver<float, float>>> solvers_;
  DeviceBuffer states_[2], bars_, start_, checkpoints_, run_states_, run_bars_, grads_, work_, radius_;
  DeviceBuffer components_, scalars_;
  std::vector<UnrolledFrame> scalars_host_;
  int dual_rows_ = 8;
  const float* map_ = nullptr;
  float alpha0_ = 0.0f, alpha1_ = 0.0f;
  bool forwarded_ = false;
};
