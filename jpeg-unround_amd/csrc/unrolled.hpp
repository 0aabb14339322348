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
// The adjoint of one iteration runs in two kernels, as the forward does, each recomputing
// what it needs from the iteration's input state and x-:
//   1. the dual step's: the balls' Jacobians (p^, r^ and w- made again), the radius and
//      alpha0 sensitivities, grad x-, grad w, grad r and grad p but div's part, as a walk
//      down the rows; grad p after the iteration completed on the way in with div^T of the
//      last prox adjoint's tau g grad v;
//   2. the data term's block: grad x~ to grad zeta by the DCT, the clip's slope (from the
//      marks the run's recomputed prox left, a byte a coefficient), back by the inverse
//      DCT to grad v; grad x, and tau g grad v for the next dual adjoint.
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

// The adjoint's gradients of one frame: of the state (x, w, p, r).
struct UnrolledGrads {
  float* x;
  float* w;
  float* p;
  float* r;
};
// The planes between the adjoint's two kernels, and the sensitivities it adds up.
struct UnrolledWork {
  float* xbar;     // grad x-, from the dual step's adjoint to the prox's
  float* v;        // tau g grad v, from the prox's adjoint to the next dual adjoint (div's transpose)
  float* radius;   // += the radius's sensitivity, per sample
  float* alpha0;   // += alpha0's, per run of rows and column
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
  float* const w = u.work + static_cast<size_t>(f) * 2 * N;
  UnrolledBack v;
  v.s = UnrolledState{kept, kept + N, kept + 3 * N, kept + 5 * N};
  v.xbar = u.bars + f * N;
  v.after = UnrolledGrads{a, a + N, a + 3 * N, a + 5 * N};
  v.before = UnrolledGrads{b, b + N, b + 3 * N, b + 5 * N};
  v.work = UnrolledWork{w, w + N, u.sensitivity + f * N, u.sensitivity + (static_cast<size_t>(u.count) + f) * N};
  v.k = u.frames[f].k;
  return v;
}

// A wavefront's own columns in the dual step's adjoint: its 64 lanes take columns
// 60 s - 2 .. 60 s + 61 of strip s and write 60 s .. 60 s + 59; lanes 0-1 and 62-63 make
// again what the neighbouring strips make, the two columns each side its own need.
constexpr int kBackColumns = 60;

// 1. The adjoint of the dual step of one iteration, for one frame (grid y; the work-groups
// taken in ordered_frames' order): a wavefront walks rows first .. first + rows - 1 of its
// strip, a lane per column. Per row i it makes the forward's w-, p^ and r^ of row i + 1
// again (from the iteration's input state and x-, loading rows i + 1 and i + 2), the
// balls' Jacobians there (grad p^, grad r^ from the gradients after the iteration, the
// radius and alpha0 sensitivities added up), and then row i's gradients: grad x- = sigma g
// grad^T grad p^, grad w- = sigma g (E^T grad r^ - grad p^), grad w~ = rho grad w+ + 2
// grad w-, and grad w, grad p (less div's part) and grad r = (1 - rho) grad r+ + grad r^ +
// div_2^T (tau g grad w~) before the iteration. The gradient of p after the iteration is
// completed on the way in with div^T of the last prox adjoint's tau g grad v. Neighbours
// in the row come from the neighbouring lanes; the walk starts two rows above its first so
// that the rows it carries are whole there.
__global__ __launch_bounds__(64) void unrolled_dual_back_kernel(UnrolledFrames u, const float* __restrict__ radius,
                                                                int H, int W, int rows, int strips) {
  // No multiply-add is fused: the walk's row step is compiled once per slot and again for
  // the first and last rows, and every copy must round alike, so that a row's gradients do
  // not depend on which copy makes it (which the run length decides).
#pragma clang fp contract(off)
  // Neighbouring strips and runs, which read each other's edge columns and rows again,
  // share an XCD's L2.
  const device::FrameTile place = device::ordered_frames();
  const int f = place.frame;
  const int strip = place.tile % strips, run = place.tile / strips;
  const int lane = static_cast<int>(threadIdx.x);
  const int j = strip * kBackColumns - 2 + lane;
  const bool column = j >= 0 && j < W;
  const bool own = column && lane >= 2 && lane < 2 + kBackColumns;
  const int first = run * rows, end = min(first + rows, H);
  const size_t N = static_cast<size_t>(H) * W;
  const UnrolledBack v = back_of(u, f, N);
  const UnrolledSteps k = v.k;
  const float rest = 1.0f - k.rho;
  // The fields through buffer resources, addressed by one offset a row for all of them; a
  // load beyond the frame reads 0 and a store there is dropped. A frame's state, and its
  // gradients after and before the iteration, are each one resource of 8 planes (x, w_x,
  // w_y, p_x, p_y, r_xx, r_yy, r_xy: at most 2^31 bytes, as the solve's checks hold), and
  // the two work planes one more, so that the kernel's resources fit its scalar registers.
  // The default cache policy throughout: the neighbouring strips' and runs' walks read the
  // edge columns and rows again.
  using Planes = device::BufferPlanesOf<float, device::kTemporal, device::kTemporal>;
  constexpr int kW = 1, kP = 3, kR = 5;  // the first plane of w, p and r in a state
  const Planes state(v.s.x, 8, N), after(v.after.x, 8, N), before(v.before.x, 8, N);
  const Planes xbar(const_cast<float*>(v.xbar), 1, N), map(const_cast<float*>(radius), 1, N);
  const Planes work(v.work.xbar, 2, N);  // grad x- (plane 0), tau g grad v (plane 1)
  constexpr int kGradBar = 0, kTgv = 1;
  const auto index = [&](int i) { return static_cast<uint32_t>(i) * static_cast<uint32_t>(W) + static_cast<uint32_t>(j); };
  const auto inside = [&](int i) { return column && i >= 0 && i < H; };
  // w- of row i from w, p and r of the row and r of the row below.
  const auto w_bar = [&](int i, float w0, float w1, float p0, float p1, const float (&r)[3], const float (&below)[3],
                         float (&bar)[2]) {
    const float xx_right = from_right(r[0]), xy_right = from_right(r[2]);
    float dx = j + 1 < W ? xx_right - r[0] : 0.0f;
    dx += i + 1 < H ? below[2] - r[2] : 0.0f;
    float dy = j + 1 < W ? xy_right - r[2] : 0.0f;
    dy += i + 1 < H ? below[1] - r[1] : 0.0f;
    const float t0 = w0 + k.descent * (p0 + dx), t1 = w1 + k.descent * (p1 + dy);
    bar[0] = 2.0f * t0 - w0;
    bar[1] = 2.0f * t1 - w1;
  };

  // The rows the walk carries: r, x- and tau g grad v of row i + 1; w- of row i; grad p^,
  // grad r^, grad p and grad r after the iteration of row i; grad p^'s y of row i - 1 and
  // tau g grad w~ of row i - 1.
  const int top = first - 2;
  const bool in1 = inside(top + 1), in0 = inside(top);
  const uint32_t at1 = index(top + 1), at0 = index(top);
  // The loads of the step at row i (r, x- and tau g grad v of row i + 2; w, p, grad p, grad
  // r and the radius of row i + 1; grad w of row i), issued a step ahead of their use into
  // slot (i + 1 - top) % 3 of three that the rows take in turn; the step at row i reads
  // r, x- and tau g grad v of row i + 1 from the slot before (the row before's). No loaded
  // value is moved between registers while its load is in flight (the move would wait for
  // it, and with it for every access of the row). Beyond the run (`live` false) nothing is
  // read, with no branch.
  struct Loads {
    float r[3], xb, v, w[2], p[2], ap[2], ar[3], R, aw[2];
  };
  Loads slots[3];
  for (int e = 0; e < 3; ++e) slots[2].r[e] = state.load_if(in1, kR + e, at1);
  slots[2].xb = xbar.load_if(in1, 0, at1);
  slots[2].v = work.load_if(in1, kTgv, at1);
  float wb0[2];
  {
    const float r0[3] = {state.load_if(in0, kR, at0), state.load_if(in0, kR + 1, at0), state.load_if(in0, kR + 2, at0)};
    w_bar(top, state.load_if(in0, kW, at0), state.load_if(in0, kW + 1, at0), state.load_if(in0, kP, at0),
          state.load_if(in0, kP + 1, at0), r0,
          slots[2].r, wb0);
  }
  float gp0[2] = {0.0f, 0.0f}, gr0[3] = {0.0f, 0.0f, 0.0f}, ap0[2] = {0.0f, 0.0f}, ar0[3] = {0.0f, 0.0f, 0.0f};
  float gp_up = 0.0f, d_up[2] = {0.0f, 0.0f};
  float alpha0_sum = 0.0f;
  const auto fetch = [&](int i, bool live) {
    Loads l;
    const bool in2 = live && inside(i + 2), in1 = live && inside(i + 1), in0 = live && inside(i);
    const uint32_t at2 = index(i + 2), at1 = index(i + 1), at0 = index(i);
    for (int e = 0; e < 3; ++e) l.r[e] = state.load_if(in2, kR + e, at2);
    l.xb = xbar.load_if(in2, 0, at2);
    l.v = work.load_if(in2, kTgv, at2);
    for (int e = 0; e < 2; ++e) {
      l.w[e] = state.load_if(in1, kW + e, at1);
      l.p[e] = state.load_if(in1, kP + e, at1);
      l.ap[e] = after.load_if(in1, kP + e, at1);
      l.aw[e] = after.load_if(in0, kW + e, at0);
    }
    for (int e = 0; e < 3; ++e) l.ar[e] = after.load_if(in1, kR + e, at1);
    l.R = map.load_if(in1, 0, at1);
    return l;
  };
  slots[0] = fetch(top, true);
  // Row i, whose loads lie in slot PH = (i - top) % 3; it fetches row i + 1's into the next.
  const auto step = [&](auto phase, int i) {
    constexpr int PH = decltype(phase)::value;
    const int n1 = i + 1;
    slots[(PH + 1) % 3] = fetch(i + 1, i + 1 < end);
    const Loads& l = slots[PH];
    const Loads& prior = slots[(PH + 2) % 3];
    const float(&r1)[3] = prior.r;
    const float xb1 = prior.xb, v1 = prior.v;
    const float(&r2)[3] = l.r;
    const float xb2 = l.xb, v2 = l.v;
    const float w1x = l.w[0], w1y = l.w[1], p1x = l.p[0], p1y = l.p[1];
    float ap1[2] = {l.ap[0], l.ap[1]};
    const float(&ar1)[3] = l.ar;
    const float R = l.R;
    const float(&aw0)[2] = l.aw;
    // The forward's w-, p^ and r^ of row i + 1.
    float wb1[2];
    w_bar(n1, w1x, w1y, p1x, p1y, r1, r2, wb1);
    const float xb_right = from_right(xb1);
    const float g1 = j + 1 < W ? xb_right - xb1 : 0.0f;
    const float g2 = n1 + 1 < H ? xb2 - xb1 : 0.0f;
    const float ph[2] = {p1x + k.ascent * (g1 - wb1[0]), p1y + k.ascent * (g2 - wb1[1])};
    const float left[2] = {from_left(wb1[0]), from_left(wb1[1])};
    const float across_x = backward_across_of(wb1[0], left[0], j, W);
    const float down_y = backward_down_of(wb1[1], wb0[1], n1, H);
    const float xy = 0.5f * (backward_down_of(wb1[0], wb0[0], n1, H) + backward_across_of(wb1[1], left[1], j, W));
    const float rh[3] = {r1[0] + k.ascent * across_x, r1[1] + k.ascent * down_y, r1[2] + k.ascent * xy};
    // grad p after the iteration, with div^T of tau g grad v: div p = B_x p_x + B_y p_y.
    const float v_right = from_right(v1);
    ap1[0] += backward_transpose(v1, v_right, j, W);
    ap1[1] += backward_transpose(v1, v2, n1, H);
    // The balls' Jacobians at row i + 1: where p^ is outside, p~ = R p^ / n.
    const bool counted = own && n1 >= first && n1 < end;
    float gp1[2], gr1[3];
    {
      const float gp[2] = {k.rho * ap1[0], k.rho * ap1[1]};
      // One reciprocal of the norm (hardware, 1 ulp) for the three quotients: the gradients
      // are held to 1e-3 of binary64's, far from binary32's last places.
      const float n = sqrt(ph[0] * ph[0] + ph[1] * ph[1]);
      if (n > R) {
        const float inverse = __builtin_amdgcn_rcpf(n);
        const float dot = ph[0] * gp[0] + ph[1] * gp[1];
        const float along = dot * inverse, scale = R * inverse;
        for (int e = 0; e < 2; ++e) gp1[e] = scale * (gp[e] - ph[e] * (along * inverse));
        // Added where it lies (an atomic add with no return: nothing waits for it). One
        // lane adds to a sample in a launch, and the launches run in turn, so the sums
        // come in a fixed order.
        if (counted) {
          __hip_atomic_fetch_add(v.work.radius + index(n1), along, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      } else {
        gp1[0] = gp[0];
        gp1[1] = gp[1];
      }
      // r~ = P(r^) in the norm with xy counted twice.
      const float gr[3] = {k.rho * ar1[0], k.rho * ar1[1], k.rho * ar1[2]};
      const float nr = sqrt(rh[0] * rh[0] + rh[1] * rh[1] + 2.0f * (rh[2] * rh[2]));
      if (nr > k.alpha0) {
        const float inverse = __builtin_amdgcn_rcpf(nr);
        const float dot = rh[0] * gr[0] + rh[1] * gr[1] + rh[2] * gr[2];
        const float along = dot * inverse, scale = k.alpha0 * inverse;
        const float weight[3] = {1.0f, 1.0f, 2.0f};
        for (int e = 0; e < 3; ++e) gr1[e] = scale * (gr[e] - weight[e] * rh[e] * (along * inverse));
        if (counted) alpha0_sum += along;
      } else {
        gr1[0] = gr[0];
        gr1[1] = gr[1];
        gr1[2] = gr[2];
      }
    }
    // Row i's gradients. p^ = ... + sigma g (F_x x-, F_y x-).
    const float gpx_left = from_left(gp0[0]);
    const float gxbar = k.ascent * (forward_transpose(gpx_left, gp0[0], j, W) + forward_transpose(gp_up, gp0[1], i, H));
    // r^ = ... + sigma g (B_x w-_x, B_y w-_y, (B_y w-_x + B_x w-_y) / 2).
    const float grxx_right = from_right(gr0[0]), grxy_right = from_right(gr0[2]);
    const float e_x = backward_transpose(gr0[0], grxx_right, j, W) + 0.5f * backward_transpose(gr0[2], gr1[2], i, H);
    const float e_y = backward_transpose(gr0[1], gr1[1], i, H) + 0.5f * backward_transpose(gr0[2], grxy_right, j, W);
    const float bar[2] = {-k.ascent * gp0[0] + k.ascent * e_x, -k.ascent * gp0[1] + k.ascent * e_y};
    float d0[2], bw[2], bp[2];
    for (int e = 0; e < 2; ++e) {
      const float tilde = k.rho * aw0[e] + 2.0f * bar[e];
      bw[e] = rest * aw0[e] - bar[e] + tilde;
      bp[e] = rest * ap0[e] + gp0[e] + k.descent * tilde;
      d0[e] = k.descent * tilde;
    }
    // div_2 r = (F_x r_xx + F_y r_xy, F_x r_xy + F_y r_yy).
    const float dx_left = from_left(d0[0]), dy_left = from_left(d0[1]);
    const float xx = forward_transpose(dx_left, d0[0], j, W);
    const float yy = forward_transpose(d_up[1], d0[1], i, H);
    const float txy = forward_transpose(d_up[0], d0[0], i, H) + forward_transpose(dy_left, d0[1], j, W);
    const bool writes = own && i >= first;
    const uint32_t at = index(i);
    work.store_if<Planes::kNext>(writes, kGradBar, at, gxbar);  // the prox adjoint reads it next
    for (int e = 0; e < 2; ++e) {
      before.store_if(writes, kW + e, at, bw[e]);
      before.store_if(writes, kP + e, at, bp[e]);
    }
    before.store_if(writes, kR, at, rest * ar0[0] + gr0[0] + xx);
    before.store_if(writes, kR + 1, at, rest * ar0[1] + gr0[1] + yy);
    before.store_if(writes, kR + 2, at, rest * ar0[2] + gr0[2] + txy);
    gp_up = gp0[1];
    d_up[0] = d0[0];
    d_up[1] = d0[1];
    for (int e = 0; e < 2; ++e) {
      gp0[e] = gp1[e];
      ap0[e] = ap1[e];
      wb0[e] = wb1[e];
    }
    for (int e = 0; e < 3; ++e) {
      gr0[e] = gr1[e];
      ar0[e] = ar1[e];
    }
  };
  // The first row before the loop, so that the loop is entered as it goes round (a fetch
  // in flight), and the rows three a turn.
  using Zero = std::integral_constant<int, 0>;
  using One = std::integral_constant<int, 1>;
  using Two = std::integral_constant<int, 2>;
  step(Zero{}, top);
  int i = top + 1;
#pragma unroll 1
  for (; i + 2 < end; i += 3) {
    step(One{}, i);
    step(Two{}, i + 1);
    step(Zero{}, i + 2);
  }
  if (i < end) step(One{}, i);
  if (i + 1 < end) step(Two{}, i + 1);
  // alpha0's sensitivity of the lane's column over the run's rows, into the run's slot.
  if (own) v.work.alpha0[static_cast<size_t>(run) * W + j] += alpha0_sum;
}

// 2. The adjoint of the prox of one iteration, for one frame (blockIdx.y), tiled as the
// forward's prox: a wavefront takes a row of 8 blocks (64 columns, a lane per column, 8
// rows in registers), the 8 x 8 transforms across lanes through LDS. The clip's slope is
// r = 1 / (1 + t) where the forward's prox marked the coefficient inside its interval
// (`inside`, [frame][N] bytes in the kept coefficients' layout), 0 elsewhere: grad x~ =
// 2 grad x- + rho grad x+, grad zeta = DCT grad x~, grad e = slope grad zeta, grad v =
// IDCT grad e; grad x = grad v - grad x- + (1 - rho) grad x+ before the iteration, and
// tau g grad v for the next dual adjoint's div^T.
__global__ __launch_bounds__(256) void unrolled_prox_back_kernel(UnrolledFrames u, const uint8_t* __restrict__ inside,
                                                                 int H, int W, int strips) {
#pragma clang fp contract(off)
  __shared__ float room[4][8][65];
  const int f = static_cast<int>(blockIdx.y);
  const size_t N = static_cast<size_t>(H) * W;
  const UnrolledBack views = back_of(u, f, N);
  const UnrolledSteps& k = views.k;
  const float step = u.frames[f].step;
  const Component& K = u.components[f];
  const int wave = static_cast<int>(threadIdx.x) >> 6, lane = static_cast<int>(threadIdx.x) & 63;
  const int strip = static_cast<int>(blockIdx.x) % strips;
  const int ty = (static_cast<int>(blockIdx.x) / strips) * 4 + wave;
  const int b = lane >> 3, n = lane & 7;
  const int tx = strip * 8 + b;
  const bool live = ty < K.rows && tx < K.columns;
  const int j = 8 * tx + n, top = 8 * ty;
  const auto at = [&](int i) { return static_cast<uint32_t>(i) * static_cast<uint32_t>(W) + static_cast<uint32_t>(j); };
  // Every load first (buffer resources: a lane beyond the blocks reads 0, with no branch),
  // then grad x~ of the lane's column, to grad zeta (the inverse DCT's transpose is the DCT).
  using Planes = device::BufferPlanesOf<float, device::kStreaming, device::kStreaming>;
  const Planes bars(views.work.xbar, 1, N), after(views.after.x, 1, N);
  float grad_bar[8], after_x[8], column[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    grad_bar[m] = bars.load_if(live, 0, at(top + m));
    after_x[m] = after.load_if(live, 0, at(top + m));
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) column[m] = 2.0f * grad_bar[m] + k.rho * after_x[m];
  // A lane's row n of block b: the marks and the slopes.
  const size_t block = live ? static_cast<size_t>(ty) * K.columns + tx : 0;
  const uint64_t marks = live ? *reinterpret_cast<const uint64_t*>(inside + f * N + 64 * block + 8 * n) : 0;
  const bool flat = live && K.flat[block] != 0;
  // The row's 8 weights loaded together, then the slopes.
  const double* weights = (flat ? K.flat_weights : K.weights) + 8 * n;
  double weight[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) weight[c] = weights[c];
  float slope[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const float t = step * static_cast<float>(weight[c]);
    slope[c] = ((marks >> (8 * c)) & 0xFFu) != 0 ? 1.0f / (1.0f + t) : 0.0f;
  }
  float y[8], row[8];
  device::forward8(column, y);
  // Each wavefront transposes through its own part of the room: no work-group barriers.
#pragma unroll
  for (int q = 0; q < 8; ++q) room[wave][q][lane] = y[q];
  device::wave_sync();
#pragma unroll
  for (int c = 0; c < 8; ++c) row[c] = room[wave][n][8 * b + c];
  device::wave_sync();
  float grad_zeta[8], grad_e[8];
  device::forward8(row, grad_zeta);
#pragma unroll
  for (int c = 0; c < 8; ++c) grad_e[c] = slope[c] * grad_zeta[c];
  // grad e to grad v (the DCT's transpose is the inverse DCT).
  device::inverse8(grad_e, row);
#pragma unroll
  for (int c = 0; c < 8; ++c) room[wave][n][8 * b + c] = row[c];
  device::wave_sync();
#pragma unroll
  for (int q = 0; q < 8; ++q) y[q] = room[wave][q][lane];
  float grad_v[8];
  device::inverse8(y, grad_v);
  const float rest = 1.0f - k.rho;
  const Planes before(views.before.x, 1, N), tgv(views.work.v, 1, N);
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    before.store_if(live, 0, at(top + m), grad_v[m] - grad_bar[m] + rest * after_x[m]);
    tgv.store_if<device::kTemporal>(live, 0, at(top + m), k.descent * grad_v[m]);  // the next launch reads it
  }
}

// The forward prox of every frame (frame blockIdx.y): x of the state relaxed into the next
// state, x- into `bars`; no tilde copies. MARKS: the clip's marks of every frame's
// coefficients also go to `inside` ([frame][N] bytes) for the adjoint (an instance of its
// own, with the registers the marks take). PLAIN: every frame's component has midpoint
// centres (the unrolled solve takes no slack cost).
template <bool MARKS, bool PLAIN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MARKS ? 3 : kProxWaves<float, 1, 1>))) void
unrolled_prox_kernel(UnrolledFrames u, int H, int W, uint8_t* inside) {
  const int f = static_cast<int>(blockIdx.y);
  const size_t N = static_cast<size_t>(H) * W;
  const UnrolledFrame scalars = u.frames[f];
  float* const now = frame_state(u.state, f, N);
  Fields<float, float> F{};
  F.x = now;
  F.xbar = u.bars + f * N;
  F.tx = F.xbar;  // not written: no tilde copies here
  F.p = now + 3 * N;
  prox_tile<float, float, 1, 1, device::BufferPlanes, PLAIN>(
      u.components + f, F, 0, nullptr, frame_state(u.next, f, N), H, W, scalars.k.descent, scalars.step,
      scalars.k.rho, 1.0f - scalars.k.rho, 0, device::ordered(blockIdx.x, gridDim.x), MARKS ? inside + f * N : nullptr);
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
           const Settings& settings, int iterations, int every, bool timed = false)
      : frames_(frames), plans_(plans), settings_(settings), iterations_(iterations), every_(std::max(1, every)),
        timed_(timed) {
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
          !planes::prox_fits(samples, sizeof(float), sizeof(float)) || !planes::fits(8 * samples * sizeof(float)))
        throw Refused("the unrolled solve takes frames of at most 2^31 / 32 samples");
    }
    if (settings_.method != Method::tgv) throw Refused("the unrolled solve is TGV's");
    const int B = static_cast<int>(frames_.size());
    hipStream_t s = stream_.stream;
    for (int f = 0; f < B; ++f) {
      solvers_.push_back(std::make_unique<Solver<float, float>>(*frames_[f], settings_, *plans_[f], Timing::none, s));
      solvers_.back()->prepare(nullptr);
    }
    // Every frame's state side by side (two of them, the iteration's input and output);
    // x-; the start, kept; checkpoints; a run's states, x- and clip marks; gradients; work
    // planes. The frames' components side by side, their level tables staying their
    // solvers'.
    const size_t state = 8 * N_;
    states_[0] = DeviceBuffer(sizeof(float) * B * state);
    states_[1] = DeviceBuffer(sizeof(float) * B * state);
    bars_ = DeviceBuffer(sizeof(float) * B * N_);
    start_ = DeviceBuffer(sizeof(float) * B * state);
    // The checkpoints after the start's (which is the start itself): runs 1 .. the last.
    checkpoints_ = DeviceBuffer(sizeof(float) * B * state * std::max((iterations_ - 1) / every_, 1));
    // A run's input states after its first (whose is its checkpoint's).
    run_states_ = DeviceBuffer(sizeof(float) * B * state * std::max(every_ - 1, 1));
    run_bars_ = DeviceBuffer(sizeof(float) * B * N_ * every_);
    run_inside_ = DeviceBuffer(static_cast<size_t>(B) * N_ * every_);
    grads_ = DeviceBuffer(sizeof(float) * B * state * 2);
    work_ = DeviceBuffer(sizeof(float) * B * N_ * 4);
    radius_ = DeviceBuffer(sizeof(float) * N_);
    // The dual step's rows per wavefront: about 8192 wavefronts over all the frames (one
    // channel a row: four times the wavefronts of three channels' 2048 for the bytes in
    // flight), unless the environment says.
    const long strips = (W_ + kStripColumns - 1) / kStripColumns;
    while (dual_rows_ < 64 && B * strips * ((H_ + 2 * dual_rows_ - 1) / (2 * dual_rows_)) >= 8192) dual_rows_ *= 2;
    if (const char* rows = std::getenv("UNROUND_DUAL_ROWS")) dual_rows_ = std::max(1, std::atoi(rows));
    if (const char* rows = std::getenv("UNROUND_BACK_ROWS")) back_rows_ = std::max(1, std::atoi(rows));
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
    release_timings();
    map_ = map;
    alpha0_ = static_cast<float>(alpha0);
    alpha1_ = static_cast<float>(alpha1);
    scale_kernel<<<static_cast<unsigned>((N_ + 255) / 256), 256, 0, s>>>(map, alpha1_, radius_.as<float>(), N_);
    HIP_CHECK(hipGetLastError());
    upload_scalars();
    const int B = frames();
    // The runs before the last, each iteration into the next state: the first state of each
    // run into its checkpoint, the others into the two states in turn (the steps only read
    // their input state, so the first run reads the kept start itself). The last run as the
    // backward recomputes a run, its states, x- and clip marks kept for it.
    const int last_run = (iterations_ - 1) / every_;
    float* now = checkpoint(0);
    float* last = now;  // with no iterations, the start
    int spare = 0;
    for (int it = 0; it < iterations_; ++it) {
      if (it == last_run * every_) {
        last = states_[spare].as<float>();
        run_into_slots(it, iterations_, now, last);
        break;
      }
      float* next = nullptr;
      if ((it + 1) % every_ == 0) {
        next = checkpoint((it + 1) / every_);
      } else {
        next = states_[spare].as<float>();
        spare = 1 - spare;
      }
      step(now, next, bars_.as<float>());
      now = next;
    }
    last_run_kept_ = true;
    // x of every frame: the first of its 8 planes.
    HIP_CHECK(hipMemcpy2DAsync(canvas, sizeof(float) * N_, last, sizeof(float) * 8 * N_, sizeof(float) * N_,
                               static_cast<size_t>(B), hipMemcpyDeviceToDevice, s));
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
    // The work planes: grad x- and tau g grad v per frame, 0 for the last iteration's div^T;
    // then the sensitivities, the radius's [B][N] and alpha0's [B][runs of rows][W].
    fill_kernel<<<static_cast<unsigned>((4 * B * N_ + 255) / 256), 256, 0, s>>>(work_.as<float>(), 0.0f, 4 * B * N_);
    HIP_CHECK(hipGetLastError());
    float* sensitivities = work_.as<float>() + 2 * B * N_;
    int parity = 0;  // which half of grads_ holds the gradients after the current iteration
    const int runs = (iterations_ + every_ - 1) / every_;
    for (int run = runs - 1; run >= 0; --run) {
      const int first = run * every_, last = std::min(iterations_, first + every_);
      // Recompute the run (the forward kept the last one's, for the first backward after it).
      if (run != runs - 1 || !last_run_kept_) run_into_slots(first, last, checkpoint(run), states_[0].as<float>());
      for (int it = last - 1; it >= first; --it) {
        UnrolledFrames u = frames_view();
        u.state = run_state(it - first, checkpoint(run));
        u.bars = run_bars_.as<float>() + static_cast<size_t>(it - first) * B * N_;
        u.after = g + static_cast<size_t>(parity) * all;
        u.before = g + static_cast<size_t>(1 - parity) * all;
        u.work = work_.as<float>();
        u.sensitivity = sensitivities;
        step_back(u, run_inside_.as<uint8_t>() + static_cast<size_t>(it - first) * B * N_);
        parity = 1 - parity;
      }
    }
    last_run_kept_ = false;  // the slots now hold the first run's
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

  // The seconds and launches since the last forward began of the dual step's adjoint, the
  // prox's adjoint and the forward dual step, where launches are timed.
  bool kernel_times(double* out) {
    if (!timed_) return false;
    HIP_CHECK(hipStreamSynchronize(stream_.stream));
    for (int kind = 0; kind < 3; ++kind) {
      out[2 * kind] = 0.0;
      out[2 * kind + 1] = 0.0;
    }
    for (const Timed& t : timings_) {
      float ms = 0.0f;
      HIP_CHECK(hipEventElapsedTime(&ms, t.begin, t.end));
      out[2 * t.kind] += 1e-3 * ms;
      out[2 * t.kind + 1] += 1.0;
    }
    return true;
  }

  ~Unrolled() { release_timings(); }

 private:
  // A timed launch: its kind (0 the dual step's adjoint, 1 the prox's, 2 the forward dual
  // step) and the events around it.
  struct Timed {
    int kind;
    hipEvent_t begin, end;
  };

  // `launch` on the solve's stream, between two events where launches are timed.
  template <typename Launch>
  void timed(int kind, Launch&& launch) {
    if (!timed_) {
      launch();
      return;
    }
    Timed t{kind, nullptr, nullptr};
    HIP_CHECK(hipEventCreate(&t.begin));
    HIP_CHECK(hipEventCreate(&t.end));
    timings_.push_back(t);
    HIP_CHECK(hipEventRecord(t.begin, stream_.stream));
    launch();
    HIP_CHECK(hipEventRecord(t.end, stream_.stream));
  }

  void release_timings() {
    for (const Timed& t : timings_) {
      (void)hipEventDestroy(t.begin);
      (void)hipEventDestroy(t.end);
    }
    timings_.clear();
  }

  // Work on `to` waits for what `from` has queued.
  static void join(hipStream_t from, hipStream_t to) {
    if (from == to) return;
    hipEvent_t event = nullptr;
    HIP_CHECK(hipEventCreateWithFlags(&event, hipEventDisableTiming));
    HIP_CHECK(hipEventRecord(event, from));
    HIP_CHECK(hipStreamWaitEvent(to, event, 0));
    HIP_CHECK(hipEventDestroy(event));
  }

  // The states of every frame at the start of run `index`: the kept start for the first.
  float* checkpoint(int index) {
    return index == 0 ? start_.as<float>()
                      : checkpoints_.as<float>() + static_cast<size_t>(index - 1) * frames() * 8 * N_;
  }

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

  // The input states of iteration k of a run (k from 0): the run's checkpoint `start` for
  // the first, the run's slots for the others.
  float* run_state(int k, float* start) {
    return k == 0 ? start : run_states_.as<float>() + static_cast<size_t>(k - 1) * frames() * 8 * N_;
  }

  // Iterations first .. last - 1 from the checkpoint `from` (which the steps only read),
  // each later iteration's input state into its own slot of the run, every iteration's x-
  // and clip marks beside; the last one's output into `out`.
  void run_into_slots(int first, int last, float* from, float* out) {
    const int B = frames();
    for (int it = first; it < last; ++it) {
      float* const next = it + 1 < last ? run_state(it + 1 - first, from) : out;
      step(run_state(it - first, from), next, run_bars_.as<float>() + static_cast<size_t>(it - first) * B * N_,
           run_inside_.as<uint8_t>() + static_cast<size_t>(it - first) * B * N_);
    }
  }

  // One forward iteration of every frame from the states at `now` into those at `next`:
  // the prox (x relaxed, x- into `bars`), then the dual step with the radius per pixel.
  void step(float* now, float* next, float* bars, uint8_t* inside = nullptr) {
    hipStream_t s = stream_.stream;
    UnrolledFrames u = frames_view();
    u.state = now;
    u.next = next;
    u.bars = bars;
    const unsigned B = static_cast<unsigned>(frames());
    const dim3 tiles(static_cast<unsigned>(solvers_[0]->prox_groups(0)), B);
    bool plain = true;
    for (const FrameModel* f : frames_) plain = plain && !f->components[0].costed && !f->components[0].mmse;
    if (inside != nullptr) {
      plain ? unrolled_prox_kernel<true, true><<<tiles, 256, 0, s>>>(u, H_, W_, inside)
            : unrolled_prox_kernel<true, false><<<tiles, 256, 0, s>>>(u, H_, W_, inside);
    } else {
      plain ? unrolled_prox_kernel<false, true><<<tiles, 256, 0, s>>>(u, H_, W_, nullptr)
            : unrolled_prox_kernel<false, false><<<tiles, 256, 0, s>>>(u, H_, W_, nullptr);
    }
    HIP_CHECK(hipGetLastError());
    const int strips = (W_ + kStripColumns - 1) / kStripColumns, across = (strips + 3) / 4;
    const int groups = across * ((H_ + dual_rows_ - 1) / dual_rows_);
    timed(2, [&] {
      unrolled_dual_kernel<<<dim3(static_cast<unsigned>(groups), B), 256, 0, s>>>(u, radius_.as<float>(), H_, W_,
                                                                                  dual_rows_, across);
      HIP_CHECK(hipGetLastError());
    });
  }

  // The adjoint of one iteration of every frame, from its input state and x- (u.state,
  // u.bars), the gradients after it (u.after) into those before it (u.before): the dual
  // step's adjoint, then the prox's.
  void step_back(const UnrolledFrames& u, const uint8_t* inside) {
    hipStream_t s = stream_.stream;
    const unsigned B = static_cast<unsigned>(frames());
    const int strips = (W_ + kBackColumns - 1) / kBackColumns, runs = (H_ + back_rows_ - 1) / back_rows_;
    timed(0, [&] {
      unrolled_dual_back_kernel<<<dim3(static_cast<unsigned>(strips * runs), B), 64, 0, s>>>(
          u, radius_.as<float>(), H_, W_, back_rows_, strips);
      HIP_CHECK(hipGetLastError());
    });
    const ComponentModel& c = frames_[0]->components[0];
    const int tile_strips = static_cast<int>((c.columns + 7) / 8), tile_rows = static_cast<int>((c.rows + 3) / 4);
    timed(1, [&] {
      unrolled_prox_back_kernel<<<dim3(static_cast<unsigned>(tile_strips * tile_rows), B), 256, 0, s>>>(
          u, inside, H_, W_, tile_strips);
      HIP_CHECK(hipGetLastError());
    });
  }

  const std::vector<const FrameModel*> frames_;
  const std::vector<const Plan*> plans_;
  const Settings settings_;
  const int iterations_, every_;
  int H_ = 0, W_ = 0;
  size_t N_ = 0;
  StreamHolder stream_;
  std::vector<std::unique_ptr<Solver<float, float>>> solvers_;
  DeviceBuffer states_[2], bars_, start_, checkpoints_, run_states_, run_bars_, run_inside_, grads_, work_, radius_;
  DeviceBuffer components_, scalars_;
  std::vector<UnrolledFrame> scalars_host_;
  int dual_rows_ = 8;
  int back_rows_ = 16;  // the rows a wavefront of the dual step's adjoint walks
  const float* map_ = nullptr;
  float alpha0_ = 0.0f, alpha1_ = 0.0f;
  bool forwarded_ = false;
  bool last_run_kept_ = false;  // the run slots hold the last run of the last forward
  const bool timed_;
  std::vector<Timed> timings_;
};
