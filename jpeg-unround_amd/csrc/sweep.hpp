// The fused sweep (included by solver.hip, inside its anonymous namespace, after
// iterate.hpp): one launch makes the dual step of iteration n and the proximal step of
// iteration n + 1, band by band, so that p is read from HBM once an iteration.
//
// A work-group walks 4 strips of 64 columns (256 canvas columns) from row `first` to `end`,
// in bands of 8 Vmax rows (one row of blocks of every component). For each band it makes
// the dual step of the band's rows (DualWalk), then, after a barrier, the proximal step of
// the band for every component (prox_tile_at): v = x + tau gamma div p takes the p the dual
// step has just written, which this work-group's own stores left in the caches. The two
// samples of that p that another work-group writes in the same launch are made once more
// here, exactly as there: p_x of the column left of the work-group (by its first wavefront,
// every row) and p_y of the row above the run (once, in DualWalk::begin). Everything else
// the launch reads is the iteration's start state or x of the last proximal step; x- is
// read from one buffer and written to the other (the run below reads the row past its own
// end), p, w and r likewise; x is read and written in place, each sample by its own lane.
//
// Per canvas sample and iteration this moves the state once each way and x- once each way
// (73 B in binary32 TGV 4:2:0 against the two-kernel iteration's 81): p's second read and a
// launch are gone. It is taken only where asked for (UNROUND_SWEEP=1): on one MI355X it
// measured slower than the two kernels (2048 x 2048 RGB 4:2:0 TGV: 236 us a launch against
// 146.5 + 55 us; 4096 x 4096 grey: 352-407 us against 195 + 75 us). Its tiles hold 134
// (grey) to 247 (4:2:0) registers, which leave 2 or 3 wavefronts a SIMD where the grey
// dual step alone runs 8, and its walk and prox phases keep little in flight; the two
// kernels' next-launch stores meanwhile stay in the 256 MB cache, which serves p's second
// read. Each band's walk starts afresh (its first rows, and the row above for
// w-, loaded again, mostly from the caches), so that the walk's carried rows are not live
// through the band's prox step, whose tiles need the registers.
//
// The frame's components: one (C = 1), or three whose first has cells of one sample and
// the other two cells of CV x CH samples (1 or 2 each).

template <int CV, int CH>
struct ChromaCells {
  static constexpr int down = CV, across = CH;
};

// The band: one row of blocks of every component.
template <int C, int CV>
constexpr int kSweepBand = C == 1 ? 8 : 8 * CV;

// The wavefronts per SIMD the sweep is compiled for: the register budget its tiles fit
// without spilling to memory (three channels spill a few registers to AGPRs at 2).
template <int C>
constexpr int kSweepWaves = C == 1 ? 3 : 2;

template <typename T, typename S, int C, bool TGV, bool COUPLED, int CV, int CH, template <typename> class Planes>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kSweepWaves<C>))) void sweep_kernel(const Component* __restrict__ components, Fields<T, S> F, int H,
                                                    int W, int rows, int groups_across, Steps<T> steps,
                                                    ProxSteps<T> prox, int keep_dual, int keep_prox) {
  __shared__ EdgeRoom<T, C> room;
  __shared__ SweepRoom<T, C> sweep;
  __shared__ ProxRooms<T> rooms;
  __shared__ device::StagedArrays staged_arrays[C];
  constexpr int kBand = kSweepBand<C, CV>;
  const int g = device::ordered(blockIdx.x, gridDim.x);
  const int group = g % groups_across;
  const int first = (g / groups_across) * rows, end = min(first + rows, H);
  // The wavefront's index, the same in all its lanes (said so: what it chooses stays scalar).
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  // Every component's constants in LDS, read after the first barrier below.
#pragma unroll
  for (int c = 0; c < C; ++c) (void)device::stage(components[c], staged_arrays[c]);
  // The prox step reads the p this launch writes and writes the other buffer of x-.
  Fields<T, S> G = F;
  G.p = F.p_next;
  G.xbar = F.xbar_next;
  const int wg_first_column = group * 4 * kStripColumns;
  // The prox step of tile row ty, strip `strip` of component c, in the band from row
  // band_top; its left column's p_x is slot `slot` of the sweep's room.
  const auto prox_of = [&](auto shape, int which, int ty, int strip, int slot, int band_top) {
    constexpr int kDV = decltype(shape)::down, kDH = decltype(shape)::across;
    // The component is the same in every lane (it comes from the wavefront's index): said so,
    // its planes' addresses stay scalar.
    const int c = __builtin_amdgcn_readfirstlane(which);
    ProxEdges<T> edges;
    edges.left = &sweep.left[slot][0][c];
    edges.left_stride = C;
    edges.left_top = band_top;
    if (8 * ty * kDV == first && first > 0) {  // the tile's first row is the run's
      edges.up = &sweep.up[c][0];
      edges.up_first = wg_first_column;
    }
    const Component& K = components[c];
    const device::StagedArrays& a = staged_arrays[c];
    const PreStaged staged{
        device::Staged{a.steps, a.weights, a.flat_weights, a.shrink, a.costs, K.slack, K.mmse, K.costed}};
    // Entries chosen by comparison, not by a run-time index: an array indexed at run time
    // would be copied to private memory.
    const auto pick = [c](const auto& three) { return c == 0 ? three[0] : (c == 1 ? three[1] : three[2]); };
    prox_tile_at<T, S, kDV, kDH, Planes>(K, staged, rooms, G, c, pick(F.zeta), F.x, H, W, pick(prox.descent),
                                          pick(prox.step), steps.rho, steps.rest, keep_prox, ty, strip, edges);
  };
#pragma unroll 1
  for (int top = first; top < end; top += kBand) {
    const int bottom = min(top + kBand, end);
    {
      // The band's walk, which ends with it.
      DualWalk<T, S, C, TGV, COUPLED, Planes, false, true> walk(room, &sweep, F, H, W, top, bottom, group, steps,
                                                                keep_dual, nullptr);
      walk.begin_stretch(top == first);
      walk.walk();
    }
    // The band's p complete in the caches, and its edges in LDS, before any wavefront's
    // prox step reads them.
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if constexpr (C == 1) {
      // One tile row of 8 rows, a wavefront's own strip.
      prox_of(ChromaCells<1, 1>{}, 0, top / 8, group * 4 + wave, wave, top);
    } else {
      // Luma: CV tile rows of the wavefront's own strip.
#pragma unroll
      for (int t = 0; t < CV; ++t) prox_of(ChromaCells<1, 1>{}, 0, top / 8 + t, group * 4 + wave, wave, top);
      // A chroma tile strip is 64 CH canvas columns wide.
      if constexpr (CH == 1) {
        prox_of(ChromaCells<CV, CH>{}, 1, top / (8 * CV), group * 4 + wave, wave, top);
        prox_of(ChromaCells<CV, CH>{}, 2, top / (8 * CV), group * 4 + wave, wave, top);
      } else {
        // Two strips of 128 columns: wavefronts 0 and 1 take Cb's, 2 and 3 Cr's.
        const int half = wave & 1;
        prox_of(ChromaCells<CV, CH>{}, 1 + (wave >> 1), top / (8 * CV), group * 2 + half, 2 * half, top);
      }
    }
    // The band's LDS is written again by the next band's rows.
    __syncthreads();
  }
}
