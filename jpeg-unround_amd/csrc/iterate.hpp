// The kernels of one iteration (included by solver.hip, inside its anonymous namespace,
// after Fields, Steps and kMaxCell).
//
// Both kernels give each lane of a wavefront one column of a 64-column strip and walk
// down rows, so that every load of a row is one coalesced 256-byte stretch (binary32).
// A value of the neighbouring column comes from the neighbouring lane (a DPP move); of the
// row above, from the lane's own registers; only the lanes at a strip's two edges read a
// column beyond it from memory. Each field is then read about once per iteration.

// The value of the lane on the left (lane 0: its own) and on the right (lane 63: its own):
// DPP's whole-wavefront shifts by one lane (wave_shr:1, wave_shl:1), a VALU move of each
// 32-bit word with no trip through LDS; a lane with no source keeps its own word.
// lane_shifted_or gives a lane with no source `old` instead.
template <int CONTROL, typename T>
__device__ __forceinline__ T lane_shifted_or(T value, T old) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8);
  if constexpr (sizeof(T) == 4) {
    const int word = __builtin_bit_cast(int, value), kept = __builtin_bit_cast(int, old);
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(kept, word, CONTROL, 0xF, 0xF, false));
  } else {
    struct Words {
      int low, high;
    };
    Words w = __builtin_bit_cast(Words, value);
    const Words k = __builtin_bit_cast(Words, old);
    w.low = __builtin_amdgcn_update_dpp(k.low, w.low, CONTROL, 0xF, 0xF, false);
    w.high = __builtin_amdgcn_update_dpp(k.high, w.high, CONTROL, 0xF, 0xF, false);
    return __builtin_bit_cast(T, w);
  }
}
template <int CONTROL, typename T>
__device__ __forceinline__ T lane_shifted(T value) {
  return lane_shifted_or<CONTROL>(value, value);
}
constexpr int kWaveShiftRight = 0x138, kWaveShiftLeft = 0x130;  // DPP controls wave_shr:1, wave_shl:1
template <typename T>
__device__ __forceinline__ T from_left(T value) {
  return lane_shifted<kWaveShiftRight>(value);
}
// The value of the lane on the left, lane 0 taking `edge` (and on the right, lane 63
// taking `edge`): with no select on the lane, so that a load of `edge` is not moved into a
// branch of lane 0's own and waited for there.
template <typename T>
__device__ __forceinline__ T from_left_or(T value, T edge) {
  return lane_shifted_or<kWaveShiftRight>(value, edge);
}
template <typename T>
__device__ __forceinline__ T from_right_or(T value, T edge) {
  return lane_shifted_or<kWaveShiftLeft>(value, edge);
}
template <typename T>
__device__ __forceinline__ T from_right(T value) {
  return lane_shifted<kWaveShiftLeft>(value);
}
// The value of `value` in lane `source` (a constant), as a scalar.
template <typename T>
__device__ __forceinline__ T lane_value(T value, int source) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8);
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, value), source));
  } else {
    struct Words {
      int low, high;
    };
    Words w = __builtin_bit_cast(Words, value);
    w.low = __builtin_amdgcn_readlane(w.low, source);
    w.high = __builtin_amdgcn_readlane(w.high, source);
    return __builtin_bit_cast(T, w);
  }
}

// delta_x f at column j from f there and at j - 1 (docs/math.md:337).
template <typename T>
__device__ __forceinline__ T backward_across_of(T here, T left, int j, int W) {
  // Every case made and one chosen by selects: no branch between the lanes.
  const T inner = here - left, last = -left;
  const T value = j == 0 ? here : (j == W - 1 ? last : inner);
  return W == 1 ? T(0) : value;
}

// delta_y f at row i from f there and at i - 1.
template <typename T>
__device__ __forceinline__ T backward_down_of(T here, T up, int i, int H) {
  const bool first = i == 0, last = i == H - 1;
  const T inner = here - up, bottom = -up;
  const T value = first ? here : (last ? bottom : inner);
  return (first && last) ? T(0) : value;
}

// The levels of a lane's row of a block in the coefficient phase, row k of block (ty, bx)
// (one 16-byte load), and the block's flat mark; zeros beyond the blocks. level_words()
// loads them, level_row() takes them apart where they are used, so that the loads can be
// issued before others without anything waiting for them early.
struct LevelRow {
  int16_t q[8];
  bool flat;
};
struct LevelWords {
  unsigned int pairs[4];
  uint8_t flat;
  bool inside;
};
__device__ __forceinline__ LevelWords level_words(const Component& K, int ty, int bx, int k, bool inside) {
  // Beyond the blocks the lane reads block 0 and keeps zeros, so that no branch parts the
  // loads from the others.
  const size_t block = static_cast<size_t>(inside ? ty : 0) * K.columns + (inside ? bx : 0);
  using Words = unsigned int __attribute__((ext_vector_type(4)));
  const Words words =
      *reinterpret_cast<const __attribute__((address_space(1))) Words*>(device::global_view(K.levels) + 64 * block + 8 * k);
  const uint8_t flat = device::global_view(K.flat)[block];
  // The words go through plain variables: a bit cast of a vector element is lowered by
  // clang as a load from the vector's first element.
  const unsigned int x = words[0], y = words[1], z = words[2], w = words[3];
  return LevelWords{{x, y, z, w}, flat, inside};
}
__device__ __forceinline__ LevelRow level_row(const LevelWords& words) {
  LevelRow row;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    row.q[2 * m] = words.inside ? static_cast<int16_t>(words.pairs[m] & 0xffffu) : int16_t{0};
    row.q[2 * m + 1] = words.inside ? static_cast<int16_t>(words.pairs[m] >> 16) : int16_t{0};
  }
  row.flat = words.inside && words.flat != 0;
  return row;
}
__device__ __forceinline__ LevelRow level_row(const Component& K, int ty, int bx, int k, bool inside) {
  return level_row(level_words(K, ty, bx, k, inside));
}

// Where a prox tile takes its component's constants: OwnStaging loads them and copies them
// into the work-group's LDS (load() before the tile's samples are loaded, store() after,
// with the work-group's barrier: every wavefront of the work-group takes part);
// PreStaged hands over constants already in LDS.
struct OwnStaging {
  const Component& K;
  device::StagedArrays& into;
  __device__ __forceinline__ device::StageLoaded load() const { return device::stage_load(K); }
  __device__ __forceinline__ device::Staged store(const device::StageLoaded& loaded) const {
    const device::Staged staged = device::stage_store(K, into, loaded);
    __syncthreads();
    return staged;
  }
};
struct PreStaged {
  struct Nothing {};
  device::Staged staged;
  __device__ __forceinline__ Nothing load() const { return {}; }
  __device__ __forceinline__ device::Staged store(Nothing) const { return staged; }
};

// ---------------------------------------------------------------------------------------
// The proximal step of one component whose cells are DV x DH samples (1 or 2 each)
// (primal_dual.hpp:536-546; frames.cpp:490-548). A wavefront takes 8 rows of blocks of a
// strip 64 component columns wide, a lane per column; the cell's samples of a lane stay
// in its registers from the descent to the write-back. Every load and store is issued by
// every lane, a lane beyond the tiles addressing outside its resource (zero read, store
// dropped), so that the loads of all rows are in flight together.
//
// prox_tile is work-group g's part of one component's step; prox_kernel runs a component
// (work-groups in a run on each XCD), and the batch's kernel runs tiles of many images.
// The LDS a prox tile transposes its blocks through, a wavefront's own part each.
template <typename T>
struct ProxRooms {
  T forward[4][8][65];
  T inverse[4][8][65];
};

// Where a prox tile takes the samples of p beside its own that another work-group writes
// in the same launch (the fused sweep's): p_x of the column left of the tile, row by row
// from `left` (left[(i - left_top) * left_stride]), and p_y of the row above the tile from
// `up` (up[j - up_first]), each null to read them from p in memory. Where `inside` is not
// null, the tile also writes there, a byte per coefficient in the layout of the kept
// coefficients, 1 where the prox's minimizer lay inside the interval (the clip's slope is
// the quadratic's there, 0 elsewhere: the unrolled solve's adjoint reads it).
template <typename T>
struct ProxEdges {
  const T* left = nullptr;
  int left_stride = 1;
  int left_top = 0;
  const T* up = nullptr;
  int up_first = 0;
  uint8_t* inside = nullptr;
};

// The tile of this wavefront in tile row ty and strip `strip` (8 tiles across) of
// component c. `staging` gives the component's constants in LDS (OwnStaging, whose store
// ends with the work-group's barrier, or PreStaged, staged before). The tile's first loads
// are the constants and its levels, which it takes first; then its samples. Its transposes
// go through the wavefront's own part of `rooms`, so past the staging no wavefront waits
// for another's loads. PLAIN: the component has no slack cost and midpoint centres
// (device::prox_of's PLAIN).
template <typename T, typename S, int DV, int DH, template <typename> class Planes, bool PLAIN = false,
          typename Staging>
__device__ __forceinline__ void prox_tile_at(const Component& K, Staging&& staging, ProxRooms<T>& rooms,
                                             Fields<T, S> F, int c, T* __restrict__ zeta_kept, T* x_out, int H,
                                             int W, T descent, T step, T rho, T rest, int keep, int ty, int strip,
                                             const ProxEdges<T>& edges) {
#pragma clang fp contract(off)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int b = lane >> 3, n = lane & 7;
  const int tx = strip * 8 + b;  // the tile of the lane's column
  const bool live = (ty < K.tiles_y) & (tx < K.tiles_x);
  const size_t N = static_cast<size_t>(H) * W;
  const Planes<T> x(F.x + c * N, 1, N), xbar(F.xbar + c * N, 1, N), tilde_x(F.tx + c * N, 1, N);
  const Planes<T> x_next(x_out + c * N, 1, N);  // the relaxed x: in place, or elsewhere
  const Planes<S> p(F.p + (2 * c) * N, 2, N);  // p_x, p_y
  const int first_column = (8 * tx + n) * DH;  // the lane's first canvas column
  const int top = 8 * ty * DV;                 // the wave's first canvas row
  const auto at = [&](int i, int j) { return static_cast<uint32_t>(i) * static_cast<uint32_t>(W) + j; };

  // The loads: p_y of the canvas row above, and per canvas row x, p_x and p_y at the lane's
  // columns and p_x of the column on its left (lane 0 only; the others take their left
  // lane's).
  // Cells two samples wide take each pair of samples of a row in one load of twice the width
  // (the pair is aligned: the first column and the canvas width are even).
  // p stays as stored until it is taken: widening a BF16 sample as it arrives made the
  // compiler wait for each load before issuing the next.
  // In the coefficient phase a lane takes row n of block b of the strip: its levels.
  const bool inside = (ty < K.rows) & (tx < K.columns);
  const auto staging_loaded = staging.load();
  const LevelWords own_words = level_words(K, ty, tx, n, inside);
  T up[DH], here[8][DV][DH];
  S across[8][DV][DH], down[8][DV][DH], edge[8][DV];
  // p_x of the column left of the strip, for lane 0: read from memory, lane k holds canvas
  // row top + k's (one register, not one a row); from edges.left, one a row.
  S edge_column{};
  if (edges.left == nullptr) {
    const int strip_column = 64 * strip * DH;  // lane 0's first column
    const bool edge_live = (ty < K.tiles_y) & (8 * strip < K.tiles_x) & (lane < 8 * DV) & (strip_column > 0);
    edge_column = p.load_if(edge_live, 0, at(top + min(lane, 8 * DV - 1), strip_column - 1));
  }
  const auto load_row = [&](const auto& planes, bool valid, int plane, uint32_t where, auto (&out)[DH]) {
    using Out = std::remove_cv_t<std::remove_reference_t<decltype(out[0])>>;
    if constexpr (DH == 2) {
      decltype(planes.load(0, 0)) first, second;
      planes.load_pair_if(valid, plane, where, first, second);
      if constexpr (std::is_same_v<Out, decltype(first)>) {
        out[0] = first;
        out[1] = second;
      } else {
        out[0] = device::widen(first);
        out[1] = device::widen(second);
      }
    } else {
#pragma unroll
      for (int bb = 0; bb < DH; ++bb) {
        const auto loaded = planes.load_if(valid, plane, where + bb);
        if constexpr (std::is_same_v<Out, std::remove_cv_t<decltype(loaded)>>) {
          out[bb] = loaded;
        } else {
          out[bb] = device::widen(loaded);
        }
      }
    }
  };
  if (edges.up != nullptr) {
#pragma unroll
    for (int bb = 0; bb < DH; ++bb) up[bb] = (live && top > 0) ? edges.up[first_column + bb - edges.up_first] : T(0);
  } else {
    load_row(p, live && top > 0, 1, at(top - 1, first_column), up);
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
#pragma unroll
    for (int a = 0; a < DV; ++a) {
      const int i = top + m * DV + a;
      load_row(x, live, 0, at(i, first_column), here[m][a]);
      load_row(p, live, 0, at(i, first_column), across[m][a]);
      load_row(p, live, 1, at(i, first_column), down[m][a]);
      if (edges.left != nullptr) {
        // Made from a stored sample, so its storage type holds it exactly.
        edge[m][a] = device::narrow<S>(edges.left[(i - edges.left_top) * edges.left_stride]);
      }
    }
  }
  const device::Staged staged = staging.store(staging_loaded);
  const LevelRow own = level_row(own_words);

  T vs[8][DV * DH], column[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    T partial[DH];
#pragma unroll
    for (int a = 0; a < DV; ++a) {
      const int i = top + m * DV + a;
      // p_x of the column on the left: the lane's own previous column, or the lane on the
      // left's last one, or beyond the strip, the edge sample.
      const T edge_here = edges.left != nullptr ? device::widen(edge[m][a])
                                                : lane_value(device::widen(edge_column), m * DV + a);
      const T left = from_left_or(device::widen(across[m][a][DH - 1]), edge_here);
#pragma unroll
      for (int bb = 0; bb < DH; ++bb) {
        const int j = first_column + bb;
        const T before = bb == 0 ? left : device::widen(across[m][a][bb - 1]);
        const T below = device::widen(down[m][a][bb]);
        T div = backward_down_of(below, up[bb], i, H);
        div += backward_across_of(device::widen(across[m][a][bb]), before, j, W);
        const T v = device::fused(descent, div, here[m][a][bb]);
        up[bb] = below;
        vs[m][a * DH + bb] = v;
        // The cell's sum: the rows of a column first, then the columns.
        partial[bb] = a == 0 ? v : partial[bb] + v;
      }
    }
    T sum = partial[0];
#pragma unroll
    for (int bb = 1; bb < DH; ++bb) sum += partial[bb];
    column[m] = DV * DH > 1 ? sum / static_cast<T>(DV * DH) : sum;
  }
  T y[8];
  device::forward8(column, y);
#pragma unroll
  for (int k = 0; k < 8; ++k) rooms.forward[wave][k][lane] = y[k];
  // The wavefront's own part of the room: no other wavefront waits for this one's loads.
  device::wave_sync();
  // A lane now takes row n of block b of the strip.
  T row[8], coefficient[8], zeta[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) row[u] = rooms.forward[wave][n][8 * b + u];
  device::forward8(row, coefficient);
  uint8_t* const marks =
      (edges.inside != nullptr && inside) ? edges.inside + 64 * (static_cast<size_t>(ty) * K.columns + tx) + 8 * n : nullptr;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    bool within = false;
    const T z = device::prox_of<PLAIN>(staged, own.q[u], own.flat, 8 * n + u, coefficient[u], step, within);
    zeta[u] = inside ? z : T(0);
    if (marks != nullptr) marks[u] = within;
    // One sample a cell: each coefficient's map before the next one's, not interleaved
    // (fewer registers held at once; the 1 x 1 tile 2 % faster, the 2 x 2 one no faster).
    if constexpr (DV * DH == 1) __builtin_amdgcn_sched_barrier(0);
  }
  if (keep && inside) {
    T* kept = zeta_kept + 64 * (static_cast<size_t>(ty) * K.columns + tx) + 8 * n;
#pragma unroll
    for (int u = 0; u < 8; ++u) kept[u] = zeta[u];
  }
  T back_row[8];
  device::inverse8(zeta, back_row);
#pragma unroll
  for (int v = 0; v < 8; ++v) rooms.inverse[wave][n][8 * b + v] = back_row[v];
  device::wave_sync();
  T back_column[8], back[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) back_column[m] = rooms.inverse[wave][m][lane];
  device::inverse8(back_column, back);
  // The lane's column of tiles is one of the blocks' (a tile beyond them keeps v): the
  // lane's block in the coefficient phase is the one of its column here (b = lane / 8).
  const bool own_inside = inside;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    // x~: the inverse DCT of zeta (one sample to a cell), or v moved by the change of its
    // cell (the inverse DCT less the mean); v beyond the blocks.
    const T change = back[m] - column[m];
#pragma unroll
    for (int a = 0; a < DV; ++a) {
      T bars[DH], relaxed[DH], tildes[DH];
#pragma unroll
      for (int bb = 0; bb < DH; ++bb) {
        const T v = vs[m][a * DH + bb], old = here[m][a][bb];
        const T tilde = own_inside ? (DV * DH > 1 ? v + change : back[m]) : v;
        bars[bb] = T(2) * tilde - old;
        relaxed[bb] = device::fused(rho, tilde, rest * old);
        tildes[bb] = tilde;
      }
      const uint32_t where = at(top + m * DV + a, first_column);
      if constexpr (DH == 2) {
        xbar.template store_pair_if<Planes<T>::kNext>(live, 0, where, bars[0], bars[1]);
        x_next.store_pair_if(live, 0, where, relaxed[0], relaxed[1]);
        if (keep) tilde_x.store_pair_if(live, 0, where, tildes[0], tildes[1]);
      } else {
#pragma unroll
        for (int bb = 0; bb < DH; ++bb) {
          xbar.template store_if<Planes<T>::kNext>(live, 0, where + bb, bars[bb]);
          x_next.store_if(live, 0, where + bb, relaxed[bb]);
          if (keep) tilde_x.store_if(live, 0, where + bb, tildes[bb]);
        }
      }
    }
  }
}

// Work-group g's part of one component's step: 4 tile rows of one strip, its constants
// staged here.
template <typename T, typename S, int DV, int DH, template <typename> class Planes, bool PLAIN = false>
__device__ __forceinline__ void prox_tile(const Component* __restrict__ component, Fields<T, S> F, int c,
                                          T* __restrict__ zeta_kept, T* x_out, int H, int W, T descent, T step, T rho,
                                          T rest, int keep, int g, uint8_t* inside = nullptr) {
  __shared__ ProxRooms<T> rooms;
  __shared__ device::StagedArrays staged_arrays;
  const Component& K = *component;
  const int wave = threadIdx.x >> 6;
  const int strips = (K.tiles_x + 7) >> 3;
  ProxEdges<T> edges;
  edges.inside = inside;
  prox_tile_at<T, S, DV, DH, Planes, PLAIN>(
      K, OwnStaging{K, staged_arrays}, rooms, F, c, zeta_kept, x_out, H, W, descent, step, rho,
      rest, keep, (g / strips) * 4 + wave, g % strips, edges);
}

// The wavefronts per SIMD a prox kernel is compiled for: as many as its registers allow
// without spilling (the scheduler otherwise trades them for latency it does not need).
template <typename T, int DV, int DH>
constexpr int kProxWaves =
    sizeof(T) == 4 ? (DV * DH == 1 ? 8 : 3) : (DV * DH == 1 ? 4 : (DV * DH == 2 ? 2 : 1));

// The steps of the components one prox launch takes, from its first: component c + k has
// descent[k] (tau gamma_c) and step[k] (tau over its cell's samples).
template <typename T>
struct ProxSteps {
  T descent[3];
  T step[3];
};

// Components c .. c + gridDim.y - 1 of one cell shape and size, blockIdx.y the component: the
// chroma of 4:2:0 in one launch, so that its work-groups fill the GPU together. PLAIN: every
// one of them has no slack cost and midpoint centres.
template <typename T, typename S, int DV, int DH, template <typename> class Planes, bool PLAIN = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kProxWaves<T, DV, DH>))) void prox_kernel(
    const Component* __restrict__ component, Fields<T, S> F, int c, int H, int W, ProxSteps<T> steps, T rho, T rest,
    int keep) {
  const int k = static_cast<int>(blockIdx.y);
  prox_tile<T, S, DV, DH, Planes, PLAIN>(component + k, F, c + k, F.zeta[c + k], F.x, H, W, steps.descent[k],
                                         steps.step[k], rho, rest, keep, device::ordered(blockIdx.x, gridDim.x));
}

// The same for any cell up to 4 x 4, its samples read again for the write-back.
template <typename T, typename S>
__global__ __launch_bounds__(256) void prox_any_cell_kernel(const Component* __restrict__ component, Fields<T, S> F,
                                                            int c, int H, int W, T descent, T step, T rho, T rest,
                                                            int keep) {
  __shared__ T forward_room[4][8][65];
  __shared__ T inverse_room[4][8][65];
  __shared__ device::StagedArrays staged_arrays;
  const Component& K = *component;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int strips = (K.tiles_x + 7) >> 3;
  const int g = device::ordered(blockIdx.x, gridDim.x);
  const int strip = g % strips;
  const int ty = (g / strips) * 4 + wave;
  const int tx = strip * 8 + (lane >> 3);
  const int n = lane & 7;
  const bool live = ty < K.tiles_y && tx < K.tiles_x;
  const int b = lane >> 3, k = lane & 7;
  const int bx = strip * 8 + b;
  const bool inside = ty < K.rows && bx < K.columns;
  const device::Staged staged = device::stage(K, staged_arrays);
  const LevelRow own = level_row(K, ty, bx, k, inside);
  const int dv = K.down, dh = K.across;
  const size_t N = static_cast<size_t>(H) * W;
  T* x = F.x + c * N;
  const S* px = F.p + (2 * c) * N;
  const S* py = F.p + (2 * c + 1) * N;
  T column[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    T value = T(0);
    if (live) {
      T sum = T(0);
      for (int bb = 0; bb < dh; ++bb) {
        T partial = T(0);
        for (int a = 0; a < dv; ++a) {
          const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + bb;
          const T v = x[static_cast<size_t>(i) * W + j] + descent * device::divergence_at<T>(px, py, i, j, H, W);
          partial = a == 0 ? v : partial + v;
        }
        sum = bb == 0 ? partial : sum + partial;
      }
      value = sum / static_cast<T>(dv * dh);
    }
    column[m] = value;
  }
  T y[8];
  device::forward8(column, y);
#pragma unroll
  for (int k = 0; k < 8; ++k) forward_room[wave][k][lane] = y[k];
  __syncthreads();
  T row[8], coefficient[8], zeta[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) row[u] = forward_room[wave][k][8 * b + u];
  device::forward8(row, coefficient);
  if (inside) {
    const size_t block = static_cast<size_t>(ty) * K.columns + bx;
#pragma unroll
    for (int u = 0; u < 8; ++u) zeta[u] = device::prox_of(staged, own.q[u], own.flat, 8 * k + u, coefficient[u], step);
    if (keep) {
      T* kept = F.zeta[c] + 64 * block + 8 * k;
#pragma unroll
      for (int u = 0; u < 8; ++u) kept[u] = zeta[u];
    }
  } else {
#pragma unroll
    for (int u = 0; u < 8; ++u) zeta[u] = T(0);
  }
  T back_row[8];
  device::inverse8(zeta, back_row);
#pragma unroll
  for (int v = 0; v < 8; ++v) inverse_room[wave][k][8 * b + v] = back_row[v];
  __syncthreads();
  T back_column[8], back[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) back_column[m] = inverse_room[wave][m][lane];
  device::inverse8(back_column, back);
  if (!live) return;
  const bool own_inside = ty < K.rows && tx < K.columns;
  T* xbar = F.xbar + c * N;
  T* tilde_x = F.tx + c * N;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const T change = back[m] - column[m];
    for (int bb = 0; bb < dh; ++bb) {
      for (int a = 0; a < dv; ++a) {
        const int i = (8 * ty + m) * dv + a, j = (8 * tx + n) * dh + bb;
        const size_t at = static_cast<size_t>(i) * W + j;
        const T here = x[at];
        const T v = here + descent * device::divergence_at<T>(px, py, i, j, H, W);
        const T tilde = own_inside ? (dv * dh > 1 ? v + change : back[m]) : v;
        xbar[at] = T(2) * tilde - here;
        x[at] = rho * tilde + rest * here;
        if (keep) tilde_x[at] = tilde;
      }
    }
  }
}

// ---------------------------------------------------------------------------------------
// The dual step of every channel, with TGV's step of w (primal_dual.hpp:547-557, 572-735),
// from the buffers of the iteration's start into the other ones.
//
// A wavefront walks down `rows` rows of a strip of 64 columns, a lane per column, so that
// each row of each field is one aligned 256-byte load. Per lane and channel it carries x-
// and r of the row and the next one and w- of the row above, and it loads the row after
// next while it works on the row. A neighbour in the strip comes from the neighbouring lane;
// the few samples beyond the strip's two edges (x- and r on the right for lane 63; w, p
// and r on the left, from which lane 0 makes w- of the column on its left) come from one
// gather per row, a lane per sample, read back as scalars.
constexpr int kStripColumns = 64;

// The arithmetic of the dual step at one sample, shared by every path that makes it (the
// row walk, and the fused sweep's making of a neighbour's p once more), with its fused
// multiply-adds written out (device::fused) and no other contraction, so that every path
// gives the same bits (docs/math.md:599-611; primal_dual.hpp:121-182, 276-327).
using device::fused;
// w~ = w + tau (p + div_2 r) and w- = 2 w~ - w at a sample from w, p and div_2 r (dx, dy).
template <typename T>
__device__ __forceinline__ void w_extrapolation(const T old[2], const T q[2], T dx, T dy, T descent, T tilde[2],
                                                T bar[2]) {
#pragma clang fp contract(off)
  tilde[0] = fused(descent, q[0] + dx, old[0]);
  tilde[1] = fused(descent, q[1] + dy, old[1]);
  bar[0] = T(2) * tilde[0] - old[0];
  bar[1] = T(2) * tilde[1] - old[1];
}

// p + sigma gamma (grad x- - w-) (w- zero for TV), one entry.
template <typename T>
__device__ __forceinline__ T p_ascent(T p, T ascent, T difference, T bar_w) {
#pragma clang fp contract(off)
  return fused(ascent, difference - bar_w, p);
}

// |q|^2 of two entries, and of the three of a symmetric tensor (the off-diagonal twice).
template <typename T>
__device__ __forceinline__ T square_of(T a, T b) {
#pragma clang fp contract(off)
  return fused(b, b, a * a);
}
template <typename T>
__device__ __forceinline__ T tensor_square_of(T xx, T yy, T xy) {
#pragma clang fp contract(off)
  return fused(T(2) * xy, xy, fused(yy, yy, xx * xx));
}

// rho q + (1 - rho) old.
template <typename T>
__device__ __forceinline__ T relaxed_of(T rho, T projected, T rest, T old) {
#pragma clang fp contract(off)
  return fused(rho, projected, rest * old);
}

// The gather of a work-group's edge samples for one row: the left ones for its first
// wavefront (`left`), the right ones for its last (`right`); the strips inside the
// work-group hand theirs over through LDS (EdgeRoom). Lanes kT c + e of the T gather hold
// w_x, w_y on the left (e = 0, 1), x- on the right (e = 2) and, for the fused sweep, x- on
// the left (e = 3) of channel c; lanes 9c + e of the S gather hold p_x, p_y, r_xx, r_yy,
// r_xy on the left (e = 0..4), r_yy and r_xy on the left one row down (e = 5, 6), and r_xx,
// r_xy on the right (e = 7, 8); TV gathers p_x and p_y on the left for the sweep only.
template <typename T, typename S, int C, bool TGV, bool SWEEP = false>
struct EdgeGather {
  static constexpr int kT = SWEEP ? 4 : 3;
  static constexpr bool kS = TGV || SWEEP;
  const T* t_source;
  const S* s_source;
  int t_down = 0, s_down = 0;
  size_t t_column = 0, s_column = 0;
  bool gathers;  // the same in every lane of the wavefront

  __device__ EdgeGather(const Fields<T, S>& F, int lane, int H, int W, int j0, bool left_side, bool right_side)
      : t_source(F.xbar), s_source(F.p), gathers(left_side || right_side) {
    const size_t N = static_cast<size_t>(H) * W;
    const size_t left = static_cast<size_t>(max(j0 - 1, 0));
    const size_t right = static_cast<size_t>(min(j0 + kStripColumns, W - 1));
    if (lane < kT * C) {
      const int c = lane / kT, e = lane % kT;
      if (e == 2 && right_side) {
        t_source = F.xbar + c * N;
        t_column = right;
      } else if (e == 3 && left_side) {
        t_source = F.xbar + c * N;
        t_column = left;
      } else if (TGV && e < 2 && left_side) {
        t_source = F.w + (2 * c + e) * N;
        t_column = left;
      }
    }
    if (kS && lane < 9 * C) {
      const int c = lane / 9, e = lane % 9;
      // The entry of p (e < 2) or r for each e, two bits each: 0, 1, 0, 1, 2, 1, 2, 0, 2.
      const int entry = (0x22644 >> (2 * e)) & 3;
      if ((TGV || e < 2) && (e < 7 ? left_side : right_side)) {
        s_source = (e < 2 ? F.p + (2 * c) * N : F.r + (3 * c) * N) + entry * N;
        s_down = (e == 5 || e == 6) ? 1 : 0;
        s_column = e < 7 ? left : right;
      }
    }
  }

  // This lane's samples for row i (rows clamped into the canvas), loaded with no branch
  // between the lanes: a lane without a sample of its own reads the first of x- (or p),
  // which no one takes. S is kept as stored and widened where it is taken, a row later,
  // so that nothing waits for these loads now. A wavefront that gathers nothing loads
  // nothing.
  __device__ __forceinline__ void load(int i, int H, int W, T& t, S& s) const {
    if (!gathers) return;
    t = t_source[static_cast<size_t>(min(i + t_down, H - 1)) * W + t_column];
    if constexpr (kS) s = s_source[static_cast<size_t>(min(i + s_down, H - 1)) * W + s_column];
  }
};

// The dual step's `keep`: kKeepTilde keeps w~ and r~ of the step (the tilde point's, which
// the records read), kKeepDualP p~ too (which only TV's records and the partial gap read).
constexpr int kKeepTilde = 1, kKeepDualP = 2;

// The samples beside each strip of a work-group, row by row: left(entry w) the column left
// of strip w (w_x, w_y, p_x, p_y, r_xx, r_yy, r_xy of the row, r_yy, r_xy of the row
// below), right(entry w + 1) the column right of it (x-, r_xx, r_xy). Lane 63 of strip w
// writes left(w + 1) and lane 0 right(w); the first strip's gather lanes write left(0) and
// the last's right(4) from memory. Every strip then reads its own two entries, the same
// address in every lane. Two rows' worth, taken in turn, so that one barrier a row parts a
// row's writes from the reads of the row before.
template <typename T, int C>
struct EdgeRoom {
  static constexpr int kLeft = 5 * C * 9, kWords = kLeft + 5 * C * 3;
  T words[2][kWords];
  __device__ __forceinline__ static constexpr int left_at(int entry, int c, int e) { return (entry * C + c) * 9 + e; }
  __device__ __forceinline__ static constexpr int right_at(int entry, int c, int e) {
    return kLeft + (entry * C + c) * 3 + e;
  }
  __device__ __forceinline__ T& left(int slot, int entry, int c, int e) { return words[slot][left_at(entry, c, e)]; }
  __device__ __forceinline__ T& right(int slot, int entry, int c, int e) { return words[slot][right_at(entry, c, e)]; }
};

// What the fused sweep takes from the dual step of a band for its prox step: p_x of the
// column left of each strip of the work-group, row by row (slot s: the column left of
// strip s; slot 0 made once more by the first wavefront, the others handed over by each
// strip's lane 63), and p_y of the row above a run, for the work-group's 256 columns.
template <typename T, int C>
struct SweepRoom {
  T left[5][16][C];
  T up[C][4 * kStripColumns];
};

//
// The dual step of every channel, with TGV's step of w (primal_dual.hpp:547-557, 572-735),
// from the buffers of the iteration's start into the other ones, walked row by row: a
// wavefront takes a strip of 64 columns, a lane per column, so that each row of each field
// is one aligned 256-byte load. Per lane and channel it carries x- and r of the row and the
// next one and w- of the row above, and it loads the row after next while it works on the
// row. A neighbour in the strip comes from the neighbouring lane; the samples beyond the
// strip's two edges (x-, r_xx, r_xy on the right for lane 63; w, p and r on the left, from
// which lane 0 makes w- of the column on its left) come through LDS from the wavefronts
// beside, or for the work-group's outer strips from one gather per row, a lane per sample,
// read back as scalars.
//
// DualWalk is the walk of the 4 strips of strip group `group` from row `first` to `end`:
// begin() sets it up, step(i) makes row i (rows in order from `first`). Every wavefront of
// the work-group makes the same barriers, those beyond the canvas too.
//
// With MAP, p's ball has a radius per sample, radius_map's (the unrolled solve's).
//
// With SWEEP (the fused sweep, whose prox step follows each band of rows in the same
// launch), the walk also hands the sweep what its prox step reads of p beside the
// work-group's own (SweepRoom): begin() makes p_y of row first - 1 once more (the run
// above writes it in the same launch), and step() gives p_x of the column left of each
// strip, made once more by the first wavefront for the work-group's left edge.
//
// KEEPS false leaves the tilde point's stores out of the walk altogether (the iterations
// that keep none: fewer scalar registers held through the rows).
template <typename T, typename S, int C, bool TGV, bool COUPLED, template <typename> class Planes, bool MAP = false,
          bool SWEEP = false, bool KEEPS = true>
struct DualWalk {
  static_assert(!(MAP && SWEEP), "the fused sweep takes one radius");
  using Gather = EdgeGather<T, S, C, TGV, SWEEP>;
  static constexpr int kT = Gather::kT;
  static constexpr int kNext = Planes<T>::kNext;

  EdgeRoom<T, C>& room;
  SweepRoom<T, C>* sweep;
  Fields<T, S> F;
  int H, W, first, end;
  Steps<T> steps;
  int keep;
  const T* radius_map;
  int wave, lane, j0, j;
  bool outside, left_gathered, right_gathered, writes, right;
  uint32_t column;
  Planes<T> xbar;
  Planes<S> p, p_next, tp;
  Planes<T> w, w_next, tw;
  Planes<S> r, r_next, tr;
  Gather edges;
  // The rows' loaded values, in three slots that the rows take in turn: row i's values lie
  // in slot (i - first) % 3 (x- and r from two rows before, p, w and the edge samples from
  // the row before), so that each value stays in one register from its load to its last
  // use. A value moved between registers while its load is in flight makes the move wait
  // for it, and with it for every load and store of the row. S samples stay as stored
  // until they are taken.
  T bar_s[3][C];
  S r_s[3][C][3], p_s[3][C][2];
  T w_s[3][C][2];
  T edge_t_s[3] = {T(0), T(0), T(0)};
  S edge_s_s[3] = {};
  // w- of the row above, carried from row to row.
  T w_up[C][2];
  // Where this lane's gathered samples go in the room (-1: nowhere), within one slot.
  int t_into = -1, s_into = -1;

  __device__ __forceinline__ DualWalk(EdgeRoom<T, C>& room_, SweepRoom<T, C>* sweep_, const Fields<T, S>& F_, int H_,
                                      int W_, int first_, int end_, int group, const Steps<T>& steps_, int keep_,
                                      const T* radius_map_)
      : room(room_),
        sweep(sweep_),
        F(F_),
        H(H_),
        W(W_),
        first(first_),
        end(end_),
        steps(steps_),
        keep(keep_),
        radius_map(radius_map_),
        wave(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6))),
        lane(threadIdx.x & 63),
        j0((group * 4 + wave) * kStripColumns),
        j(j0 + (threadIdx.x & 63)),
        outside(j0 >= W_),
        left_gathered(wave == 0),
        right_gathered(wave == 3),
        writes(j < W_),
        right(j + 1 < W_),
        column(static_cast<uint32_t>(min(j, W_ - 1))),
        xbar(F_.xbar, C, static_cast<size_t>(H_) * W_),
        p(F_.p, 2 * C, static_cast<size_t>(H_) * W_),
        p_next(F_.p_next, 2 * C, static_cast<size_t>(H_) * W_),
        tp(F_.tp, 2 * C, static_cast<size_t>(H_) * W_),
        w(F_.w, 2 * C, static_cast<size_t>(H_) * W_),
        w_next(F_.w_next, 2 * C, static_cast<size_t>(H_) * W_),
        tw(F_.tw, 2 * C, static_cast<size_t>(H_) * W_),
        r(F_.r, 3 * C, static_cast<size_t>(H_) * W_),
        r_next(F_.r_next, 3 * C, static_cast<size_t>(H_) * W_),
        tr(F_.tr, 3 * C, static_cast<size_t>(H_) * W_),
        edges(F_, threadIdx.x & 63, H_, W_, j0, wave == 0, wave == 3) {
    // The gather's lanes (EdgeGather's order) and the room's entries: the first strip's
    // left column into left(0), the last strip's right column into right(4). x- on the
    // left (the sweep's) stays in the gather.
    if (lane < kT * C) {
      const int c = lane / kT, e = lane % kT;
      if (TGV && e < 2 && left_gathered) t_into = EdgeRoom<T, C>::left_at(0, c, e);
      if (e == 2 && right_gathered) t_into = EdgeRoom<T, C>::right_at(4, c, 0);
    }
    if (Gather::kS && lane < 9 * C) {
      const int c = lane / 9, e = lane % 9;
      if ((TGV || e < 2) && e < 7 && left_gathered) s_into = EdgeRoom<T, C>::left_at(0, c, e + 2);
      if (TGV && e >= 7 && right_gathered) s_into = EdgeRoom<T, C>::right_at(4, c, e - 6);
    }
  }

  // The gathered samples of a row (t, s) into slot `slot` of the room.
  __device__ __forceinline__ void hand_gathered(int slot, T t, S s) {
    if (t_into >= 0) room.words[slot][t_into] = t;
    if constexpr (Gather::kS) {
      if (s_into >= 0) room.words[slot][s_into] = device::widen(s);
    }
  }

  // The index of the lane's sample in row i, i clamped into the canvas (a clamped load only
  // feeds differences that the canvas' edges take as zero).
  __device__ __forceinline__ uint32_t index(int i) const {
    return static_cast<uint32_t>(min(i, H - 1)) * static_cast<uint32_t>(W) + column;
  }

  // w~ and w- at a sample from w, p and r there, r of the row below and r_xx, r_xy of the
  // column on the right.
  __device__ __forceinline__ void w_step(const T old[2], const T q[2], const T here[3], const T next[3], T right_xx,
                                         T right_xy, bool across, bool below, T descent, T tilde[2], T bar_w[2]) const {
    const T dx = (across ? right_xx - here[0] : T(0)) + (below ? next[2] - here[2] : T(0));
    const T dy = (across ? right_xy - here[2] : T(0)) + (below ? next[1] - here[1] : T(0));
    w_extrapolation(old, q, dx, dy, descent, tilde, bar_w);
  }

  // The scale of p's ball for the steps of every channel (coupled), or of channel c.
  __device__ __forceinline__ T coupled_scale(const T step_p[C][2], T radius) const {
    T square = T(0);
#pragma unroll
    for (int c = 0; c < C; ++c) square += square_of(step_p[c][0], step_p[c][1]);
    return device::ball_scale(square, radius);
  }

  // The rows of the start: x- and r of the first two rows, p and w of the first, the edge
  // samples of the first row, and w- of the row above (and for the sweep p_y of the row
  // above, made as the run above makes it).
  __device__ __forceinline__ void begin() { begin_stretch(true); }

  // up_row: for the sweep, p_y of the row above is made once more (a run's first stretch;
  // the sweep walks each band as a stretch of its own).
  __device__ __forceinline__ void begin_stretch(bool up_row) {
    if (outside) {
      // The whole wavefront beyond the canvas: it keeps the work-group's barriers only.
      if ((TGV || SWEEP) && first > 0) __syncthreads();
      return;
    }
    T up_t = T(0);
    S up_s{};
    // x- and r of rows first and first + 1 into slots 0 and 1, p, w and the edge samples of
    // row first into slot 0.
    edges.load(first, H, W, edge_t_s[0], edge_s_s[0]);
    if ((TGV || SWEEP) && first > 0) edges.load(first - 1, H, W, up_t, up_s);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      bar_s[0][c] = xbar.load(c, index(first));
      bar_s[1][c] = xbar.load(c, index(first + 1));
#pragma unroll
      for (int e = 0; e < 2; ++e) p_s[0][c][e] = p.load(2 * c + e, index(first));
      if constexpr (TGV) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          r_s[0][c][e] = r.load(3 * c + e, index(first));
          r_s[1][c][e] = r.load(3 * c + e, index(first + 1));
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) w_s[0][c][e] = w.load(2 * c + e, index(first));
        w_up[c][0] = w_up[c][1] = T(0);
      }
    }
    if (!((TGV || SWEEP) && first > 0)) {
      settle();
      return;
    }
    // The row above: w- of it (TGV), whose right column's r_xx, r_xy come from the wavefront
    // on the right (its lane 0) through the room's slot 1 (the run's first row takes slot
    // 0); for the sweep also x- of it and of its right column, and p there.
    T r_up[C][3], p_up[C][2], old_up[C][2], bar_up[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
#pragma unroll
      for (int e = 0; e < 2; ++e) p_up[c][e] = device::widen(p.load(2 * c + e, index(first - 1)));
      if constexpr (TGV) {
#pragma unroll
        for (int e = 0; e < 3; ++e) r_up[c][e] = device::widen(r.load(3 * c + e, index(first - 1)));
#pragma unroll
        for (int e = 0; e < 2; ++e) old_up[c][e] = w.load(2 * c + e, index(first - 1));
      }
      if constexpr (SWEEP) bar_up[c] = xbar.load(c, index(first - 1));
      if (lane == 0) {
        if constexpr (SWEEP) room.right(1, wave, c, 0) = bar_up[c];
        if constexpr (TGV) {
          room.right(1, wave, c, 1) = r_up[c][0];
          room.right(1, wave, c, 2) = r_up[c][2];
        }
      }
    }
    hand_gathered(1, up_t, up_s);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if constexpr (TGV) {
        T tilde_up[2], r_first[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) r_first[e] = device::widen(r_s[0][c][e]);
        const T right_xx = from_right_or(r_up[c][0], room.right(1, wave + 1, c, 1));
        const T right_xy = from_right_or(r_up[c][2], room.right(1, wave + 1, c, 2));
        w_step(old_up[c], p_up[c], r_up[c], r_first, right_xx, right_xy, right, true, steps.descent[c], tilde_up,
               w_up[c]);
      }
    }
    if (SWEEP && up_row) {
      // p of row first - 1 made as the run above makes it (step() below), its p_y kept for
      // the prox step of this run's first band.
      T step_p[C][2];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const T bar_right = from_right_or(bar_up[c], room.right(1, wave + 1, c, 0));
        const T g1 = right ? bar_right - bar_up[c] : T(0);
        const T g2 = bar_s[0][c] - bar_up[c];  // row first lies in the canvas
        const T wx = TGV ? w_up[c][0] : T(0), wy = TGV ? w_up[c][1] : T(0);
        step_p[c][0] = p_ascent(p_up[c][0], steps.ascent[c], g1, wx);
        step_p[c][1] = p_ascent(p_up[c][1], steps.ascent[c], g2, wy);
      }
      const T shared_scale = COUPLED ? coupled_scale(step_p, steps.radius) : T(1);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        T scale = shared_scale;
        if constexpr (!COUPLED) scale = device::ball_scale(square_of(step_p[c][0], step_p[c][1]), steps.radius);
        const S stored = device::narrow<S>(relaxed_of(steps.rho, step_p[c][1] * scale, steps.rest, p_up[c][1]));
        sweep->up[c][wave * kStripColumns + lane] = device::widen(stored);
      }
    }
    settle();
  }

  // The start's loads done before the first row: the compiler's waits in the walk's loop
  // hold for every way into it, and with the start's loads still in flight at its entry
  // they would make each row wait for the row before's stores too.
  __device__ __forceinline__ static void settle() { __builtin_amdgcn_s_waitcnt(0); }

  // Rows first .. end - 1 in order, three a turn, so that each row's slot is known at
  // compile time; the last one or two after the loop, so that every way round the loop
  // leaves the slots alike.
  __device__ __forceinline__ void walk() {
    if (outside) {
      // The whole wavefront beyond the canvas keeps the work-group's barriers only, one a
      // row. Apart from the rows' loop, so that no way round that loop skips a row's work.
#pragma unroll 1
      for (int i = first; i < end; ++i) __syncthreads();
      return;
    }
    // The first row before the loop, so that the loop is entered as it goes round: with a
    // row's stores behind its loads.
    step<0>(first);
    int i = first + 1;
#pragma unroll 1
    for (; i + 2 < end; i += 3) {
      step<1>(i);
      step<2>(i + 1);
      step<0>(i + 2);
    }
    if (i < end) step<1>(i);
    if (i + 1 < end) step<2>(i + 1);
  }

  // Row i, whose slot is PH = (i - first) % 3; for the sweep, row i - first of its band.
  template <int PH>
  __device__ __forceinline__ void step(int i) {
    constexpr int kNextSlot = (PH + 1) % 3, kAfterSlot = (PH + 2) % 3;
    const int band_row = i - first;
    const bool below = i + 1 < H;
    // The loads of the next rows: x- and r two rows down into slot PH + 2, p, w and the
    // edge samples one row down into slot PH + 1; none in the run's last row, whose next
    // rows belong to the run below (an unwanted buffer load reads nothing from memory), but
    // for the sweep's edge samples, whose x- of the left column one row down the row itself
    // takes.
    const bool ahead = i + 1 < end;
    if (ahead || SWEEP) edges.load(i + 1, H, W, edge_t_s[kNextSlot], edge_s_s[kNextSlot]);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      bar_s[kAfterSlot][c] = xbar.load_if(ahead, c, index(i + 2));
#pragma unroll
      for (int e = 0; e < 2; ++e) p_s[kNextSlot][c][e] = p.load_if(ahead, 2 * c + e, index(i + 1));
      if constexpr (TGV) {
#pragma unroll
        for (int e = 0; e < 3; ++e) r_s[kAfterSlot][c][e] = r.load_if(ahead, 3 * c + e, index(i + 2));
#pragma unroll
        for (int e = 0; e < 2; ++e) w_s[kNextSlot][c][e] = w.load_if(ahead, 2 * c + e, index(i + 1));
      }
    }
    // Row i's values and row i + 1's x- and r, loaded by the rows before (the waits fall
    // here).
    T bar[C], bar_next[C], r_here[C][3], r_below[C][3], p_here[C][2], w_here[C][2];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      bar[c] = bar_s[PH][c];
      bar_next[c] = bar_s[kNextSlot][c];
#pragma unroll
      for (int e = 0; e < 2; ++e) p_here[c][e] = device::widen(p_s[PH][c][e]);
      if constexpr (TGV) {
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          r_here[c][e] = device::widen(r_s[PH][c][e]);
          r_below[c][e] = device::widen(r_s[kNextSlot][c][e]);
        }
#pragma unroll
        for (int e = 0; e < 2; ++e) w_here[c][e] = w_s[PH][c][e];
      }
    }
    const T edge_t = edge_t_s[PH];
    const S edge_s = edge_s_s[PH];

    // This row's edge samples for the wavefronts beside (and the gathered ones), then
    // theirs for this one.
    const int slot = (i - first) & 1;
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        room.right(slot, wave, c, 0) = bar[c];
        if constexpr (TGV) {
          room.right(slot, wave, c, 1) = r_here[c][0];
          room.right(slot, wave, c, 2) = r_here[c][2];
        }
      }
    }
    if (TGV && lane == 63) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const T handed[9] = {w_here[c][0], w_here[c][1], p_here[c][0], p_here[c][1], r_here[c][0],
                             r_here[c][1], r_here[c][2], r_below[c][1], r_below[c][2]};
#pragma unroll
        for (int e = 0; e < 9; ++e) room.left(slot, wave + 1, c, e) = handed[e];
      }
    }
    hand_gathered(slot, edge_t, edge_s);
    __syncthreads();
    // The left column's sample e of channel c (p_x, p_y, r_xx, r_yy, r_xy, r_yy and r_xy
    // below, then w_x and w_y at 7, 8) and the right column's (x-, r_xx, r_xy).
    const auto left_of = [&](int c, int e) { return room.left(slot, wave, c, e < 7 ? e + 2 : e - 7); };
    const auto right_of = [&](int c, int e) { return room.right(slot, wave + 1, c, e); };
    T step_p[C][2], step_r[C][3], tilde_w[C][2], bar_w[C][2], bar_left[C][2];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const T a = steps.ascent[c];
      const T bar_right = from_right_or(bar[c], right_of(c, 0));
      const T g1 = right ? bar_right - bar[c] : T(0);
      const T g2 = below ? bar_next[c] - bar[c] : T(0);
      if constexpr (TGV) {
        const T right_xx = from_right_or(r_here[c][0], right_of(c, 1));
        const T right_xy = from_right_or(r_here[c][2], right_of(c, 2));
        w_step(w_here[c], p_here[c], r_here[c], r_below[c], right_xx, right_xy, right, below, steps.descent[c],
               tilde_w[c], bar_w[c]);
        // w- of the column on the left: the lane on the left's, or for lane 0, made from the
        // edge samples and its own r (every lane makes it; lane 0 keeps it).
        const T old_left[2] = {left_of(c, 7), left_of(c, 8)};
        const T p_left[2] = {left_of(c, 0), left_of(c, 1)};
        const T r_left[3] = {left_of(c, 2), left_of(c, 3), left_of(c, 4)};
        const T r_left_below[3] = {T(0), left_of(c, 5), left_of(c, 6)};
        T tilde_left[2];
        w_step(old_left, p_left, r_left, r_left_below, r_here[c][0], r_here[c][2], true, below, steps.descent[c],
               tilde_left, bar_left[c]);
        const T left_x = from_left_or(bar_w[c][0], bar_left[c][0]);
        const T left_y = from_left_or(bar_w[c][1], bar_left[c][1]);
        // E w- (docs/math.md:355): backward differences.
        const T across_x = backward_across_of(bar_w[c][0], left_x, j, W);
        const T down_y = backward_down_of(bar_w[c][1], w_up[c][1], i, H);
        const T exy =
            T(0.5) * (backward_down_of(bar_w[c][0], w_up[c][0], i, H) + backward_across_of(bar_w[c][1], left_y, j, W));
        step_p[c][0] = p_ascent(p_here[c][0], a, g1, bar_w[c][0]);
        step_p[c][1] = p_ascent(p_here[c][1], a, g2, bar_w[c][1]);
        step_r[c][0] = p_ascent(r_here[c][0], a, across_x, T(0));
        step_r[c][1] = p_ascent(r_here[c][1], a, down_y, T(0));
        step_r[c][2] = p_ascent(r_here[c][2], a, exy, T(0));
      } else {
        step_p[c][0] = p_ascent(p_here[c][0], a, g1, T(0));
        step_p[c][1] = p_ascent(p_here[c][1], a, g2, T(0));
      }
    }
    const T rho = steps.rho, rest = steps.rest;
    T radius = steps.radius;
    if constexpr (MAP) radius = radius_map[index(i)];
    const T shared_scale = COUPLED ? coupled_scale(step_p, radius) : T(1);
    T shared_tensor_scale = T(1);
    if constexpr (COUPLED && TGV) {
      T tensor_square = T(0);
#pragma unroll
      for (int c = 0; c < C; ++c) tensor_square += tensor_square_of(step_r[c][0], step_r[c][1], step_r[c][2]);
      shared_tensor_scale = device::ball_scale(tensor_square, steps.radius0);
    }
    const uint32_t at = index(i);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      T scale = shared_scale;
      if constexpr (!COUPLED) scale = device::ball_scale(square_of(step_p[c][0], step_p[c][1]), radius);
      S stored_x{};
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const T projected = step_p[c][e] * scale;
        const S stored = device::narrow<S>(relaxed_of(rho, projected, rest, p_here[c][e]));
        if (e == 0) stored_x = stored;
        if (KEEPS && (keep & kKeepDualP)) {
          tp.template store_if<device::kStreaming>(writes, 2 * c + e, at, device::narrow<S>(projected));
        }
        p_next.template store_if<kNext>(writes, 2 * c + e, at, stored);
      }
      if constexpr (SWEEP) {
        // p_x of the strip's last column, the left column of the strip on the right.
        if (lane == 63) sweep->left[wave + 1][band_row][c] = device::widen(stored_x);
      }
      if constexpr (TGV) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          if (KEEPS && keep) tw.template store_if<device::kStreaming>(writes, 2 * c + e, at, tilde_w[c][e]);
          w_next.template store_if<kNext>(writes, 2 * c + e, at, relaxed_of(rho, tilde_w[c][e], rest, w_here[c][e]));
        }
        T tensor_scale = shared_tensor_scale;
        if constexpr (!COUPLED) {
          tensor_scale =
              device::ball_scale(tensor_square_of(step_r[c][0], step_r[c][1], step_r[c][2]), steps.radius0);
        }
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          const T projected = step_r[c][e] * tensor_scale;
          if (KEEPS && keep) {
            tr.template store_if<device::kStreaming>(writes, 3 * c + e, at, device::narrow<S>(projected));
          }
          r_next.template store_if<device::kStreaming>(
              writes, 3 * c + e, at, device::narrow<S>(relaxed_of(rho, projected, rest, r_here[c][e])));
        }
      }
    }
    if constexpr (SWEEP) {
      if (left_gathered) {
        // p_x of the column left of the work-group, made as the work-group on the left
        // makes it (above), from the edge samples: lane 0's own x- is its right neighbour.
        T left_step[C][2];
#pragma unroll
        for (int c = 0; c < C; ++c) {
          const T x_left = lane_value(edge_t, kT * c + 3), x_left_below = lane_value(edge_t_s[kNextSlot], kT * c + 3);
          const T x_right = lane_value(bar[c], 0);
          const T g1 = x_right - x_left;  // the column left of a strip has one on its right
          const T g2 = below ? x_left_below - x_left : T(0);
          const T wx = TGV ? bar_left[c][0] : T(0), wy = TGV ? bar_left[c][1] : T(0);
          left_step[c][0] = p_ascent(left_of(c, 0), steps.ascent[c], g1, wx);
          left_step[c][1] = p_ascent(left_of(c, 1), steps.ascent[c], g2, wy);
        }
        const T left_shared = COUPLED ? coupled_scale(left_step, radius) : T(1);
#pragma unroll
        for (int c = 0; c < C; ++c) {
          T scale = left_shared;
          if constexpr (!COUPLED) scale = device::ball_scale(square_of(left_step[c][0], left_step[c][1]), radius);
          const S stored = device::narrow<S>(relaxed_of(rho, left_step[c][0] * scale, rest, left_of(c, 0)));
          if (lane == 0) sweep->left[0][band_row][c] = device::widen(stored);
        }
      }
    }
    // This row's w- is the next one's row above.
    if constexpr (TGV) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
#pragma unroll
        for (int e = 0; e < 2; ++e) w_up[c][e] = bar_w[c][e];
      }
    }
  }
};

// The work of 4 strips across (strip group `group`) from row `first` to `end`.
template <typename T, typename S, int C, bool TGV, bool COUPLED, template <typename> class Planes, bool MAP = false,
          bool KEEPS = true>
__device__ __forceinline__ void dual_run(Fields<T, S> F, int H, int W, int first, int end, int group, Steps<T> steps,
                                         int keep, const T* __restrict__ radius_map = nullptr) {
  __shared__ EdgeRoom<T, C> room;
  DualWalk<T, S, C, TGV, COUPLED, Planes, MAP, false, KEEPS> walk(room, nullptr, F, H, W, first, end, group, steps,
                                                                  keep, radius_map);
  walk.begin();
  walk.walk();
}

// dual_tile is work-group g's part: 4 strips across and `rows` rows down; dual_kernel runs
// a frame (work-groups in a run on each XCD), and the batch's kernel runs tiles of many.
template <typename T, typename S, int C, bool TGV, bool COUPLED, template <typename> class Planes, bool MAP = false,
          bool KEEPS = true>
__device__ __forceinline__ void dual_tile(Fields<T, S> F, int H, int W, int rows, int groups_across,
                                          Steps<T> steps, int keep, int g, const T* __restrict__ radius_map = nullptr) {
  const int first = (g / groups_across) * rows;
  dual_run<T, S, C, TGV, COUPLED, Planes, MAP, KEEPS>(F, H, W, first, min(first + rows, H), g % groups_across, steps,
                                                      keep, radius_map);
}

// KEEPS false: for the iterations that keep no tilde point (keep 0).
template <typename T, typename S, int C, bool TGV, bool COUPLED, template <typename> class Planes, bool KEEPS = true>
__global__ __launch_bounds__(256) void dual_kernel(Fields<T, S> F, int H, int W, int rows, int groups_across,
                                                   Steps<T> steps, int keep) {
  dual_tile<T, S, C, TGV, COUPLED, Planes, false, KEEPS>(F, H, W, rows, groups_across, steps, keep,
                                                         device::ordered(blockIdx.x, gridDim.x));
}
