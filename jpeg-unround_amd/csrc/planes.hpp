// Which addressing the kernels take for the planes of a field: one buffer resource per
// kind of field (a 32-bit lane offset, the plane a scalar offset) or 64-bit addresses.
//
// A buffer resource reads zero and drops stores at a lane offset beyond its size, and the
// kernels use that for lanes with nothing to do: such a lane addresses kOutsideOffset
// (2^31). That offset lies beyond the resource only where the kind's planes span at most
// 2^31 bytes, so every kind of field a kernel reads or writes through buffer resources
// must span at most that; a larger one takes the 64-bit addresses. One predicate serves
// every kernel (a frame's prox and dual step, the batch's, the unrolled solve's), so that
// they cannot disagree. Host code only: no HIP here.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>

namespace unround {
namespace planes {

// The lane offset of "no sample" and the largest span the buffer addressing takes.
constexpr uint32_t kOutsideOffset = 0x80000000u;
constexpr size_t kBufferSpanLimit = size_t{1} << 31;

constexpr bool fits(size_t bytes) { return bytes <= kBufferSpanLimit; }

// A component's prox step: x, x-, x~ and the relaxed x (one plane of `t` bytes a sample
// each) and p (two planes of `s` bytes).
constexpr bool prox_fits(size_t samples, size_t t, size_t s) { return fits(samples * t) && fits(2 * samples * s); }

// The dual step of C channels: x- (C planes of T), p (2C of S), w (2C of T), r (3C of S).
constexpr bool dual_fits(size_t channels, size_t samples, size_t t, size_t s) {
  return fits(channels * samples * t) && fits(2 * channels * samples * s) && fits(2 * channels * samples * t) &&
         fits(3 * channels * samples * s);
}

// UNROUND_WIDE_PLANES=1 makes every kernel that has both forms take the 64-bit addresses,
// whatever the spans: the same values by another route, for testing that route.
inline bool wide_forced() {
  const char* value = std::getenv("UNROUND_WIDE_PLANES");
  return value != nullptr && std::strcmp(value, "1") == 0;
}

inline bool prox_buffers(size_t samples, size_t t, size_t s) { return !wide_forced() && prox_fits(samples, t, s); }
inline bool dual_buffers(size_t channels, size_t samples, size_t t, size_t s) {
  return !wide_forced() && dual_fits(channels, samples, t, s);
}

}  // namespace planes
}  // namespace unround
