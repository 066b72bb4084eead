// Host-side helpers shared by libmpm_hip and its host check library (plain
// C++, no HIP types).
#pragma once
#include <cstddef>

namespace mpm {

// Bytes per part when a host copy of `bytes` bytes is split over `nt`
// threads: ceil(bytes / nt) rounded up to a multiple of 64, so parts start
// on cache-line boundaries and nt of them cover every byte.  Part t spans
// [min(bytes, t * part), t + 1 == nt ? bytes : min(bytes, (t + 1) * part)).
inline size_t readback_part_bytes(size_t bytes, unsigned nt) {
  const size_t per = (bytes + nt - 1) / nt;
  return (per + 63) / 64 * 64;
}

}  // namespace mpm
