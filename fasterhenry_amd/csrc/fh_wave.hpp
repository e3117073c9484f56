// Appending to a global work queue with one atomic per wavefront: a ballot
// counts the lanes that append, the first of them reserves that many slots,
// and each takes the slot of its rank among them.  Every lane of the
// wavefront calls it together (converged control flow); the slot of a lane
// that does not append is meaningless.
#pragma once

#include <hip/hip_runtime.h>

namespace fh {

__device__ __forceinline__ unsigned wave_append(bool want, unsigned *count) {
  const unsigned long long votes = __ballot(want);
  if (votes == 0ull) return 0u;
  const int lane = static_cast<int>(__lane_id());
  const int leader = __ffsll(static_cast<unsigned long long>(votes)) - 1;
  unsigned base = 0u;
  if (lane == leader) base = atomicAdd(count, static_cast<unsigned>(__popcll(votes)));
  base = static_cast<unsigned>(__shfl(static_cast<int>(base), leader));
  return base + static_cast<unsigned>(__popcll(votes & ((1ull << lane) - 1ull)));
}

}  // namespace fh
