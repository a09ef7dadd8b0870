// pmg_norms.hip -- deterministic fold of per-block norm partials.
#include "pmg_common.h"

namespace pmg {

// One block per output; each thread keeps four independent accumulators so
// the loads of a whole slot are in flight together (the fold is latency-,
// not bandwidth-bound).  The summation order is fixed by the thread and
// unroll assignment, so repeated runs give identical norms.
__global__ __launch_bounds__(kFoldThreads) void k_norm_fold(NormWs ws, int slot0, int fold, double *out) {
  double ts[4] = {0.0, 0.0, 0.0, 0.0}, tc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int f = 0; f < fold; f++) {
    const int slot = slot0 + blockIdx.x * fold + f;
    const int nb = ws.slot_nb[slot];
    const double2 *p = reinterpret_cast<const double2 *>(ws.partials + (size_t)slot * ws.cap * 2);
    for (int b0 = threadIdx.x; b0 < nb; b0 += 4 * kFoldThreads) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int b = b0 + u * kFoldThreads;
        if (b < nb) {
          const double2 v = p[b];
          ts[u] += v.x;
          tc[u] += v.y;
        }
      }
    }
  }
  double s = (ts[0] + ts[1]) + (ts[2] + ts[3]);
  double c = (tc[0] + tc[1]) + (tc[2] + tc[3]);
  block_reduce2<kFoldThreads>(s, c);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(s / c);
}

}  // namespace pmg
