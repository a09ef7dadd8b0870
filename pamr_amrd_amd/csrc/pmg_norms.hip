// pmg_norms.hip -- deterministic fold of per-block norm partials.
#include "pmg_common.h"

namespace pmg {

__global__ __launch_bounds__(256) void k_norm_fold(NormWs ws, int slot0, int fold, double *out) {
  double ts = 0.0, tc = 0.0;
  for (int f = 0; f < fold; f++) {
    const int slot = slot0 + blockIdx.x * fold + f;
    const int nb = ws.slot_nb[slot];
    const double *p = ws.partials + (size_t)slot * ws.cap * 2;
    for (int b = threadIdx.x; b < nb; b += 256) {
      ts += p[2 * b];
      tc += p[2 * b + 1];
    }
  }
  block_reduce2<256>(ts, tc);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(ts / tc);
}

}  // namespace pmg
