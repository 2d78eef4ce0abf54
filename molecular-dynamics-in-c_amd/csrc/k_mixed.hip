// k_mixed.hip -- FP32 pair and k-space kernels of the mixed-precision path.
// Positions, structure-factor sums beyond one LDS tile, per-ion force sums
// across tiles and columns, energies and virials stay FP64.
#include "kernels.h"
#include "launch.h"

namespace mdg {

#define MDG_GRID(n, t) dim3((unsigned)(n)), dim3((unsigned)(t)), 0, s
#define MDG_GRID_LDS(n, t, b) dim3((unsigned)(n)), dim3((unsigned)(t)), (unsigned)(b), s

void launch_real32(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s) {
  if (n <= 0) return;
  if (analytic)
    hipLaunchKernelGGL(k_real<kRealAnalytic32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
  else
    hipLaunchKernelGGL(k_real<kRealTable32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
}

void launch_sk32(const BoxDesc* boxes, const Item* items, int n, int hs, int hsw, hipStream_t s) {
  if (n <= 0) return;
  const size_t lds = (size_t)kSkTile * (2 * hs + hsw) * sizeof(float2);
  hipLaunchKernelGGL(k_sk_seg<float>, MDG_GRID_LDS(n, kSkThreads, lds), boxes, items, hs, hsw);
}

void launch_kforce32(int hz, const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n <= 0) return;
  switch (hz) {
    case 8: hipLaunchKernelGGL((k_kforce<float, 8>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 16: hipLaunchKernelGGL((k_kforce<float, 16>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 24: hipLaunchKernelGGL((k_kforce<float, 24>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 32: hipLaunchKernelGGL((k_kforce<float, 32>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 48: hipLaunchKernelGGL((k_kforce<float, 48>), MDG_GRID(n, kKfThreads), boxes, items); break;
    default: hipLaunchKernelGGL((k_kforce<float, 64>), MDG_GRID(n, kKfThreads), boxes, items); break;
  }
}

}  // namespace mdg
