// k_mixed.hip -- FP32 pair and k-space kernels of the mixed-precision path.
// Positions, structure-factor sums beyond a thread's own ions, per-ion
// force sums across tiles/columns, energies and virials stay FP64.
#include "kernels.h"
#include "launch.h"

namespace mdg {

#define MDG_GRID(n, t) dim3((unsigned)(n)), dim3((unsigned)(t)), 0, s

void launch_real32(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s) {
  if (n <= 0) return;
  if (analytic)
    hipLaunchKernelGGL(k_real<kRealAnalytic32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
  else
    hipLaunchKernelGGL(k_real<kRealTable32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
}

void launch_sk_partial32(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sk_partial<float>, MDG_GRID(n, kSkThreads), boxes, items);
}

void launch_kforce32(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_kforce<float>, MDG_GRID(n, kKfThreads), boxes, items);
}

}  // namespace mdg
