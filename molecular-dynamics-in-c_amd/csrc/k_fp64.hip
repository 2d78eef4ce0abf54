// k_fp64.hip -- FP64 kernels of the plain path and the FP64 kernels both
// precision modes share.  Built with -ffp-contract=off: products and sums
// written unfused round as mdc's SSE2 code does (explicit fma() calls in
// the k-space recurrences stay fused).
#include "kernels_common.h"
#include "launch.h"

namespace mdg {

#define MDG_GRID(n, t) dim3((unsigned)(n)), dim3((unsigned)(t)), 0, s

void launch_cell_build(const BoxDesc* boxes, const int* ids, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_cell_build, MDG_GRID(n, kCellThreads), boxes, ids);
}

void launch_phase_tables(bool fp32, const BoxDesc* boxes, const Item* items, int n,
                         hipStream_t s) {
  if (n <= 0) return;
  if (fp32) hipLaunchKernelGGL(k_phase_tables<float2>, MDG_GRID(n, kIonThreads), boxes, items);
  else hipLaunchKernelGGL(k_phase_tables<double2>, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_real64(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s) {
  if (n <= 0) return;
  if (analytic)
    hipLaunchKernelGGL(k_real<kRealAnalytic64>, MDG_GRID(n, kRealThreads), boxes, items, mode);
  else
    hipLaunchKernelGGL(k_real<kRealTable64>, MDG_GRID(n, kRealThreads), boxes, items, mode);
}

void launch_sk_partial64(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sk_partial<double>, MDG_GRID(n, kSkThreads), boxes, items);
}

void launch_kforce64(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_kforce<double>, MDG_GRID(n, kKfThreads), boxes, items);
}

void launch_sk_finalize(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_sk_finalize, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_force_vv(bool vv, const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n <= 0) return;
  if (vv) hipLaunchKernelGGL(k_force_vv<1>, MDG_GRID(n, kIonThreads), boxes, items);
  else hipLaunchKernelGGL(k_force_vv<0>, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_clear_foc(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_clear_foc, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_next_v(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_next_v, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_next_r(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_next_r, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_box_scalars(const BoxDesc* boxes, const int* ids, int n, int flags, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_box_scalars, MDG_GRID(n, kScalThreads), boxes, ids, flags);
}

void launch_moment_correction(const BoxDesc* boxes, const int* ids, int n, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_moment_correction, MDG_GRID(n, kScalThreads), boxes, ids);
}

}  // namespace mdg
