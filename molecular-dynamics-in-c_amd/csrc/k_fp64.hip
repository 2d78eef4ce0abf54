// k_fp64.hip -- FP64 kernels of the plain path and the FP64 kernels both
// precision modes share.  Built with -ffp-contract=off: products and sums
// written unfused round as mdc's SSE2 code does (explicit fma() calls in
// the k-space products stay fused).
#include "kernels_common.h"
#include "launch.h"

namespace mdg {

#define MDG_GRID(n, t) dim3((unsigned)(n)), dim3((unsigned)(t)), 0, s
#define MDG_GRID_LDS(n, t, b) dim3((unsigned)(n)), dim3((unsigned)(t)), (unsigned)(b), s

int kforce_hz_for(int hmax) {
  const int need = hmax + 1;
  const int opts[6] = {8, 16, 24, 32, 48, 64};
  for (int h : opts)
    if (need <= h) return h;
  return -1;
}

void launch_cell_build(const BoxDesc* boxes, const int* ids, int nids, const Item* items,
                       int nitems, const Item* cells, int ncells, hipStream_t s) {
  if (nids <= 0) return;
  if (nitems > 0) MDG_LAUNCH(k_cell_bin, MDG_GRID(nitems, kIonThreads), boxes, items);
  MDG_LAUNCH(k_cell_count, MDG_GRID(nids, kCellThreads), boxes, ids);
  if (nitems > 0) MDG_LAUNCH(k_cell_scatter, MDG_GRID(nitems, kIonThreads), boxes, items);
  if (ncells > 0) MDG_LAUNCH(k_cell_place, MDG_GRID(ncells, kWave), boxes, cells);
}

void launch_phase_tables(bool fp32, const BoxDesc* boxes, const Item* items, int n,
                         hipStream_t s) {
  if (n <= 0) return;
  if (fp32) MDG_LAUNCH(k_phase_tables<float2>, MDG_GRID(n, kIonThreads), boxes, items);
  else MDG_LAUNCH(k_phase_tables<double2>, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_real64(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s) {
  if (n <= 0) return;
  if (analytic)
    MDG_LAUNCH(k_real<kRealAnalytic64>, MDG_GRID(n, kRealThreads), boxes, items, mode);
  else
    MDG_LAUNCH(k_real<kRealTable64>, MDG_GRID(n, kRealThreads), boxes, items, mode);
}

void launch_sk64(const BoxDesc* boxes, const Item* items, int n, int hs, int hsw, hipStream_t s) {
  if (n <= 0) return;
  const size_t lds = (size_t)kSkTile * (2 * hs + hsw) * sizeof(double2);
  static bool raised = false;  // allow beyond 64 KB of dynamic LDS (gfx950 has 160 KB)
  if (lds > 65536 && !raised) {
    (void)hipFuncSetAttribute((const void*)k_sk_seg<double>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    raised = true;
  }
  MDG_LAUNCH(k_sk_seg<double>, MDG_GRID_LDS(n, kSkThreads, lds), boxes, items, hs, hsw);
}

void launch_kforce64(int hz, const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n <= 0) return;
  switch (hz) {
    case 8: MDG_LAUNCH((k_kforce<double, 8>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 16: MDG_LAUNCH((k_kforce<double, 16>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 24: MDG_LAUNCH((k_kforce<double, 24>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 32: MDG_LAUNCH((k_kforce<double, 32>), MDG_GRID(n, kKfThreads), boxes, items); break;
    case 48: MDG_LAUNCH((k_kforce<double, 48>), MDG_GRID(n, kKfThreads), boxes, items); break;
    default: MDG_LAUNCH((k_kforce<double, 64>), MDG_GRID(n, kKfThreads), boxes, items); break;
  }
}

void launch_sk_finalize(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_sk_finalize, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_force_vv(bool vv, const BoxDesc* boxes, const Item* items, int n, hipStream_t s,
                     int terms, bool scal) {
  if (n <= 0) return;
  if (scal) {
    if (vv) MDG_LAUNCH((k_force_vv<1, 1>), MDG_GRID(n, kIonThreads), boxes, items, terms);
    else MDG_LAUNCH((k_force_vv<0, 1>), MDG_GRID(n, kIonThreads), boxes, items, terms);
  } else {
    if (vv) MDG_LAUNCH(k_force_vv<1>, MDG_GRID(n, kIonThreads), boxes, items, terms);
    else MDG_LAUNCH(k_force_vv<0>, MDG_GRID(n, kIonThreads), boxes, items, terms);
  }
}

void launch_clear_foc(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_clear_foc, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_next_v(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_next_v, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_next_r(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_next_r, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_box_scalars(const BoxDesc* boxes, const int* ids, int nids, const Item* items,
                        int nitems, int flags, hipStream_t s, bool parts) {
  if (nids <= 0 || nitems <= 0) return;
  if (parts) MDG_LAUNCH(k_scal_part, MDG_GRID(nitems, kIonThreads), boxes, items);
  MDG_LAUNCH(k_box_scalars, MDG_GRID(nids, kScalThreads), boxes, ids, flags);
}

void launch_moment_correction(const BoxDesc* boxes, const int* ids, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_moment_correction, MDG_GRID(n, kScalThreads), boxes, ids);
}

void launch_scale_v(const BoxDesc* boxes, const Item* items, int n, double scale,
                    hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_scale_v, MDG_GRID(n, kIonThreads), boxes, items, scale);
}

void launch_snapshot(const BoxDesc* boxes, const Item* items, int n, const long long* off,
                     double* dst, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_snapshot, MDG_GRID(n, kIonThreads), boxes, items, off, dst);
}

void launch_observe(const BoxDesc* boxes, const int* ids, int n, long long step, int phases,
                    hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_observe, MDG_GRID(n, kScalThreads), boxes, ids, step, phases);
}

// Plain copy of n16 16-byte words: the reference kernel of known traffic
// for scaling hardware byte counters, and an HBM bandwidth yardstick.
__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ src,
                                                    float4* __restrict__ dst, size_t n16) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}

void launch_copy_probe(const void* src, void* dst, size_t n16, hipStream_t s) {
  if (n16 == 0) return;
  const size_t blocks = std::min<size_t>((n16 + 255) / 256, 8192);
  MDG_LAUNCH(k_copy_probe, MDG_GRID(blocks, 256), (const float4*)src, (float4*)dst, n16);
}

}  // namespace mdg
