// k_mixed.hip -- FP32 pair and k-space kernels of the mixed-precision path,
// and its back-projection on the matrix cores (FP16 hi/lo operands, FP32
// accumulation).  Positions, structure-factor sums beyond one LDS tile,
// per-ion force sums across tiles, energies and virials stay FP64.
#include "kernels.h"
#include "kernels_mfma.h"
#include "launch.h"

namespace mdg {

#define MDG_GRID(n, t) dim3((unsigned)(n)), dim3((unsigned)(t)), 0, s
#define MDG_GRID_LDS(n, t, b) dim3((unsigned)(n)), dim3((unsigned)(t)), (unsigned)(b), s

void launch_real32(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s) {
  if (n <= 0) return;
  if (analytic)
    MDG_LAUNCH(k_real_win<kRealAnalytic32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
  else
    MDG_LAUNCH(k_real_win<kRealTable32>, MDG_GRID(n, kRealThreads), boxes, items, mode);
}

void launch_mixed_tables(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_mixed_tables, MDG_GRID(n, kIonThreads), boxes, items);
}

void launch_kf_pack(const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n > 0) MDG_LAUNCH(k_kf_pack, MDG_GRID(n, kWave), boxes, items);
}

void launch_kforce_mfma(int kb, const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n <= 0) return;
  const int t = kKfMfmaWaves * kWave;
#define MDG_KF(K) \
  MDG_LAUNCH((k_kforce_mfma<K>), MDG_GRID(n, t), boxes, items)
  switch (kb) {
    case 1: MDG_KF(1); break;
    case 2: MDG_KF(2); break;
    case 3: MDG_KF(3); break;
    case 4: MDG_KF(4); break;
    case 6: MDG_KF(6); break;
    default: MDG_KF(8); break;
  }
#undef MDG_KF
}

int kforce_mfma_ions(int) { return 32 * kKfMfmaWaves; }

void launch_sk_mfma(int hl, const BoxDesc* boxes, const Item* items, int n, hipStream_t s) {
  if (n <= 0) return;
  const int lds = sk_mfma_lds_bytes(hl);
  static int attr_set = 0;  // largest dynamic LDS size granted so far
  if (lds > 65536 && lds > attr_set) {
    (void)hipFuncSetAttribute((const void*)k_sk_mfma, hipFuncAttributeMaxDynamicSharedMemorySize,
                              lds);
    attr_set = lds;
  }
  MDG_LAUNCH(k_sk_mfma, MDG_GRID_LDS(n, kSkmThreads, lds), boxes, items, hl);
}

}  // namespace mdg
