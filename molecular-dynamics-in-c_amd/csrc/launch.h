// launch.h -- host-side launchers of libmdgpu's kernels.  Each launches one
// workgroup per work item on `s`; the caller owns the item arrays.
#pragma once

#include <hip/hip_ext.h>

#include "mdg_internal.h"

namespace mdg {

// Kernel timing.  While `start` is set, the launchers hand these events to
// hipExtLaunchKernel, which stamps them with the kernel's own start and end
// (its dispatch timestamps, no marker packets between kernels), and set
// `used`.  Set only around eager launches, never during graph capture.
struct LaunchEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool used = false;
};
extern thread_local LaunchEvents g_launch_events;

// k_fp64.hip (compiled with -ffp-contract=off)
// ids: the boxes with cell lists; items / cells: the per-ion and per-cell
// items of those built in four launches (not cell_fused; none: one launch
// does every box)
void launch_cell_build(const BoxDesc* boxes, const int* ids, int nids, const Item* items,
                       int nitems, const Item* cells, int ncells, hipStream_t s);
void launch_phase_tables(bool fp32, const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_real64(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s);
// hs, hsw: LDS row strides (complex elements) of the X/Y and W tiles
void launch_sk64(const BoxDesc* boxes, const Item* items, int n, int hs, int hsw, hipStream_t s);
// hz: register row length of the back-projection (8, 16, 24, 32, 48 or 64)
void launch_kforce64(int hz, const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_sk_finalize(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_copy_probe(const void* src, void* dst, size_t n16, hipStream_t s);
// scal: also k_scal_part's per-item partials (then launch_box_scalars
// with parts = false finishes the box scalars)
void launch_force_vv(bool vv, const BoxDesc* boxes, const Item* items, int n, hipStream_t s,
                     int terms = 3, bool scal = false);
void launch_clear_foc(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_next_v(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_next_r(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
// ids: the boxes; items: the per-ion items of exactly those boxes; parts:
// compute the per-item partials first (false when the force assembly did)
void launch_box_scalars(const BoxDesc* boxes, const int* ids, int nids, const Item* items,
                        int nitems, int flags, hipStream_t s, bool parts = true);
void launch_moment_correction(const BoxDesc* boxes, const int* ids, int n, hipStream_t s);
// r, v, scalars and observables of every box into one staging buffer
// (box b at dst + off[b]: [r 3N][v 3N][kScalCount][kObsLen] doubles)
void launch_snapshot(const BoxDesc* boxes, const Item* items, int n, const long long* off,
                     double* dst, hipStream_t s);
void launch_scale_v(const BoxDesc* boxes, const Item* items, int n, double scale,
                    hipStream_t s);
// mdc's calc_msd, control_temp and moment_correction after a step
// (main.c:44-60); step < 0: the box's own step count; phases: bit 0 MSD,
// bit 1 thermostat, bit 2 moment correction, bit 3 MSD origin only at step 1
void launch_observe(const BoxDesc* boxes, const int* ids, int n, long long step, int phases,
                    hipStream_t s);

// k_mixed.hip (FP32 pair and k-space arithmetic)
void launch_real32(bool analytic, const BoxDesc* boxes, const Item* items, int n, int mode,
                   hipStream_t s);
// MFMA back-projection: FP16 hi/lo W operands, packed weight tiles, then
// the tiles against 32-ion blocks; kb = K blocks (8 hz) of the widest box
void launch_mixed_tables(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_kf_pack(const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
void launch_kforce_mfma(int kb, const BoxDesc* boxes, const Item* items, int n, hipStream_t s);
// ions per back-projection workgroup for kb K blocks
int kforce_mfma_ions(int kb);
// MFMA structure factor: Z W records, then the partial sums; hl = hmax + 1
// of the widest box (sizes the LDS tiles)
void launch_sk_mfma(int hl, const BoxDesc* boxes, const Item* items, int n, hipStream_t s);

// register row lengths the back-projection is instantiated for
int kforce_hz_for(int hmax);

}  // namespace mdg

// MDG_LAUNCH(kernel, grid, block, lds, stream, args...): a kernel launch
// that carries g_launch_events when they are set.  Of several launches
// under one set of events the first stamps the start and the last the stop,
// so the events span the group.
#define MDG_LAUNCH(...) MDG_LAUNCH_(__VA_ARGS__)
#define MDG_LAUNCH_(kernel, grid, block, lds, strm, ...)                                 \
  do {                                                                                  \
    ::mdg::LaunchEvents& ev_ = ::mdg::g_launch_events;                                  \
    if (ev_.start) {                                                                    \
      hipExtLaunchKernelGGL(kernel, grid, block, lds, strm, ev_.used ? nullptr : ev_.start, \
                            ev_.stop, 0u, __VA_ARGS__);                                 \
      ev_.used = true;                                                                  \
    } else {                                                                            \
      hipLaunchKernelGGL(kernel, grid, block, lds, strm, __VA_ARGS__);                  \
    }                                                                                   \
  } while (0)
