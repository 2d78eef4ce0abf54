// kernels_common.h -- FP64 kernels shared by both precision modes: cell
// lists, phase tables, structure-factor finalisation, force assembly with
// the velocity-Verlet update, and per-box reductions.  Included only by
// k_fp64.hip, which is compiled with -ffp-contract=off: the integrator then
// rounds exactly as mdc's next_v / next_r do on the host.
#pragma once

#include "kernels.h"

namespace mdg {

// Fixed-order workgroup sum of NV doubles (wave butterflies, then waves in
// index order).  Result valid in thread 0.
template <int NT, int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NT / 64][NV];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) sh[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int q = 0; q < NT / 64; ++q) s += sh[q][k];
      v[k] = s;
    }
  }
  __syncthreads();
}

// ----------------------------------------------------------------------
// Cell lists.  Binning is integer and exact: c_a = (int)(r_a * (n_a / L_a))
// clamped to n_a - 1 (r_a == L_a is a legal coordinate after mdc's fold,
// rv.c:59-64).  Within a cell the ions are listed in ascending index, so
// the order never depends on timing.
//
// A box of up to kCellFusedN ions (cell_fused) is built entirely by one
// workgroup of k_cell_count: bin, count, scan, an unordered scatter into
// LDS slots by LDS cursors, and the placement below.  A larger box is built
// in four launches so that it spreads over the chip:
//   k_cell_bin     per ion: its cell;
//   k_cell_count   per box: counts, cell_start (exclusive scan) and the
//                  scatter's fill cursors;
//   k_cell_scatter per ion: a slot of its cell by an atomic cursor (one
//                  atomic per cell and wave); the slots within a cell come
//                  in whatever order the atomics took;
//   k_cell_place   per cell: each of its ions goes to cell_start + the
//                  number of ions of its cell with a lower index, and the
//                  cell-ordered copies of positions and species are written.
// The per-ion and per-cell launches cover the items of the boxes that need
// them and skip any other box.
// ----------------------------------------------------------------------
__device__ __forceinline__ int cell_of_pos(const BoxDesc& B, double x, double y, double z,
                                           int& err) {
  if (!(x >= 0.0 && x <= B.L[0] && y >= 0.0 && y <= B.L[1] && z >= 0.0 && z <= B.L[2]))
    err |= kErrState;
  int cx = (x >= 0.0) ? (int)(x * B.cellscale[0]) : 0;
  int cy = (y >= 0.0) ? (int)(y * B.cellscale[1]) : 0;
  int cz = (z >= 0.0) ? (int)(z * B.cellscale[2]) : 0;
  cx = min(max(cx, 0), B.nc[0] - 1);
  cy = min(max(cy, 0), B.nc[1] - 1);
  cz = min(max(cz, 0), B.nc[2] - 1);
  return (cz * B.nc[1] + cy) * B.nc[0] + cx;
}

__global__ __launch_bounds__(kIonThreads) void k_cell_bin(const BoxDesc* __restrict__ boxes,
                                                          const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  if (B.all_pairs || cell_fused(B.N, B.ncell) || i >= B.N) return;
  int err = 0;
  B.cell_of[i] = cell_of_pos(B, B.rx[i], B.ry[i], B.rz[i], err);
  if (err) atomicOr(B.err, err);
}

__global__ __launch_bounds__(kCellThreads) void k_cell_count(const BoxDesc* __restrict__ boxes,
                                                             const int* __restrict__ box_ids) {
  __shared__ int s_cnt[kCellLds];
  __shared__ int s_scan[kCellThreads];
  __shared__ int s_tmp[kCellFusedN];
  const BoxDesc& B = boxes[box_ids[blockIdx.x]];
  const int N = B.N, ncell = B.ncell, tid = threadIdx.x;
  const bool fused = cell_fused(N, ncell);
  // counters in LDS for up to kCellLds cells, else in cell_count, where
  // only atomics (which act in L2) touch them until the scan has read them
  const bool lds = ncell <= kCellLds;
  int* cnt = lds ? s_cnt : B.cell_count;
  if (lds) {
    for (int c = tid; c < ncell; c += kCellThreads) s_cnt[c] = 0;
  } else {
    for (int c = tid; c < ncell; c += kCellThreads) atomicExch(&cnt[c], 0);
    __threadfence();  // every wave's global atomics done before the barrier
  }
  __syncthreads();
  // the one-workgroup box keeps its ions' cells in registers (ion
  // tid + k kCellThreads in cl[k])
  constexpr int kPer = kCellFusedN / kCellThreads;
  int cl[kPer];
  if (fused) {
    int err = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * kCellThreads;
      cl[k] = 0;
      if (i < N) {
        cl[k] = cell_of_pos(B, B.rx[i], B.ry[i], B.rz[i], err);
        atomicAdd(&s_cnt[cl[k]], 1);
      }
    }
    if (err) atomicOr(B.err, err);
  } else {
    for (int i = tid; i < N; i += kCellThreads) atomicAdd(&cnt[B.cell_of[i]], 1);
    if (!lds) __threadfence();
  }
  __syncthreads();
  // exclusive scan: thread t owns cells [t * cper, (t + 1) * cper)
  const int cper = (ncell + kCellThreads - 1) / kCellThreads;
  const int c0 = min(tid * cper, ncell), c1 = min(c0 + cper, ncell);
  int local = 0;
  for (int c = c0; c < c1; ++c) local += lds ? cnt[c] : atomicAdd(&cnt[c], 0);
  s_scan[tid] = local;
  __syncthreads();
  for (int off = 1; off < kCellThreads; off <<= 1) {
    const int add = tid >= off ? s_scan[tid - off] : 0;
    __syncthreads();
    s_scan[tid] += add;
    __syncthreads();
  }
  int run = s_scan[tid] - local;
  for (int c = c0; c < c1; ++c) {
    const int n = lds ? cnt[c] : atomicAdd(&cnt[c], 0);  // read before the cursor replaces it
    B.cell_start[c] = run;
    if (fused) s_cnt[c] = run;          // LDS cursor
    else B.cell_count[c] = run;         // the scatter's cursor
    run += n;
  }
  if (tid == kCellThreads - 1) B.cell_start[ncell] = s_scan[tid];
  if (!fused) return;
  __syncthreads();
  // The one-workgroup box: an unordered scatter into LDS slots holding
  // (cell << 13 | ion) (ions < 2^13, cells < 2^12), after which each LDS
  // cursor is its cell's end and the previous cell's cursor its start; the
  // slots of one cell share the high bits, so comparing slots compares ions.
  static_assert(kCellFusedN <= (1 << 13) && kCellLds <= (1 << 12), "slot packing");
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = tid + k * kCellThreads;
    if (i < N) s_tmp[atomicAdd(&s_cnt[cl[k]], 1)] = (cl[k] << 13) | i;
  }
  __syncthreads();
  for (int p = tid; p < N; p += kCellThreads) {
    const int v = s_tmp[p], c = v >> 13, j = v & ((1 << 13) - 1);
    const int st = c == 0 ? 0 : s_cnt[c - 1], e = s_cnt[c];
    int rank = 0;
    for (int q = st; q < e; ++q) rank += s_tmp[q] < v ? 1 : 0;
    const int d = st + rank;
    B.sorted[d] = j;
    B.sx[d] = B.rx[j]; B.sy[d] = B.ry[j]; B.sz[d] = B.rz[j];
    B.styp[d] = B.typ[j];
  }
}

__global__ __launch_bounds__(kIonThreads) void k_cell_scatter(const BoxDesc* __restrict__ boxes,
                                                              const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  if (B.all_pairs || cell_fused(B.N, B.ncell)) return;  // uniform over the block
  const int i = it.a + threadIdx.x, lane = threadIdx.x % kWave;
  const bool act = i < B.N;
  const int c = act ? B.cell_of[i] : -1;
  // one atomic per distinct cell of the wave (ions of consecutive index
  // often share a cell): the lowest lane of each cell group claims the
  // group's slots, the others take theirs by rank within the group
  const unsigned long long below = (1ull << lane) - 1ull;
  unsigned long long todo = __ballot(act);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const int cl = __shfl(c, leader, kWave);
    const unsigned long long m = __ballot(c == cl);
    int base = 0;
    if (lane == leader) base = atomicAdd(&B.cell_count[cl], __popcll(m));
    base = __shfl(base, leader, kWave);
    if (c == cl) B.cell_tmp[base + __popcll(m & below)] = i;
    todo &= ~m;
  }
}

// one wave per cell: the cell's slots staged in LDS (a cell of more than
// kCellPlaceLds ions reads them from memory), each slot's ion placed at the
// cell's start plus its rank by index among the cell's ions
constexpr int kCellPlaceLds = 4096;
__global__ __launch_bounds__(kWave) void k_cell_place(const BoxDesc* __restrict__ boxes,
                                                      const Item* __restrict__ items) {
  __shared__ int s_slot[kCellPlaceLds];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  if (B.all_pairs || cell_fused(B.N, B.ncell)) return;
  const int lane = threadIdx.x;
  const int st = B.cell_start[it.a], n = B.cell_start[it.a + 1] - st;
  const bool lds = n <= kCellPlaceLds;
  if (lds)
    for (int k = lane; k < n; k += kWave) s_slot[k] = B.cell_tmp[st + k];
  __syncthreads();
  for (int k = lane; k < n; k += kWave) {
    const int j = lds ? s_slot[k] : B.cell_tmp[st + k];
    int rank = 0;
    if (lds) {
      for (int q = 0; q < n; ++q) rank += s_slot[q] < j ? 1 : 0;
    } else {
      for (int q = 0; q < n; ++q) rank += B.cell_tmp[st + q] < j ? 1 : 0;
    }
    const int d = st + rank;
    B.sorted[d] = j;
    B.sx[d] = B.rx[j]; B.sy[d] = B.ry[j]; B.sz[d] = B.rz[j];
    B.styp[d] = B.typ[j];
  }
}

// ----------------------------------------------------------------------
// Per-ion phase tables X[h] = e^{2 pi i h x/Lx}, Y[h], W[h] for
// h = 0..hmax, from FP64 sincospi (argument reduction by exact multiples
// of 2); stored FP64 or rounded to FP32.
// ----------------------------------------------------------------------
template <typename CT>
__global__ __launch_bounds__(kIonThreads) void k_phase_tables(const BoxDesc* __restrict__ boxes,
                                                              const Item* __restrict__ items) {
  using RT = decltype(CT{}.x);
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N;
  const int i = it.a + threadIdx.x;
  if (i >= N) return;
  const double s[3] = {B.rx[i] * B.invL[0], B.ry[i] * B.invL[1], B.rz[i] * B.invL[2]};
  CT* tab[3] = {reinterpret_cast<CT*>(B.tabx), reinterpret_cast<CT*>(B.taby),
                reinterpret_cast<CT*>(B.tabw)};
  for (int d = 0; d < 3; ++d)
    for (int h = 0; h <= B.hmax; ++h) {
      double sn, cs;
      sincospi(2.0 * (double)h * s[d], &sn, &cs);
      tab[d][(size_t)h * N + i] = CT{(RT)cs, (RT)sn};
    }
}

// ----------------------------------------------------------------------
// Structure factor per k from the chunk partials (fixed chunk order), with
// mdc's doubling of both sums (rcprcl.c:77-78), the energy term
// vp exp(..)/m^2 (2Sc^2 + 2Ss^2) and the back-projection weights
// T = cf (2Ss, 2Sc) and hz T.
// ----------------------------------------------------------------------
__global__ __launch_bounds__(kIonThreads) void k_sk_finalize(const BoxDesc* __restrict__ boxes,
                                                             const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int k = it.a + threadIdx.x;
  double e[1] = {0.0};
  if (k < B.nk) {
    double sc = 0.0, ss = 0.0;
    for (int ch = 0; ch < B.nchunk; ++ch) {
      const double2 p = B.skpart[(size_t)ch * B.nk + k];
      sc += p.x;
      ss += p.y;
    }
    sc *= 2.0;
    ss *= 2.0;
    e[0] = B.kce[k] * (sc * sc + ss * ss);
    const double a = B.kcf[k] * ss, b = B.kcf[k] * sc, hz = (double)B.khz[k];
    // FP64 back-projection: quad slot layout; MFMA packing (mixed): k order
    const size_t slot = B.kfa ? (size_t)k : (size_t)B.kslot[k];
    double* t = B.kab + 4 * slot;
    t[0] = a; t[1] = b; t[2] = hz * a; t[3] = hz * b;
    if (B.kab32) {
      float* u = B.kab32 + 4 * slot;
      u[0] = (float)a; u[1] = (float)b; u[2] = (float)(hz * a); u[3] = (float)(hz * b);
    }
  }
  // energy of this block of k (fixed order); box_scalars sums the blocks
  block_sum<kIonThreads, 1>(e);
  if (threadIdx.x == 0) B.kpot[it.a / kIonThreads] = e[0];
}

// k_scal_part's work for one ion item (a whole workgroup): vx, vy, vz are
// the velocities of this thread's ion (ignored past the box's end).
__device__ __forceinline__ void scal_part_block(const BoxDesc& B, const Item& it, double vx,
                                                double vy, double vz) {
  const int tid = threadIdx.x, q = it.a / kIonThreads, nq = B.nspart;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const int s0 = (int)((long)q * B.nreal / nq), s1 = (int)((long)(q + 1) * B.nreal / nq);
  for (int s = s0 + tid; s < s1; s += kIonThreads) {
    v[0] += B.rpart[4 * s];
    v[1] += B.rpart[4 * s + 1];
    v[2] += B.rpart[4 * s + 2];
    v[3] += B.rpart[4 * s + 3];
    v[8] += (double)B.rpairs[s];
  }
  if (B.kspace) {
    const int nkb = (B.nk + kIonThreads - 1) / kIonThreads;
    const int k0 = (int)((long)q * nkb / nq), k1 = (int)((long)(q + 1) * nkb / nq);
    for (int k = k0 + tid; k < k1; k += kIonThreads) v[4] += B.kpot[k];
  }
  const int i = it.a + tid;
  if (i < B.N) {
    const double m = B.m[B.typ[i]];
    v[5] = m * vx * vx;
    v[6] = m * vy * vy;
    v[7] = m * vz * vz;
  }
  block_sum<kIonThreads, 9>(v);
  if (tid == 0)
    for (int k = 0; k < 9; ++k) B.spart[9 * (size_t)q + k] = v[k];
}

// ----------------------------------------------------------------------
// Force assembly: with split real space, f0 <- f and f <- sum of the
// partials (split order); then the reciprocal partials are added (group
// order).  With VV also mdc's next_v then next_r (rv.c:33-66) in mdc's
// operation order, with the one-fold strict-comparison wrap.
// ----------------------------------------------------------------------
// terms: bit 0 = the real-space kernel ran (its split partials or f hold
// the real-space force), bit 1 = add the reciprocal-space partials.
// SCAL: the block then does k_scal_part's work for its ions (the kinetic
// sums of the velocities just written, its share of the real-space and
// reciprocal partials), one launch less per step; k_box_scalars follows.
template <int VV, int SCAL = 0>
__global__ __launch_bounds__(kIonThreads) void k_force_vv(const BoxDesc* __restrict__ boxes,
                                                          const Item* __restrict__ items,
                                                          int terms) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N;
  const int i = it.a + threadIdx.x;
  if (!SCAL && i >= N) return;
  double vx = 0.0, vy = 0.0, vz = 0.0;
  if (i < N) {
    double fx, fy, fz;
    if (!(terms & 1) || B.fix) {
      // no real space, or the mixed real space (all of it in the fixed-point
      // array, added below)
      fx = 0.0; fy = 0.0; fz = 0.0;
      B.fx0[i] = B.fx[i]; B.fy0[i] = B.fy[i]; B.fz0[i] = B.fz[i];
    } else if (B.rsplit > 1) {
      fx = 0.0; fy = 0.0; fz = 0.0;
      for (int s = 0; s < B.rsplit; ++s) {
        const double* p = B.rpf + (size_t)s * 3 * N;
        fx += p[i];
        fy += p[N + i];
        fz += p[2 * N + i];
      }
      B.fx0[i] = B.fx[i]; B.fy0[i] = B.fy[i]; B.fz0[i] = B.fz[i];
    } else {
      fx = B.fx[i]; fy = B.fy[i]; fz = B.fz[i];
    }
    if (B.fix && (terms & 1)) {
      // the mixed real space (home sums and partner shares), exact in fixed
      // point, at the ion's position in the real space's visiting order;
      // cleared for the next evaluation
      const int p = B.rstride == 1 ? i : mul_mod(i, B.rstride_inv, N, 1.0 / (double)N);
      const double s = 1.0 / kFixScale;
      fx += (double)(long long)B.fix[p] * s;
      fy += (double)(long long)B.fix[N + p] * s;
      fz += (double)(long long)B.fix[2 * N + p] * s;
      B.fix[p] = 0ull; B.fix[N + p] = 0ull; B.fix[2 * N + p] = 0ull;
    }
    if (B.kspace && (terms & 2)) {
      for (int g = 0; g < B.ngroup; ++g) {
        const double* fp = B.fpart + (size_t)g * 3 * N;
        fx += fp[i];
        fy += fp[N + i];
        fz += fp[2 * N + i];
      }
    }
    B.fx[i] = fx; B.fy[i] = fy; B.fz[i] = fz;
    if (VV) {
      const double m2 = 2.0 * B.m[B.typ[i]];
      vx = B.vx[i] + B.dt * (fx + B.fx0[i]) / m2;
      vy = B.vy[i] + B.dt * (fy + B.fy0[i]) / m2;
      vz = B.vz[i] + B.dt * (fz + B.fz0[i]) / m2;
      double rx = B.rx[i] + (B.dt * vx + B.dt2 * fx / m2);
      double ry = B.ry[i] + (B.dt * vy + B.dt2 * fy / m2);
      double rz = B.rz[i] + (B.dt * vz + B.dt2 * fz / m2);
      if (rx > B.L[0]) rx = rx - B.L[0];
      if (ry > B.L[1]) ry = ry - B.L[1];
      if (rz > B.L[2]) rz = rz - B.L[2];
      if (rx < 0.0) rx = B.L[0] + rx;
      if (ry < 0.0) ry = B.L[1] + ry;
      if (rz < 0.0) rz = B.L[2] + rz;
      if (!(rx >= 0.0 && rx <= B.L[0] && ry >= 0.0 && ry <= B.L[1] && rz >= 0.0 &&
            rz <= B.L[2]))
        atomicOr(B.err, kErrState);
      B.vx[i] = vx; B.vy[i] = vy; B.vz[i] = vz;
      B.rx[i] = rx; B.ry[i] = ry; B.rz[i] = rz;
    } else if (SCAL) {
      vx = B.vx[i]; vy = B.vy[i]; vz = B.vz[i];
    }
  }
  if (SCAL) scal_part_block(B, it, vx, vy, vz);
}

// mdc's clear_foc (init.c:179-190), next_v (rv.c:33-44), next_r (rv.c:46-66)
// as separate kernels for the one-routine-per-call drop-in.
__global__ __launch_bounds__(kIonThreads) void k_clear_foc(const BoxDesc* __restrict__ boxes,
                                                           const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  if (i >= B.N) return;
  B.fx0[i] = B.fx[i]; B.fy0[i] = B.fy[i]; B.fz0[i] = B.fz[i];
  B.fx[i] = 0.0; B.fy[i] = 0.0; B.fz[i] = 0.0;
}

__global__ __launch_bounds__(kIonThreads) void k_next_v(const BoxDesc* __restrict__ boxes,
                                                        const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  if (i >= B.N) return;
  const double m2 = 2.0 * B.m[B.typ[i]];
  B.vx[i] += B.dt * (B.fx[i] + B.fx0[i]) / m2;
  B.vy[i] += B.dt * (B.fy[i] + B.fy0[i]) / m2;
  B.vz[i] += B.dt * (B.fz[i] + B.fz0[i]) / m2;
}

__global__ __launch_bounds__(kIonThreads) void k_next_r(const BoxDesc* __restrict__ boxes,
                                                        const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  if (i >= B.N) return;
  const double m2 = 2.0 * B.m[B.typ[i]];
  double rx = B.rx[i] + (B.dt * B.vx[i] + B.dt2 * B.fx[i] / m2);
  double ry = B.ry[i] + (B.dt * B.vy[i] + B.dt2 * B.fy[i] / m2);
  double rz = B.rz[i] + (B.dt * B.vz[i] + B.dt2 * B.fz[i] / m2);
  if (rx > B.L[0]) rx = rx - B.L[0];
  if (ry > B.L[1]) ry = ry - B.L[1];
  if (rz > B.L[2]) rz = rz - B.L[2];
  if (rx < 0.0) rx = B.L[0] + rx;
  if (ry < 0.0) ry = B.L[1] + ry;
  if (rz < 0.0) rz = B.L[2] + rz;
  B.rx[i] = rx; B.ry[i] = ry; B.rz[i] = rz;
}

// ----------------------------------------------------------------------
// Per-box scalars in two launches.  k_scal_part, per ion item q of a box
// (nspart of them): the kinetic sums m v^2 of its ions and its share of
// the real-space item partials (pot, diagonal virial, pairs) and of the
// reciprocal energy blocks, each a fixed-order workgroup sum.  k_box_scalars,
// one workgroup per box, sums the nspart partials (fixed order) into
// real-space pot and virial, reciprocal energy, kinetic energy as mdc's
// calc_kin (rv.c:68-86) and pressure as calc_press (pt.c:8-21).  A step
// runs k_scal_part's work inside the force assembly (k_force_vv<VV, 1>).
// flags bit0: count a completed step.
// ----------------------------------------------------------------------
__global__ __launch_bounds__(kIonThreads) void k_scal_part(const BoxDesc* __restrict__ boxes,
                                                           const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  double vx = 0.0, vy = 0.0, vz = 0.0;
  if (i < B.N) { vx = B.vx[i]; vy = B.vy[i]; vz = B.vz[i]; }
  scal_part_block(B, it, vx, vy, vz);
}

__global__ __launch_bounds__(kScalThreads) void k_box_scalars(const BoxDesc* __restrict__ boxes,
                                                              const int* __restrict__ box_ids,
                                                              int flags) {
  const BoxDesc& B = boxes[box_ids[blockIdx.x]];
  const int tid = threadIdx.x;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int q = tid; q < B.nspart; q += kScalThreads)
    for (int k = 0; k < 9; ++k) v[k] += B.spart[9 * (size_t)q + k];
  block_sum<kScalThreads, 9>(v);
  if (tid == 0) {
    double* S = B.scal;
    const double pself = B.ewald ? B.self_pot : 0.0;
    const double precip = B.kspace ? v[4] : 0.0;
    S[kScalPotReal] = v[0];
    S[kScalPotRecip] = precip;
    S[kScalPotSelf] = pself;
    S[kScalPot] = v[0] + precip + pself;
    S[kScalVirX] = v[1]; S[kScalVirY] = v[2]; S[kScalVirZ] = v[3];
    S[kScalKinX] = v[5] / 2.0; S[kScalKinY] = v[6] / 2.0; S[kScalKinZ] = v[7] / 2.0;
    const double kin = S[kScalKinX] + S[kScalKinY] + S[kScalKinZ];
    S[kScalKin] = kin;
    S[kScalEtot] = kin + v[0] + 0.25 * precip + pself;
    S[kScalPresX] = (2.0 * S[kScalKinX] + v[1]) / B.vol;
    S[kScalPresY] = (2.0 * S[kScalKinY] + v[2]) / B.vol;
    S[kScalPresZ] = (2.0 * S[kScalKinZ] + v[3]) / B.vol;
    S[kScalPres] = (S[kScalPresX] + S[kScalPresY] + S[kScalPresZ]) / 3.0;
    if (flags & 1) S[kScalStep] += 1.0;
    S[kScalPairs] = v[8];
  }
}

// Copies each box's r, v, scalars and observables into one staging
// buffer, box b at dst + off[b] as [r 3N][v 3N][scalars][observables]
// (one launch over the per-ion items, so a frame of many boxes is one
// contiguous device-to-host copy).
__global__ __launch_bounds__(kIonThreads) void k_snapshot(const BoxDesc* __restrict__ boxes,
                                                          const Item* __restrict__ items,
                                                          const long long* __restrict__ off,
                                                          double* __restrict__ dst) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N;
  double* o = dst + off[it.box];
  const int i = it.a + threadIdx.x;
  if (i < N) {
    o[i] = B.rx[i];
    o[N + i] = B.ry[i];
    o[2 * N + i] = B.rz[i];
    o[3 * N + i] = B.vx[i];
    o[4 * N + i] = B.vy[i];
    o[5 * N + i] = B.vz[i];
  }
  if (it.a == 0) {
    const int t = threadIdx.x;
    if (t < kScalCount) o[6 * N + t] = B.scal[t];
    if (t < kObsLen) o[6 * N + kScalCount + t] = B.obs[t];
  }
}

// v <- scale * v (the velocity scaling of control_temp, pt.c:101-105)
__global__ __launch_bounds__(kIonThreads) void k_scale_v(const BoxDesc* __restrict__ boxes,
                                                         const Item* __restrict__ items,
                                                         double scale) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int i = it.a + threadIdx.x;
  if (i >= B.N) return;
  B.vx[i] *= scale;
  B.vy[i] *= scale;
  B.vz[i] *= scale;
}

// mdc's moment_correction (rv.c:88-116) of box B by one workgroup: the
// mean momentum per ion is removed, v -= (sum m v / N) / m.
__device__ __forceinline__ void moment_correct_box(const BoxDesc& B) {
  __shared__ double mean[3];
  const int tid = threadIdx.x;
  double v[3] = {0, 0, 0};
  for (int i = tid; i < B.N; i += kScalThreads) {
    const double m = B.m[B.typ[i]];
    v[0] += m * B.vx[i];
    v[1] += m * B.vy[i];
    v[2] += m * B.vz[i];
  }
  block_sum<kScalThreads, 3>(v);
  if (tid == 0) {
    mean[0] = v[0] / (double)B.N;
    mean[1] = v[1] / (double)B.N;
    mean[2] = v[2] / (double)B.N;
  }
  __syncthreads();
  for (int i = tid; i < B.N; i += kScalThreads) {
    const double m = B.m[B.typ[i]];
    B.vx[i] -= mean[0] / m;
    B.vy[i] -= mean[1] / m;
    B.vz[i] -= mean[2] / m;
  }
}

__global__ __launch_bounds__(kScalThreads) void k_moment_correction(
    const BoxDesc* __restrict__ boxes, const int* __restrict__ box_ids) {
  moment_correct_box(boxes[box_ids[blockIdx.x]]);
}

// ----------------------------------------------------------------------
// The rest of one iteration of mdc's step loop after the integration and
// calc_kin/calc_press (main.c:44-60), one box per workgroup, in mdc's
// order:
//  * calc_msd (msd.c:20-81): at the origin step the positions are
//    recorded; afterwards each ion's displacement since the last sample,
//    folded once by the strict half-box tests, is added to its summed
//    displacement, and the MSD per species is the mean of |d|^2;
//  * control_temp (pt.c:85-112) while step < t_control_until: the kinetic
//    energy of every step is summed, and every t_control_step steps the
//    velocities are scaled by sqrt(target / mean) and the sum restarts;
//  * moment_correction (rv.c:88-116) every moment_every steps.
// step >= 0 is mdc's sys.step of the iteration; step < 0 takes the box's
// own count of completed steps.  phases: bit 0 MSD, bit 1 thermostat,
// bit 2 moment correction, each also subject to the box's own switches;
// bit 3: the MSD origin is recorded only at step 1 (mdc's rule), else also
// at the first sample after the observables were (re)armed.
// ----------------------------------------------------------------------
__global__ __launch_bounds__(kScalThreads) void k_observe(const BoxDesc* __restrict__ boxes,
                                                          const int* __restrict__ box_ids,
                                                          long long step_arg, int phases) {
  const BoxDesc& B = boxes[box_ids[blockIdx.x]];
  const int tid = threadIdx.x;
  const int N = B.N;
  double* O = B.obs;
  const long long step = step_arg >= 0 ? step_arg : (long long)B.scal[kScalStep];
  if ((phases & 1) && B.msd_on) {
    const bool origin = step == 1 || (!(phases & 8) && O[kObsMsdInit] == 0.0);
    __syncthreads();  // every thread has read the origin flag
    double* ox = B.msd_old;
    double* oy = ox + N;
    double* oz = oy + N;
    if (origin) {
      for (int i = tid; i < N; i += kScalThreads) {
        ox[i] = B.rx[i];
        oy[i] = B.ry[i];
        oz[i] = B.rz[i];
      }
      if (tid == 0) O[kObsMsdInit] = 1.0;
    } else {
      double* dx = B.msd_d;
      double* dy = dx + N;
      double* dz = dy + N;
      double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // |d|^2 sums, then ion counts, per species
      for (int i = tid; i < N; i += kScalThreads) {
        const double rx = B.rx[i], ry = B.ry[i], rz = B.rz[i];
        double ddx = ox[i] - rx, ddy = oy[i] - ry, ddz = oz[i] - rz;
        ox[i] = rx;
        oy[i] = ry;
        oz[i] = rz;
        if (ddx > B.halfL[0]) ddx -= B.L[0];
        if (ddx < -B.halfL[0]) ddx += B.L[0];
        if (ddy > B.halfL[1]) ddy -= B.L[1];
        if (ddy < -B.halfL[1]) ddy += B.L[1];
        if (ddz > B.halfL[2]) ddz -= B.L[2];
        if (ddz < -B.halfL[2]) ddz += B.L[2];
        const double sx = dx[i] + ddx, sy = dy[i] + ddy, sz = dz[i] + ddz;
        dx[i] = sx;
        dy[i] = sy;
        dz[i] = sz;
        const double d2 = sx * sx + sy * sy + sz * sz;
        const int t = B.typ[i];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          v[k] += t == k ? d2 : 0.0;
          v[4 + k] += t == k ? 1.0 : 0.0;
        }
      }
      block_sum<kScalThreads, 8>(v);
      if (tid == 0)
        for (int k = 0; k < B.kinds && k < 4; ++k) {
          O[kObsMsd + k] = v[k] / v[4 + k];
          O[kObsMsdIons + k] = v[4 + k];
        }
    }
  }
  if ((phases & 2) && B.t_control_step > 0 && step < B.t_control_until) {
    __shared__ double scale;
    __shared__ int scaled;
    if (tid == 0) {
      double count = O[kObsCount] + 1.0;
      double oldk = O[kObsOldK] + B.scal[kScalKin];
      scaled = 0;
      if (step % B.t_control_step == 0) {
        const double averaged = oldk / count;
        oldk = 0.0;
        count = 0.0;
        scale = sqrt(B.t_control_ke / averaged);
        scaled = 1;
        O[kObsScalings] += 1.0;
      }
      O[kObsCount] = count;
      O[kObsOldK] = oldk;
    }
    __syncthreads();
    if (scaled) {
      const double sc = scale;
      for (int i = tid; i < N; i += kScalThreads) {
        B.vx[i] *= sc;
        B.vy[i] *= sc;
        B.vz[i] *= sc;
      }
    }
  }
  if ((phases & 4) && B.moment_every > 0 && step % B.moment_every == 0) {
    __syncthreads();  // the scaled velocities are visible to every thread
    moment_correct_box(B);
  }
}

}  // namespace mdg
