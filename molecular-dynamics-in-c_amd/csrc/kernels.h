// kernels.h -- CDNA4 kernels of libmdgpu (included by k_fp64.hip and
// k_mixed.hip; the FP64 unit is compiled with -ffp-contract=off so that
// expressions mdc evaluates without fusion round the same way on gfx950).
//
// All kernels walk a list of work items; an item names a box (BoxDesc) and
// a tile of that box, so one launch serves any mix of box sizes.  Every
// reduction has a fixed order that depends on the box alone: a box gives
// identical bits alone or inside any batch.
#pragma once

#include "mdg_internal.h"

namespace mdg {

// ------------------------------------------------------------ complex
template <typename R> struct Cx { R x, y; };

__device__ __forceinline__ Cx<double> cmul(Cx<double> a, Cx<double> b) {
  return {fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ Cx<float> cmul(Cx<float> a, Cx<float> b) {
  return {fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)};
}
template <typename R> __device__ __forceinline__ Cx<R> conj_if(Cx<R> a, bool c) {
  return {a.x, c ? -a.y : a.y};
}
template <typename R> __device__ __forceinline__ Cx<R> ld_cx(const void* tab, long idx);
template <> __device__ __forceinline__ Cx<double> ld_cx<double>(const void* tab, long idx) {
  double2 v = reinterpret_cast<const double2*>(tab)[idx];
  return {v.x, v.y};
}
template <> __device__ __forceinline__ Cx<float> ld_cx<float>(const void* tab, long idx) {
  float2 v = reinterpret_cast<const float2*>(tab)[idx];
  return {v.x, v.y};
}

// ------------------------------------------------------------ erfc
// Mori's erfc (step h = 1/2), the form mdc tabulates (mdc/program/ext.c:73-104):
//   erfc(x) ~ (x/pi) e^{-x^2} [1/(2x^2) + sum_{n=1..13} e^{-n^2/4}/(n^2/4 + x^2)]
//             - 2/(e^{4 pi x} - 1)
// Weights e^{-n^2/4} at the 16 digits mdc uses.
__device__ __forceinline__ double mori_erfc(double x) {
  const double w[13] = {
      0.7788007830714049,    0.3678794411714423,    0.1053992245618643,
      0.01831563888873417,   0.001930454136227709,  0.0001234098040866795,
      4.785117392129008e-6,  1.125351747192591e-7,  1.605228055185611e-9,
      1.388794386496401e-11, 7.287724095819693e-14, 2.319522830243569e-16,
      4.4777324417183e-19};
  double x2 = x * x;
  double s = 0.5 / x2;
#pragma unroll
  for (int n = 1; n <= 13; ++n) s += w[n - 1] / (0.25 * (double)(n * n) + x2);
  double v = -2.0 / (-1.0 + exp(12.56637061435917 * x)) + (0.3183098861837907 * x * s) / exp(x2);
  return x >= 0.0 ? v : 2.0 - v;
}

// erfc(z) for z >= 0 in FP32: t = 1/(1 + z/2), erfc = t exp(-z^2 + P(t)) with
// the Chebyshev-fitted P of Numerical Recipes in C (2nd ed., sec. 6.2),
// fractional error below 1.2e-7 in exact arithmetic.
__device__ __forceinline__ float erfc_cheb(float z) {
  float t = __frcp_rn(1.0f + 0.5f * z);
  float p = fmaf(t, 0.17087277f, -0.82215223f);
  p = fmaf(t, p, 1.48851587f);
  p = fmaf(t, p, -1.13520398f);
  p = fmaf(t, p, 0.27886807f);
  p = fmaf(t, p, -0.18628806f);
  p = fmaf(t, p, 0.09678418f);
  p = fmaf(t, p, 0.37409196f);
  p = fmaf(t, p, 1.00002368f);
  p = fmaf(t, p, -1.26551223f);
  return t * __expf(fmaf(-z, z, p));
}

// ------------------------------------------------------------ pair terms
// -U'(r)/r and U(r) of erfc-screened Coulomb + Born-Mayer-Huggins at exact r,
// operation order of mdc's mk_table (mdc/program/control.c:219-240); used by
// the FP64 analytic mode as pr-mdc's ewald1 evaluates erfc per pair
// (pr-mdc/program/ewald.c:37-43).
__device__ __forceinline__ void pair_analytic64(const BoxDesc& B, int t, double ZZ, double dr,
                                                double& p, double& f) {
  double dr6 = dr * dr * dr * dr * dr * dr;
  double dr8 = dr6 * dr * dr;
  double adr = B.alpha * dr;
  double erfc_per_dr = mori_erfc(adr) / dr;
  double ebmh = B.cb[t] * exp((B.sig[t] - dr) / B.rho);
  p = ZZ * erfc_per_dr + (ebmh - B.Cd[t] / dr6 - B.Dd[t] / dr8);
  double fe = ZZ * (2.0 * B.a3 * exp(-adr * adr) + erfc_per_dr) / (dr * dr);
  double fr = (ebmh / B.rho - 6.0 * B.Cd[t] / (dr6 * dr) - 8.0 * B.Dd[t] / (dr8 * dr)) / dr;
  f = fe + fr;
}

// FP32 constants of the pair function per type, staged once per work item.
struct Pair32 {
  float rc2, alpha, two_a3;
  float zz[3], cb[3], sig[3], inv_rho, c6[3], d8[3];
};

__device__ __forceinline__ void pair_analytic32(const Pair32& P, int t, float r2, float& p,
                                                float& f) {
  const float kLog2e = 1.4426950408889634f;
  float rinv = __frsqrt_rn(r2);
  float r = r2 * rinv;
  float rinv2 = rinv * rinv;
  float z = P.alpha * r;
  float erfc_r = erfc_cheb(z) * rinv;
  float g = exp2f(-z * z * kLog2e);
  float eb = P.cb[t] * exp2f((P.sig[t] - r) * P.inv_rho * kLog2e);
  float r6 = rinv2 * rinv2 * rinv2;
  float r8 = r6 * rinv2;
  p = P.zz[t] * erfc_r + eb - P.c6[t] * r6 - P.d8[t] * r8;
  f = P.zz[t] * (P.two_a3 * g + erfc_r) * rinv2 +
      (eb * P.inv_rho * rinv - 6.0f * P.c6[t] * r8 - 8.0f * P.d8[t] * r8 * rinv2);
}

// ------------------------------------------------------------ reductions
// Fixed-order sum over the 64 lanes of a wave.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ======================================================================
// Real space: erfc + BMH pair forces (mdc/program/real.c:15-84).
//
// One wave per item.  The force on an ion is gathered by its own lane
// (no scatter, no atomics); partner ions are staged 64 at a time in LDS and
// read by all lanes at once.  All-pairs items visit partners in ascending
// index, which is the order in which mdc's i<j loop adds contributions to
// any given ion, so table-mode forces are bit-identical to mdc's.  Energy
// and virial count each unordered pair once (partner index > own index).
//
// MODE 0: f0 <- f, f <- real-space force (mdc's clear_foc then real_space).
// MODE 1: f += real-space force (real_space() after a caller's clear_foc).
// ======================================================================
enum { kRealTable64 = 0, kRealAnalytic64 = 1, kRealAnalytic32 = 2, kRealTable32 = 3 };

template <int KIND>
struct RealAcc {
  double fx = 0, fy = 0, fz = 0, pot = 0, vx = 0, vy = 0, vz = 0;
  unsigned long long np = 0;
};

template <int KIND>
__device__ __forceinline__ void real_tile(const BoxDesc& B, const Pair32& P, int i, int ti,
                                          double xi, double yi, double zi, bool valid,
                                          const double* jx, const double* jy, const double* jz,
                                          const int* jt, const int* jid, int cnt,
                                          RealAcc<KIND>& a, int& err) {
  const double hx = B.halfL[0], hy = B.halfL[1], hz = B.halfL[2];
  const double Lx = B.L[0], Ly = B.L[1], Lz = B.L[2];
  float ffx = 0.f, ffy = 0.f, ffz = 0.f, fpot = 0.f, fvx = 0.f, fvy = 0.f, fvz = 0.f;
  for (int u = 0; u < cnt; ++u) {
    const int j = jid[u];
    double dx = xi - jx[u], dy = yi - jy[u], dz = zi - jz[u];
    // mdc's single fold, '>' on all axes first, then '<' (real.c:43-49)
    if (dx > hx) dx -= Lx;
    if (dy > hy) dy -= Ly;
    if (dz > hz) dz -= Lz;
    if (dx < -hx) dx += Lx;
    if (dy < -hy) dy += Ly;
    if (dz < -hz) dz += Lz;
    const bool upper = j > i;
    if constexpr (KIND == kRealTable64 || KIND == kRealAnalytic64 || KIND == kRealTable32) {
      const double dr = dx * dx + dy * dy + dz * dz;  // unfused, as mdc
      if (!valid || j == i || !(dr <= B.rc2)) continue;
      const int t = B.lookup[ti * B.kinds + jt[u]];
      if constexpr (KIND == kRealTable64 || KIND == kRealTable32) {
        const int bin = (int)((sqrt(dr) - 0.5) * 1000.0);  // truncating, real.c:62
        if (bin < 0 || bin >= B.table_nb) { err |= kErrRange; continue; }
        if constexpr (KIND == kRealTable64) {
          const double2 pf = B.table[t * B.table_nb + bin];
          a.fx += pf.y * dx; a.fy += pf.y * dy; a.fz += pf.y * dz;
          if (upper) {
            a.pot += pf.x;
            a.vx += pf.y * dx * dx; a.vy += pf.y * dy * dy; a.vz += pf.y * dz * dz;
            ++a.np;
          }
        } else {
          const float2 pf = B.table32[t * B.table_nb + bin];
          const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
          ffx = fmaf(pf.y, fdx, ffx); ffy = fmaf(pf.y, fdy, ffy); ffz = fmaf(pf.y, fdz, ffz);
          if (upper) {
            fpot += pf.x;
            fvx = fmaf(pf.y * fdx, fdx, fvx); fvy = fmaf(pf.y * fdy, fdy, fvy);
            fvz = fmaf(pf.y * fdz, fdz, fvz);
            ++a.np;
          }
        }
      } else {
        double p, f;
        pair_analytic64(B, t, B.Z[ti] * B.Z[jt[u]], sqrt(dr), p, f);
        a.fx += f * dx; a.fy += f * dy; a.fz += f * dz;
        if (upper) {
          a.pot += p;
          a.vx += f * dx * dx; a.vy += f * dy * dy; a.vz += f * dz * dz;
          ++a.np;
        }
      }
    } else {  // kRealAnalytic32: fold in FP64, pair arithmetic in FP32
      const float fdx = (float)dx, fdy = (float)dy, fdz = (float)dz;
      const float r2 = fmaf(fdx, fdx, fmaf(fdy, fdy, fdz * fdz));
      if (!valid || j == i || !(r2 <= P.rc2)) continue;
      const int t = B.lookup[ti * B.kinds + jt[u]];
      float p, f;
      pair_analytic32(P, t, r2, p, f);
      ffx = fmaf(f, fdx, ffx); ffy = fmaf(f, fdy, ffy); ffz = fmaf(f, fdz, ffz);
      if (upper) {
        fpot += p;
        fvx = fmaf(f * fdx, fdx, fvx); fvy = fmaf(f * fdy, fdy, fvy); fvz = fmaf(f * fdz, fdz, fvz);
        ++a.np;
      }
    }
  }
  if constexpr (KIND == kRealAnalytic32 || KIND == kRealTable32) {
    // tile sums (at most 64 pairs) in FP32, running sums in FP64
    a.fx += ffx; a.fy += ffy; a.fz += ffz; a.pot += fpot;
    a.vx += fvx; a.vy += fvy; a.vz += fvz;
  }
}

template <int KIND>
__device__ __forceinline__ void real_store(const BoxDesc& B, int i, bool valid, int mode,
                                           const RealAcc<KIND>& a) {
  if (!valid) return;
  if (mode == 0) {
    B.fx0[i] = B.fx[i]; B.fy0[i] = B.fy[i]; B.fz0[i] = B.fz[i];
    B.fx[i] = a.fx; B.fy[i] = a.fy; B.fz[i] = a.fz;
  } else {
    B.fx[i] += a.fx; B.fy[i] += a.fy; B.fz[i] += a.fz;
  }
}

template <int KIND>
__global__ __launch_bounds__(kRealThreads) void k_real(const BoxDesc* __restrict__ boxes,
                                                       const Item* __restrict__ items, int mode) {
  __shared__ double jx[kRealThreads], jy[kRealThreads], jz[kRealThreads];
  __shared__ int jt[kRealThreads], jid[kRealThreads];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int lane = threadIdx.x;
  int err = 0;
  Pair32 P;
  if constexpr (KIND == kRealAnalytic32) {
    P.rc2 = (float)B.rc2; P.alpha = (float)B.alpha; P.two_a3 = (float)(2.0 * B.a3);
    P.inv_rho = (float)(1.0 / B.rho);
    for (int ti = 0; ti < 2; ++ti)
      for (int tj = ti; tj < 2; ++tj) {
        const int t = B.lookup[ti * B.kinds + tj];
        P.zz[t] = (float)(B.Z[ti] * B.Z[tj]);
      }
    for (int t = 0; t < 3; ++t) {
      P.cb[t] = (float)B.cb[t]; P.sig[t] = (float)B.sig[t];
      P.c6[t] = (float)B.Cd[t]; P.d8[t] = (float)B.Dd[t];
    }
  }
  RealAcc<KIND> tot;  // item totals of pot / virial (per lane, then wave)
  if (B.all_pairs) {
    const int i = it.a + lane;
    const bool valid = i < B.N;
    const int ii = valid ? i : 0;
    const double xi = B.rx[ii], yi = B.ry[ii], zi = B.rz[ii];
    const int ti = B.typ[ii];
    RealAcc<KIND> a;
    for (int j0 = 0; j0 < B.N; j0 += kRealThreads) {
      const int j = j0 + lane;
      if (j < B.N) {
        jx[lane] = B.rx[j]; jy[lane] = B.ry[j]; jz[lane] = B.rz[j];
        jt[lane] = B.typ[j]; jid[lane] = j;
      }
      __syncthreads();
      const int cnt = min(kRealThreads, B.N - j0);
      real_tile<KIND>(B, P, i, ti, xi, yi, zi, valid, jx, jy, jz, jt, jid, cnt, a, err);
      __syncthreads();
    }
    real_store<KIND>(B, i, valid, mode, a);
    tot = a;
  } else {
    const int c = it.a;
    const int ncx = B.nc[0], ncy = B.nc[1], ncz = B.nc[2];
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    const int hs = B.cell_start[c], he = B.cell_start[c + 1];
    for (int h0 = hs; h0 < he; h0 += kRealThreads) {
      const int hp = h0 + lane;
      const bool valid = hp < he;
      const int hq = valid ? hp : hs;
      const int i = B.sorted[hq];
      const double xi = B.sx[hq], yi = B.sy[hq], zi = B.sz[hq];
      const int ti = B.styp[hq];
      RealAcc<KIND> a;
      for (int oz = -1; oz <= 1; ++oz) {
        const int nz = (cz + oz + ncz) % ncz;
        for (int oy = -1; oy <= 1; ++oy) {
          const int ny = (cy + oy + ncy) % ncy;
          for (int ox = -1; ox <= 1; ++ox) {
            const int nx = (cx + ox + ncx) % ncx;
            const int nb = (nz * ncy + ny) * ncx + nx;
            const int ns = B.cell_start[nb], ne = B.cell_start[nb + 1];
            for (int j0 = ns; j0 < ne; j0 += kRealThreads) {
              const int jp = j0 + lane;
              if (jp < ne) {
                jx[lane] = B.sx[jp]; jy[lane] = B.sy[jp]; jz[lane] = B.sz[jp];
                jt[lane] = B.styp[jp]; jid[lane] = B.sorted[jp];
              }
              __syncthreads();
              const int cnt = min(kRealThreads, ne - j0);
              real_tile<KIND>(B, P, i, ti, xi, yi, zi, valid, jx, jy, jz, jt, jid, cnt, a, err);
              __syncthreads();
            }
          }
        }
      }
      real_store<KIND>(B, i, valid, mode, a);
      tot.pot += a.pot; tot.vx += a.vx; tot.vy += a.vy; tot.vz += a.vz; tot.np += a.np;
    }
  }
  const double pot = wave_sum(tot.pot), vx = wave_sum(tot.vx), vy = wave_sum(tot.vy),
               vz = wave_sum(tot.vz);
  const unsigned long long np = wave_sum_u64(tot.np);
  if (err) atomicOr(B.err, err);
  if (lane == 0) {
    double* rp = B.rpart + 4 * (size_t)it.b;
    rp[0] = pot; rp[1] = vx; rp[2] = vy; rp[3] = vz;
    B.rpairs[it.b] = np;
  }
}

// ======================================================================
// Reciprocal space, pass A: partial structure factors
//   S(k) = sum_j Z_j exp(2 pi i h.r_j / L)  over one chunk of ions
// (mdc/program/rcprcl.c:59-78).  A thread owns kSkIpt ions; for each k
// column (hx, hy) it forms Z X[hx] Y[hy] (X, Y: per-ion phase tables) and
// steps hz with one complex multiply per k.  Sub-blocks of kSkSub k
// vectors are summed over the workgroup through LDS in a fixed order.
// ======================================================================
template <typename R>
__global__ __launch_bounds__(kSkThreads) void k_sk_partial(const BoxDesc* __restrict__ boxes,
                                                           const Item* __restrict__ items) {
  __shared__ double red[2 * kSkSub][kSkThreads + 2];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, tid = threadIdx.x;
  const void* tx = B.tabx;
  const void* ty = B.taby;
  R zq[kSkIpt];
  Cx<R> w1[kSkIpt];
  int ion[kSkIpt];
#pragma unroll
  for (int q = 0; q < kSkIpt; ++q) {
    const int i = it.a * kSkChunk + q * kSkThreads + tid;
    const bool v = i < N;
    ion[q] = v ? i : 0;
    zq[q] = v ? (R)B.Z[B.typ[i]] : (R)0;
    w1[q] = ld_cx<R>(B.tabw, ion[q]);
  }
  const int v = tid >> 4, seg = tid & 15;
  for (int c = it.b; c < it.b + it.c; ++c) {
    const KCol col = B.cols[c];
    const int ax = col.hx < 0 ? -col.hx : col.hx, ay = col.hy < 0 ? -col.hy : col.hy;
    Cx<R> t[kSkIpt];
#pragma unroll
    for (int q = 0; q < kSkIpt; ++q) {
      Cx<R> X = conj_if(ld_cx<R>(tx, (long)ax * N + ion[q]), col.hx < 0);
      Cx<R> Y = conj_if(ld_cx<R>(ty, (long)ay * N + ion[q]), col.hy < 0);
      Cx<R> a = cmul(X, Y);
      if (col.hz0) a = cmul(a, w1[q]);
      t[q] = {a.x * zq[q], a.y * zq[q]};
    }
    for (int sb = 0; sb < col.len; sb += kSkSub) {
      const int nsub = min(kSkSub, col.len - sb);
      R ax_[kSkSub], ay_[kSkSub];
#pragma unroll
      for (int s = 0; s < kSkSub; ++s) { ax_[s] = 0; ay_[s] = 0; }
#pragma unroll
      for (int q = 0; q < kSkIpt; ++q) {
#pragma unroll
        for (int s = 0; s < kSkSub; ++s) {
          if (s < nsub) {
            ax_[s] += t[q].x; ay_[s] += t[q].y;
            t[q] = cmul(t[q], w1[q]);
          }
        }
      }
#pragma unroll
      for (int s = 0; s < kSkSub; ++s) {
        red[2 * s][tid] = (double)ax_[s];
        red[2 * s + 1][tid] = (double)ay_[s];
      }
      __syncthreads();
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u) acc += red[v][u * 16 + seg];
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 16);
      if (seg == 0 && (v >> 1) < nsub) {
        double* dst = reinterpret_cast<double*>(B.skpart + (size_t)it.a * B.nk + col.k0 + sb + (v >> 1));
        dst[v & 1] = acc;
      }
      __syncthreads();
    }
  }
}

// ======================================================================
// Reciprocal space, pass B: back-projection of S(k) onto each ion
//   f_i += -(2/V) sum_k e^{-pi^2 m^2/alpha^2}/m^2 (c_i Ss - s_i Sc) m
// (mdc/program/rcprcl.c:84-92), with c_i, s_i the ion's Z cos / Z sin.
// One thread per ion over a group of k columns; A = cf 2Ss and B = cf 2Sc
// are the same for all lanes and come through the scalar cache.  Column
// sums run in R, the cross-column sums in FP64.
// ======================================================================
template <typename R>
__global__ __launch_bounds__(kKfThreads) void k_kforce(const BoxDesc* __restrict__ boxes,
                                                       const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N;
  const int i = it.a + threadIdx.x;
  const bool valid = i < N;
  const int ii = valid ? i : 0;
  const Cx<R> w1 = ld_cx<R>(B.tabw, ii);
  const R* ab;
  if constexpr (sizeof(R) == 8) ab = reinterpret_cast<const R*>(B.kab);
  else ab = reinterpret_cast<const R*>(B.kab32);
  double Fx = 0.0, Fy = 0.0, Fz = 0.0;
  const int cend = it.b + it.c;
  for (int c = it.b; c < cend; ++c) {
    const KCol col = B.cols[c];
    const int ax = col.hx < 0 ? -col.hx : col.hx, ay = col.hy < 0 ? -col.hy : col.hy;
    Cx<R> t = cmul(conj_if(ld_cx<R>(B.tabx, (long)ax * N + ii), col.hx < 0),
                   conj_if(ld_cx<R>(B.taby, (long)ay * N + ii), col.hy < 0));
    if (col.hz0) t = cmul(t, w1);
    R G = 0, Gz = 0, hz = (R)col.hz0;
    const R* p = ab + 2 * (size_t)col.k0;
    for (int j = 0; j < col.len; ++j) {
      const R A = p[2 * j], Bv = p[2 * j + 1];
      const R g = t.x * A - t.y * Bv;
      G += g;
      Gz += g * hz;
      hz += (R)1;
      t = cmul(t, w1);
    }
    Fx += (double)col.hx * (double)G;
    Fy += (double)col.hy * (double)G;
    Fz += (double)Gz;
  }
  if (valid) {
    const double zq = B.Z[B.typ[i]];
    double* fp = B.fpart + (size_t)it.d * 3 * N;
    fp[i] = -zq * Fx * B.invL[0];
    fp[N + i] = -zq * Fy * B.invL[1];
    fp[2 * N + i] = -zq * Fz * B.invL[2];
  }
}

}  // namespace mdg
