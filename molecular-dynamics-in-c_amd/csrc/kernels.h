// kernels.h -- CDNA4 kernels of libmdgpu (included by k_fp64.hip and
// k_mixed.hip; the FP64 unit is compiled with -ffp-contract=off so that
// expressions mdc evaluates without fusion round the same way on gfx950).
//
// All kernels walk a list of work items; an item names a box (BoxDesc) and
// a tile of that box, so one launch serves any mix of box sizes.  Every
// reduction has a fixed order that depends on the box alone: a box gives
// identical bits alone or inside any batch.
#pragma once

#include <type_traits>

#include "mdg_internal.h"
#include "mori_erfc.h"

namespace mdg {

// ------------------------------------------------------------ complex
template <typename R> struct Cx { R x, y; };

__device__ __forceinline__ Cx<double> cmul(Cx<double> a, Cx<double> b) {
  return {fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ Cx<float> cmul(Cx<float> a, Cx<float> b) {
  return {fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)};
}
// acc += a * b (complex), four fused multiply-adds
__device__ __forceinline__ void cmac(Cx<double>& acc, Cx<double> a, Cx<double> b) {
  acc.x = fma(a.x, b.x, acc.x); acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y); acc.y = fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cmac(Cx<float>& acc, Cx<float> a, Cx<float> b) {
  acc.x = fmaf(a.x, b.x, acc.x); acc.x = fmaf(-a.y, b.y, acc.x);
  acc.y = fmaf(a.x, b.y, acc.y); acc.y = fmaf(a.y, b.x, acc.y);
}
// Read-only per-box tables (k columns, segments, per-k weights) are never
// written while a kernel that reads them runs; reading them through the
// constant address space lets uniform indices become scalar-cache loads.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
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
// mori_erfc (FP64, mdc's form): mori_erfc.h

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
// read by all lanes at once.  An item covers one home unit (64 consecutive
// ions, or one cell) and one split of its partners (a contiguous index
// range, or a contiguous run of the 27 neighbour cells).  With one split
// and all pairs, partners are visited in ascending index -- the order in
// which mdc's i<j loop adds contributions to any given ion -- so
// table-mode forces are bit-identical to mdc's.  Energy and virial count
// each unordered pair once (partner index > own index).
//
// rsplit == 1, MODE 0: f0 <- f, f <- real-space force (clear_foc + real_space)
// rsplit == 1, MODE 1: f += real-space force (real_space() after clear_foc)
// rsplit > 1: the split's partial force goes to rpf; k_force_vv sums them.
// ======================================================================
enum { kRealTable64 = 0, kRealAnalytic64 = 1, kRealAnalytic32 = 2, kRealTable32 = 3 };

struct RealAcc {
  double fx = 0, fy = 0, fz = 0, pot = 0, vx = 0, vy = 0, vz = 0;
  unsigned long long np = 0;
};

struct HomeCtx {
  int i, ti;
  bool valid;
  double xi, yi, zi;
  float xf, yf, zf;       // screening coordinates (see real_tile)
  const double2* t64[4];  // table rows against partner species 0..3
  const float2* t32[4];
  double Zi;
  int ta[4];              // analytic64 pair type against partner species
};

struct BoxCtx {
  double hx, hy, hz, Lx, Ly, Lz, rc2;
  int nb;
  float rc2f, alpha, half_alpha, two_a3, inv_rho, inv_rho_l2e;
  float gauss_l2;  // -alpha^2 log2(e): e^{-alpha^2 r^2} = 2^(gauss_l2 r^2)
  float hxf, hyf, hzf, Lxf, Lyf, Lzf;
  float rc2m;  // screening cutoff: rc^2 plus a margin above FP32 rounding
};

template <int KIND>
__device__ __forceinline__ void make_home(const BoxDesc& B, int i, int ti, bool valid, double x,
                                          double y, double z, HomeCtx& h) {
  h.i = i; h.ti = ti; h.valid = valid; h.xi = x; h.yi = y; h.zi = z;
  h.Zi = B.Z[ti];
  for (int s = 0; s < 4; ++s) {
    const int t = s < B.kinds ? B.lookup[ti * B.kinds + s] : 0;
    if constexpr (KIND == kRealTable64) h.t64[s] = B.table + (size_t)t * B.table_nb;
    if constexpr (KIND == kRealTable32) h.t32[s] = B.table32 + (size_t)t * B.table_nb;
    h.ta[s] = t;
  }
}

// mdc's single minimum-image fold: '>' on every axis first, then '<'
// (mdc/program/real.c:43-49)
__device__ __forceinline__ void fold(const BoxCtx& c, double& dx, double& dy, double& dz) {
  if (dx > c.hx) dx -= c.Lx;
  if (dy > c.hy) dy -= c.Ly;
  if (dz > c.hz) dz -= c.Lz;
  if (dx < -c.hx) dx += c.Lx;
  if (dy < -c.hy) dy += c.Ly;
  if (dz < -c.hz) dz += c.Lz;
}

// FP32 analytic pair function: -U'(r)/r and U(r) of erfc Coulomb + BMH.
// dr is the FP64 squared distance, rounded to FP32.  The Born-Mayer
// exponential multiplies r's rounding error by r / rho (about 9 for alkali
// halides), so near contact that term errs by about 1e-6 of itself, ten
// times the Coulomb term; the per-ion force error it leaves stays below
// the reciprocal part's (measured with and without an FP64 residual of r).
// The hardware reciprocal square root, reciprocal and base-2 exponential
// (v_rsq_f32, v_rcp_f32, v_exp_f32, about 1 ulp) stand in for the
// correctly rounded forms, which only add a rescaling for results in the
// denormal range; here the exponents are about -20 (Gaussian, (alpha rc)^2
// log2 e at an Ewald accuracy of 1e-6) and -32 (Born-Mayer at 10 A), and a
// result below 2^-126 would be flushed to a zero that is negligible anyway.
// Pair constants of one (home species, partner species) combination, two
// 16-byte LDS rows: zz = Z_i Z_j, cb, sigma log2(e) / rho, C, D of the BMH form, the
// force's 6 C and 8 D, and zz times the Gaussian's prefactor 2 alpha / sqrt(pi).
struct alignas(16) PairRow {
  float zz, cb, sig_l2e, c6, d8, c6x6, d8x8, zz2a3;
};

__device__ __forceinline__ void pair32_fast(const BoxCtx& c, const PairRow& w, double dr,
                                            float& p, float& f) {
  // r and 1/r in FP32 from v_rsq (about 1 ulp)
  const float r2 = (float)dr;
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * rinv;
  const float rinv2 = rinv * rinv;
  // erfc(z) = t e^{-z^2} R(t), t = 1/(1 + z/2), with R(t) = erfc(z) e^{z^2} / t
  // as a degree-8 polynomial fitted for this library (relative least squares
  // on 4000 Chebyshev nodes of z in [0, 8], coefficients rounded to FP32):
  // relative error below 1e-7 for alpha r <= 8, i.e. every pair within the
  // cutoff down to an Ewald accuracy of 1e-27 (the plan refuses larger
  // alpha rc).  The Gaussian e^{-z^2} is the force's own, so one
  // exponential serves both terms; its argument -alpha^2 log2(e) r^2 comes
  // from r^2 directly (one rounding fewer than squaring alpha r).
  const float t = __builtin_amdgcn_rcpf(fmaf(c.half_alpha, r, 1.0f));  // z = alpha r
  float q = fmaf(t, -0.0588165745f, 0.283194035f);
  q = fmaf(t, q, -0.483994156f);
  q = fmaf(t, q, 0.273955554f);
  q = fmaf(t, q, -0.0234501231f);
  q = fmaf(t, q, 0.201180458f);
  q = fmaf(t, q, 0.243530557f);
  q = fmaf(t, q, 0.282309562f);
  q = fmaf(t, q, 0.282090604f);
  const float g = __builtin_amdgcn_exp2f(c.gauss_l2 * r2);
  const float erfc_r = t * g * q * rinv;
  // Born-Mayer: 2^((sigma - r) log2(e) / rho) by one FMA from the row's
  // sigma log2(e) / rho; the FP32 exponent's rounding (about 5e-7 near
  // contact, below r's own) moves the term by 4e-8 |exponent| elsewhere,
  // large only where the term itself is negligible (exponent -32 at 10 A)
  const float eb = w.cb * __builtin_amdgcn_exp2f(fmaf(-r, c.inv_rho_l2e, w.sig_l2e));
  const float r6 = rinv2 * rinv2 * rinv2;
  const float r8 = r6 * rinv2;
  // C r^-6 + D r^-8 and its force 6 C r^-8 + 8 D r^-10
  const float ze = w.zz * erfc_r;
  p = ze + eb - r6 * fmaf(w.d8, rinv2, w.c6);
  f = fmaf(w.zz2a3, g, ze) * rinv2 + (eb * c.inv_rho * rinv - r8 * fmaf(w.d8x8, rinv2, w.c6x6));
}

// FP64 analytic pair function in the operation order of mdc's mk_table
// (mdc/program/control.c:219-240), erfc per pair as pr-mdc's ewald1 does
// (pr-mdc/program/ewald.c:37-43).
__device__ __forceinline__ void pair64(const BoxDesc& B, int t, double ZZ, double dr, double& p,
                                       double& f) {
  const double dr6 = dr * dr * dr * dr * dr * dr;
  const double dr8 = dr6 * dr * dr;
  const double adr = B.alpha * dr;
  const double erfc_per_dr = mori_erfc(adr) / dr;
  const double ebmh = B.cb[t] * exp((B.sig[t] - dr) / B.rho);
  p = ZZ * erfc_per_dr + (ebmh - B.Cd[t] / dr6 - B.Dd[t] / dr8);
  const double fe = ZZ * (2.0 * B.a3 * exp(-adr * adr) + erfc_per_dr) / (dr * dr);
  const double fr = (ebmh / B.rho - 6.0 * B.Cd[t] / (dr6 * dr) - 8.0 * B.Dd[t] / (dr8 * dr)) / dr;
  f = fe + fr;
}

// One LDS tile of partners u in [0, cnt), in two phases.
//  1. Screen: every lane tests all partners in FP32 (screening coordinates
//     jf / h.xf, a cutoff with a margin larger than their rounding) and
//     keeps a 64-bit mask of candidates.  Partners are read as broadcasts.
//  2. Evaluate: each lane pops its own candidates in increasing u (so a
//     lane's force is still summed in partner order) and applies the exact
//     FP64 fold and cutoff of the unscreened code; lanes work on different
//     partners at once instead of idling while one lane's pair runs.
// The accepted pairs and the per-ion summation order are those of a plain
// loop over all partners.  FOLD32: screening coordinates are absolute and
// need the minimum-image fold (all-pairs path); cell-list partners come
// pre-shifted to the home cell's image, relative to its corner.
template <int KIND, bool FOLD32>
__device__ __forceinline__ void real_tile(const BoxDesc& B, const BoxCtx& c, const HomeCtx& h,
                                          const double* jx, const double* jy, const double* jz,
                                          const float4* jf, const int* jt, const int* jid,
                                          int cnt, RealAcc& a, int& err) {
  unsigned long long m = 0ull;
  if (h.valid) {
    for (int u = 0; u < cnt; ++u) {
      const float4 p = jf[u];
      float dx = h.xf - p.x, dy = h.yf - p.y, dz = h.zf - p.z;
      if constexpr (FOLD32) {
        dx = dx > c.hxf ? dx - c.Lxf : (dx < -c.hxf ? dx + c.Lxf : dx);
        dy = dy > c.hyf ? dy - c.Lyf : (dy < -c.hyf ? dy + c.Lyf : dy);
        dz = dz > c.hzf ? dz - c.Lzf : (dz < -c.hzf ? dz + c.Lzf : dz);
      }
      const float dr = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      if (dr <= c.rc2m) m |= 1ull << u;
    }
  }
  if constexpr (KIND == kRealTable64) {
    // Table mode: each round a lane takes up to kTabBatch of its candidates
    // in increasing u and issues their table reads together, then adds
    // them in that same order, so an ion's force is summed exactly as in a
    // one-pair-at-a-time loop while several reads are in flight.
    constexpr int kTabBatch = 4;
    while (__ballot(m != 0ull)) {
      double2 pf[kTabBatch];
      double ex[kTabBatch], ey[kTabBatch], ez[kTabBatch];
      bool ok[kTabBatch], upper[kTabBatch];
#pragma unroll
      for (int q = 0; q < kTabBatch; ++q) {
        ok[q] = false;
        upper[q] = false;
        ex[q] = ey[q] = ez[q] = 0.0;
        pf[q] = make_double2(0.0, 0.0);
        if (m == 0ull) continue;
        const int u = __builtin_ctzll(m);
        m &= m - 1ull;
        if (jid[u] == h.i) continue;
        double dx = h.xi - jx[u], dy = h.yi - jy[u], dz = h.zi - jz[u];
        fold(c, dx, dy, dz);
        const double dr = dx * dx + dy * dy + dz * dz;  // unfused, as mdc
        if (!(dr <= c.rc2)) continue;
        const int bin = (int)((sqrt(dr) - 0.5) * 1000.0);  // truncating, real.c:62
        if (bin < 0 || bin >= c.nb) { err |= kErrRange; continue; }
        const int s = jt[u];
        const double2* row =
            s == 0 ? h.t64[0] : (s == 1 ? h.t64[1] : (s == 2 ? h.t64[2] : h.t64[3]));
        pf[q] = row[bin];
        ok[q] = true;
        upper[q] = jid[u] > h.i;
        ex[q] = dx; ey[q] = dy; ez[q] = dz;
      }
#pragma unroll
      for (int q = 0; q < kTabBatch; ++q) {
        if (!ok[q]) continue;
        const double dx = ex[q], dy = ey[q], dz = ez[q];
        a.fx += pf[q].y * dx; a.fy += pf[q].y * dy; a.fz += pf[q].y * dz;
        if (upper[q]) {
          a.pot += pf[q].x;
          a.vx += pf[q].y * dx * dx; a.vy += pf[q].y * dy * dy; a.vz += pf[q].y * dz * dz;
          ++a.np;
        }
      }
    }
    return;
  }
  while (__ballot(m != 0ull)) {
    if (m == 0ull) continue;
    const int u = __builtin_ctzll(m);
    m &= m - 1ull;
    if (jid[u] == h.i) continue;
    double dx = h.xi - jx[u], dy = h.yi - jy[u], dz = h.zi - jz[u];
    fold(c, dx, dy, dz);
    if constexpr (KIND == kRealAnalytic64) {
      const double dr = dx * dx + dy * dy + dz * dz;
      if (!(dr <= c.rc2)) continue;
      const int s = jt[u];
      const int t = s == 0 ? h.ta[0] : (s == 1 ? h.ta[1] : (s == 2 ? h.ta[2] : h.ta[3]));
      double p, f;
      pair64(B, t, h.Zi * B.Z[s], sqrt(dr), p, f);
      a.fx += f * dx; a.fy += f * dy; a.fz += f * dz;
      if (jid[u] > h.i) {
        a.pot += p;
        a.vx += f * dx * dx; a.vy += f * dy * dy; a.vz += f * dz * dz;
        ++a.np;
      }
    }
  }
}

__device__ __forceinline__ void real_store(const BoxDesc& B, const HomeCtx& h, int split,
                                           int mode, const RealAcc& a) {
  if (!h.valid) return;
  const int i = h.i;
  if (B.rsplit > 1) {
    double* p = B.rpf + (size_t)split * 3 * B.N;
    p[i] = a.fx; p[B.N + i] = a.fy; p[2 * B.N + i] = a.fz;
  } else if (mode == 0) {
    B.fx0[i] = B.fx[i]; B.fy0[i] = B.fy[i]; B.fz0[i] = B.fz[i];
    B.fx[i] = a.fx; B.fy[i] = a.fy; B.fz[i] = a.fz;
  } else {
    B.fx[i] += a.fx; B.fy[i] += a.fy; B.fz[i] += a.fz;
  }
}

template <int KIND>
__global__ __launch_bounds__(kRealThreads) void k_real(const BoxDesc* __restrict__ boxes,
                                                       const Item* __restrict__ items, int mode) {
  static_assert(KIND == kRealTable64 || KIND == kRealAnalytic64,
                "the mixed real space is k_real_win");
  __shared__ double jx[kRealThreads], jy[kRealThreads], jz[kRealThreads];
  __shared__ int jt[kRealThreads], jid[kRealThreads];
  __shared__ float4 jf[kRealThreads];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int lane = threadIdx.x;
  const int N = B.N, split = it.c, nsplit = B.rsplit;
  int err = 0;
  BoxCtx c;
  c.hx = B.halfL[0]; c.hy = B.halfL[1]; c.hz = B.halfL[2];
  c.Lx = B.L[0]; c.Ly = B.L[1]; c.Lz = B.L[2];
  c.rc2 = B.rc2;
  c.nb = B.table_nb;
  c.rc2f = (float)B.rc2; c.alpha = (float)B.alpha; c.two_a3 = (float)(2.0 * B.a3);
  c.inv_rho = (float)(1.0 / B.rho);
  c.inv_rho_l2e = (float)(1.4426950408889634 / B.rho);  // log2(e) / rho
  c.hxf = (float)c.hx; c.hyf = (float)c.hy; c.hzf = (float)c.hz;
  c.Lxf = (float)c.Lx; c.Lyf = (float)c.Ly; c.Lzf = (float)c.Lz;
  // FP32 screening coordinates are within ~1e-5 A: a relative margin of
  // 1e-4 on rc^2 (plus 1e-3 A^2) keeps every pair the exact test accepts
  c.rc2m = (float)(B.rc2 * 1.0001 + 1e-3);
  RealAcc tot;  // pot / virial / pair count of the item
  if (B.all_pairs) {
    const int i = it.a + lane;
    const bool valid = i < N;
    const int ii = valid ? i : 0;
    HomeCtx h;
    make_home<KIND>(B, i, B.typ[ii], valid, B.rx[ii], B.ry[ii], B.rz[ii], h);
    h.xf = (float)h.xi; h.yf = (float)h.yi; h.zf = (float)h.zi;
    const int js = (int)((long)split * N / nsplit), je = (int)((long)(split + 1) * N / nsplit);
    RealAcc a;
    for (int j0 = js; j0 < je; j0 += kRealThreads) {
      const int j = j0 + lane;
      if (j < je) {
        jx[lane] = B.rx[j]; jy[lane] = B.ry[j]; jz[lane] = B.rz[j];
        jt[lane] = B.typ[j]; jid[lane] = j;
        jf[lane] = make_float4((float)B.rx[j], (float)B.ry[j], (float)B.rz[j], 0.f);
      }
      __syncthreads();
      real_tile<KIND, true>(B, c, h, jx, jy, jz, jf, jt, jid, min(kRealThreads, je - j0), a,
                            err);
      __syncthreads();
    }
    real_store(B, h, split, mode, a);
    tot = a;
  } else {
    const int cell = it.a;
    const int ncx = B.nc[0], ncy = B.nc[1], ncz = B.nc[2];
    const int cx = cell % ncx, cy = (cell / ncx) % ncy, cz = cell / (ncx * ncy);
    const int hs = B.cell_start[cell], he = B.cell_start[cell + 1];
    const int n0 = split * 27 / nsplit, n1 = (split + 1) * 27 / nsplit;
    // this item's home blocks: it.d, it.d + cell_hb, ...
    for (int h0 = hs + it.d * kRealThreads; h0 < he; h0 += B.cell_hb * kRealThreads) {
      const int hp = h0 + lane;
      const bool valid = hp < he;
      const int hq = valid ? hp : hs;
      HomeCtx h;
      make_home<KIND>(B, B.sorted[hq], B.styp[hq], valid, B.sx[hq], B.sy[hq], B.sz[hq], h);
      // screening coordinates relative to the home cell's corner
      const double ox0 = cx * (B.L[0] / ncx), oy0 = cy * (B.L[1] / ncy), oz0 = cz * (B.L[2] / ncz);
      h.xf = (float)(h.xi - ox0); h.yf = (float)(h.yi - oy0); h.zf = (float)(h.zi - oz0);
      RealAcc a;
      for (int n = n0; n < n1; ++n) {
        const int oz = n / 9 - 1, oy = (n / 3) % 3 - 1, ox = n % 3 - 1;
        const int nzc = (cz + oz + ncz) % ncz, nyc = (cy + oy + ncy) % ncy,
                  nxc = (cx + ox + ncx) % ncx;
        const int nb = (nzc * ncy + nyc) * ncx + nxc;
        const int ns = B.cell_start[nb], ne = B.cell_start[nb + 1];
        // the neighbour's image next to the home cell, relative to its corner
        const double shx = (cx + ox < 0 ? -B.L[0] : (cx + ox >= ncx ? B.L[0] : 0.0)) - ox0;
        const double shy = (cy + oy < 0 ? -B.L[1] : (cy + oy >= ncy ? B.L[1] : 0.0)) - oy0;
        const double shz = (cz + oz < 0 ? -B.L[2] : (cz + oz >= ncz ? B.L[2] : 0.0)) - oz0;
        for (int j0 = ns; j0 < ne; j0 += kRealThreads) {
          const int jp = j0 + lane;
          if (jp < ne) {
            jx[lane] = B.sx[jp]; jy[lane] = B.sy[jp]; jz[lane] = B.sz[jp];
            jt[lane] = B.styp[jp]; jid[lane] = B.sorted[jp];
            jf[lane] = make_float4((float)(B.sx[jp] + shx), (float)(B.sy[jp] + shy),
                                   (float)(B.sz[jp] + shz), 0.f);
          }
          __syncthreads();
          real_tile<KIND, false>(B, c, h, jx, jy, jz, jf, jt, jid, min(kRealThreads, ne - j0), a,
                                 err);
          __syncthreads();
        }
      }
      real_store(B, h, split, mode, a);
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
// Real space, mixed precision: screened windows of partners, each pair
// evaluated once (Newton's third law, as mdc's i < j loop applies it,
// mdc/program/real.c:70-75).
//
// A home unit meets only the partners that follow it: all pairs, the ions
// from its own first position on in the box's visiting order (position p
// is ion p rstride mod N, which spreads home blocks and windows over the
// box; a later position within the unit); cell lists, its own cell
// (j > i) and the 13 cells of the half shell.  The home ion's share of a
// pair force is summed in its lane in FP64 over the item's partners, then
// added to the box's 64-bit fixed-point array (x 2^32) with integer
// atomics; the partner's share goes into fixed-point accumulators in LDS
// per window slot, then into the same array.  Integer sums are exact and
// order-free, so the result has the same bits whatever order the atomics
// land in, and no per-split partial forces are written; k_force_vv
// converts the array.  (mode is unused: the force assembly always follows.)
//
// The pairs, pair function and per-ion FP64 home sums are k_real's mixed
// kinds, organised for issue rate:
//  * Partners are staged kRealWin at a time (a window), so one lane's pop
//    loop spans four tiles and the lanes' candidate counts even out.
//  * Screening runs on 16-bit fixed-point coordinates q = x * 65536 / L
//    (mod 2^16).  The packed 16-bit difference of two such coordinates
//    wraps to the minimum image by itself, and two v_dot2_i32_i16 give
//    the squared distance against a threshold: five instructions per
//    candidate pair, no fold.  With u = min(L) / 65536 each axis' wrapped
//    difference is at most |dx| / u + 1, so every pair within rc passes
//    d^2 <= (rc / u + sqrt 3)^2 (a superset: the exact test follows).
//  * A lane's candidates are bits of 32-bit words kept in LDS; the lane
//    pops them in increasing partner order and applies the FP64
//    minimum-image fold and the FP64 cutoff test, then the FP32 pair
//    function.  The pair force f d is formed once, in FP64 from the FP32
//    f/r; the home sums, energy and virial accumulate it in FP64.
// The fold is x - L rint(x / L), which equals mdc's strict-comparison
// fold (mdc/program/real.c:43-49) except for a separation of exactly L/2
// on an axis, where both images are at the same distance.
// ======================================================================
constexpr int kRealWin = 128;                 // partners per window
constexpr int kRealWinWords = kRealWin / 32;  // candidate words per lane

struct alignas(16) WinRec {
  double x, y, z;
  int id, t;
};

// 16-bit fixed-point coordinates of one ion: (qx | qy << 16, qz)
__device__ __forceinline__ uint2 quantize(double x, double y, double z, double sx, double sy,
                                          double sz) {
  const unsigned qx = (unsigned)(int)rint(x * sx) & 0xffffu;
  const unsigned qy = (unsigned)(int)rint(y * sy) & 0xffffu;
  const unsigned qz = (unsigned)(int)rint(z * sz) & 0xffffu;
  return make_uint2(qx | (qy << 16), qz);
}

typedef short mdg_s2 __attribute__((ext_vector_type(2)));

// 32 candidates u0 .. u0+31 of one lane: bit 31-k is set when partner
// u0+k passes the screen.  thr1 = threshold + 1.
__device__ __forceinline__ unsigned screen32(uint2 hq, const uint2* pq, int u0, int thr1) {
  unsigned m = 0u;
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const uint2 p = pq[u0 + k];
    const mdg_s2 dxy = __builtin_bit_cast(mdg_s2, hq.x) - __builtin_bit_cast(mdg_s2, p.x);
    const mdg_s2 dz = __builtin_bit_cast(mdg_s2, hq.y) - __builtin_bit_cast(mdg_s2, p.y);
    int a = __builtin_amdgcn_sdot2(dxy, dxy, -thr1, true);
    a = __builtin_amdgcn_sdot2(dz, dz, a, true);   // saturates: far pairs stay positive
    m = __builtin_amdgcn_alignbit(m, (unsigned)a, 31);  // m << 1 | sign(a)
  }
  return m;
}

// x * 2^32 as a 64-bit two's complement integer, truncated below 2^-32,
// for |x| <= 1e9: floor(x) in the upper word, the fraction scaled into the
// lower one.  Done in FP64: for a negative x the fraction is 1 - frac|x|,
// which FP32 rounds to its spacing near 1 (6e-8, a large error against the
// 2^-32 resolution); in FP64 it is exact for any widened FP32 value down to
// 2^-21 in magnitude and within one unit of 2^-32 below that.
__device__ __forceinline__ unsigned long long to_fix(double x) {
  const double hi = floor(x);
  const double lo = (x - hi) * 4294967296.0;  // in [0, 2^32]; 2^32 saturates to 2^32 - 1
  return ((unsigned long long)(long long)(int)hi << 32) + (unsigned long long)(unsigned)lo;
}

// to_fix of an FP64 sum, |x| clamped to 1e9 first
__device__ __forceinline__ unsigned long long to_fix64(double x) {
  return to_fix(fmin(fmax(x, -1.0e9), 1.0e9));
}

// x * 2^32 rounded to the nearest integer, as a 64-bit two's complement
// integer, for |x| < 2^19: fma(x, 2^32, 1.5 * 2^52) lies in (2^52, 2^53),
// where FP64 numbers are the integers, so it is 1.5 * 2^52 + round(x 2^32)
// exactly and its bit pattern is 0x4338000000000000 + round(x 2^32).  The
// constant's low word is zero, so the subtraction touches the high word
// only: one FMA and one integer subtraction per value.
// (two32 = 2^32 and magic = 1.5 * 2^52, passed in registers: a VOP3 FMA
// takes neither as an inline constant)
__device__ __forceinline__ unsigned long long to_fix_near(double x, double two32, double magic) {
  const double t = fma(x, two32, magic);
  const unsigned long long b = (unsigned long long)__double_as_longlong(t);
  const unsigned hi = (unsigned)(b >> 32) - 0x43380000u;
  return ((unsigned long long)hi << 32) | (unsigned)b;
}

// Largest pair force component the partner shares carry: |f d| <= 2^18
// (force units of 1e-8 N; a pair of ions at contact exerts about 1), so
// to_fix_near's range holds with room to spare.
constexpr double kPairForceMax = 262144.0;

template <int KIND>
__global__ __launch_bounds__(kRealThreads) void k_real_win(const BoxDesc* __restrict__ boxes,
                                                           const Item* __restrict__ items,
                                                           int mode) {
  static_assert(KIND == kRealAnalytic32 || KIND == kRealTable32, "mixed kinds only");
  __shared__ WinRec pr[kRealWin];
  __shared__ uint2 pq[kRealWin];
  // candidate words; word kRealWinWords is a zero sentinel that a lane out
  // of words lands on
  __shared__ unsigned pm[kRealWinWords + 1][kRealThreads];
  __shared__ unsigned long long pacc[3][kRealWin];  // partner shares, fixed point
  __shared__ int ncum[28], nst[27];
  __shared__ PairRow prow[4];  // by home species * 2 + partner species
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int lane = threadIdx.x;
  const int N = B.N, split = it.c, nsplit = B.rsplit;
  int err = 0;
  pm[kRealWinWords][lane] = 0u;  // read back by this lane only
  if constexpr (KIND == kRealAnalytic32) {
    if (lane < 4) {
      const int ti = lane >> 1, tj = lane & 1;
      const int t = B.lookup[min(ti, B.kinds - 1) * B.kinds + min(tj, B.kinds - 1)];
      PairRow w;
      w.zz = (float)(B.Z[ti] * B.Z[tj]);
      w.cb = (float)B.cb[t];
      w.sig_l2e = (float)(B.sig[t] * 1.4426950408889634 / B.rho);
      w.c6 = (float)B.Cd[t];
      w.d8 = (float)B.Dd[t];
      w.c6x6 = (float)(6.0 * B.Cd[t]);
      w.d8x8 = (float)(8.0 * B.Dd[t]);
      w.zz2a3 = (float)(B.Z[ti] * B.Z[tj] * 2.0 * B.a3);
      prow[lane] = w;
    }
  }
  BoxCtx c;
  c.Lx = B.L[0]; c.Ly = B.L[1]; c.Lz = B.L[2];
  c.rc2 = B.rc2;
  c.nb = B.table_nb;
  c.alpha = (float)B.alpha; c.two_a3 = (float)(2.0 * B.a3);
  c.half_alpha = (float)(0.5 * B.alpha);
  c.gauss_l2 = (float)(-B.alpha * B.alpha * 1.4426950408889634);
  c.inv_rho = (float)(1.0 / B.rho);
  c.inv_rho_l2e = (float)(1.4426950408889634 / B.rho);  // log2(e) / rho
  const double ilx = B.invL[0], ily = B.invL[1], ilz = B.invL[2];
  const double sx = 65536.0 * ilx, sy = 65536.0 * ily, sz = 65536.0 * ilz;
  const double hlx = 0.5 * c.Lx, hly = 0.5 * c.Ly, hlz = 0.5 * c.Lz;
  int thr1;
  bool screen_all;
  {
    const double lmin = fmin(c.Lx, fmin(c.Ly, c.Lz));
    const double s = sqrt(B.rc2) * (65536.0 / lmin) + 1.7320508075688772 + 0.01;
    // cutoffs beyond ~0.7 L exceed the 31-bit range: every partner is a candidate
    screen_all = s * s >= 2.0e9;
    thr1 = screen_all ? 0 : (int)(s * s) + 1;
  }
  const float fsmax = (float)(kPairForceMax / sqrt(B.rc2));
  const double inv_n = 1.0 / (double)N;
  RealAcc tot;
  // partner source: all pairs -> index range; cells -> run of neighbour cells.
  // Partners before self_end (a run index) pair with the home unit's own
  // ions, so only those of higher index are taken.
  int hs, he, pbeg, pend, self_end = 0x7fffffff;
  int cx = 0, cy = 0, cz = 0;
  if (B.all_pairs) {
    // partners from the unit's first ion on, in chunks of ceil(N / nsplit)
    // (a block gets as many items as its run needs; the home-side partial
    // slots it does not reach stay zero)
    hs = it.a; he = min(N, it.a + kRealThreads);
    const int chunk = (N + nsplit - 1) / nsplit;
    pbeg = hs + split * chunk; pend = min(N, pbeg + chunk);
  } else {
    const int cell = it.a;
    const int ncx = B.nc[0], ncy = B.nc[1], ncz = B.nc[2];
    cx = cell % ncx; cy = (cell / ncx) % ncy; cz = cell / (ncx * ncy);
    hs = B.cell_start[cell]; he = B.cell_start[cell + 1];
    // stencil cells 13 (the home cell) .. 26: the half shell, split into runs
    const int n0 = 13 + split * 14 / nsplit, n1 = 13 + (split + 1) * 14 / nsplit;
    self_end = n0 == 13 ? he - hs : 0;
    const int n = n0 + lane;
    int cnt = 0, st = 0;
    if (n < n1) {
      const int oz = n / 9 - 1, oy = (n / 3) % 3 - 1, ox = n % 3 - 1;
      const int nb = (((cz + oz + ncz) % ncz) * ncy + (cy + oy + ncy) % ncy) * ncx +
                     (cx + ox + ncx) % ncx;
      st = B.cell_start[nb]; cnt = B.cell_start[nb + 1] - st;
    }
    int inc = cnt;  // inclusive prefix over the run
#pragma unroll
    for (int off = 1; off < 32; off <<= 1) {
      const int v = __shfl_up(inc, off, 64);
      if (lane >= off) inc += v;
    }
    if (lane < 27) { nst[lane] = st; ncum[lane + 1] = inc; }
    if (lane == 0) ncum[0] = 0;
    pbeg = 0; pend = __shfl(inc, 26, 64);
    __syncthreads();
  }
  // cells: this item's home blocks it.d, it.d + cell_hb, ... of its cell
  const int hstep = B.all_pairs ? kRealThreads : B.cell_hb * kRealThreads;
  for (int h0 = hs + (B.all_pairs ? 0 : it.d * kRealThreads); h0 < he; h0 += hstep) {
    const int hp = h0 + lane;
    const bool valid = hp < he;
    const int hq = valid ? hp : hs;
    HomeCtx h;
    if (B.all_pairs) {
      // the ion at position hq of the box's visiting order (p -> p rstride
      // mod N); its force accumulates at fix[hq]
      const int jh = mul_mod(hq, B.rstride, N, inv_n);
      make_home<KIND>(B, hq, B.typ[jh], valid, B.rx[jh], B.ry[jh], B.rz[jh], h);
    } else
      make_home<KIND>(B, B.sorted[hq], B.styp[hq], valid, B.sx[hq], B.sy[hq], B.sz[hq], h);
    const uint2 hqz = quantize(h.xi, h.yi, h.zi, sx, sy, sz);
    // the home ion in box units (x/L per axis)
    const double hux = h.xi * ilx, huy = h.yi * ily, huz = h.zi * ilz;
    // partner positions p (index or run slot) below self_lim pair with this
    // lane's ion in the other order; self_max bounds them for the wave
    const int self_lim = B.all_pairs ? hq + 1 : min(hq - hs + 1, self_end);
    const int self_max = B.all_pairs ? he : self_end;
    const PairRow* hrow = prow + 2 * (h.ti & 1);  // this lane's two rows
    double fx_two32 = 4294967296.0, fx_magic = 6755399441055744.0;
    __asm__ volatile("" : "+v"(fx_two32), "+v"(fx_magic));
    RealAcc a;
    for (int w0 = pbeg; w0 < pend; w0 += kRealWin) {
      const int cnt = min(kRealWin, pend - w0);
      // stage the window
      for (int s = lane; s < cnt; s += kRealThreads) {
        int j;
        double x, y, z;
        int t;
        if (B.all_pairs) {
          const int jp = mul_mod(w0 + s, B.rstride, N, inv_n);  // at position w0 + s
          x = B.rx[jp]; y = B.ry[jp]; z = B.rz[jp]; t = B.typ[jp];
          j = w0 + s;  // its force accumulates at fix[w0 + s]
        } else {
          const int g = w0 + s;
          int lo = 0, hi = 26;  // last run entry with ncum[n] <= g
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (ncum[mid] <= g) lo = mid; else hi = mid - 1;
          }
          const int jp = nst[lo] + (g - ncum[lo]);
          x = B.sx[jp]; y = B.sy[jp]; z = B.sz[jp]; t = B.styp[jp]; j = B.sorted[jp];
        }
        // staged as 1/2 - x/L per axis: the pop loop folds home - partner
        // in box units with one add and one fract
        WinRec r;
        r.x = 0.5 - x * ilx; r.y = 0.5 - y * ily; r.z = 0.5 - z * ilz; r.id = j;
        // analytic: the byte offset of the partner species' pair row
        r.t = KIND == kRealAnalytic32 ? (t & 1) * (int)sizeof(PairRow) : t;
        pr[s] = r;
        pq[s] = quantize(x, y, z, sx, sy, sz);
        pacc[0][s] = 0ull; pacc[1][s] = 0ull; pacc[2][s] = 0ull;
      }
      __syncthreads();
      // screen: one candidate word per 32 partners
      const int nw = (cnt + 31) >> 5;
      unsigned nz = 0u;
      for (int w = 0; w < nw; ++w) {
        unsigned m = screen32(hqz, pq, w * 32, thr1);
        if (screen_all) m = 0xffffffffu;
        const int nv = min(32, cnt - w * 32);
        if (nv < 32) m &= ~(0xffffffffu >> nv);
        if (!valid) m = 0u;
        // each unordered pair once: partners p < self_lim are the home
        // unit's own ions at or below this lane's (all pairs: by position in
        // the visiting order; in a cell, by slot, the ascending index order
        // within the cell)
        if (w0 + 32 * w < self_max) {
          const int cut = min(max(self_lim - (w0 + 32 * w), 0), 32);
          m &= cut >= 32 ? 0u : (0xffffffffu >> cut);
        }
        pm[w][lane] = m;
        nz |= (m != 0u ? 1u : 0u) << w;
      }
      // Evaluate: each lane pops its own candidates in increasing order.
      // The body is branch-free but for the partner's atomics; a lane out of
      // candidates evaluates a discarded pair.  Energy and virial count each
      // evaluated (unordered) pair once.
      unsigned npw = 0u;  // accepted pairs of this window (32 bits suffice)
      int w = __builtin_ctz(nz | (1u << kRealWinWords));
      unsigned cur = pm[w][lane];
      if (__ballot(cur != 0u)) do {
        const bool live = cur != 0u;
        const int b = __builtin_clz(cur | 1u);
        cur &= ~(0x80000000u >> b);
        // A lane out of candidates reads slot 0, which every window stages:
        // its discarded pair is weighted by zero, and a slot the window did
        // not stage may hold another kernel's LDS bits (a NaN as FP64 would
        // survive the multiplication by zero).
        const int slot = live ? w * 32 + b : 0;
        const WinRec P = pr[slot];
        if (live && cur == 0u) {
          // w is nz's lowest bit: drop it and move to the next non-empty
          // word, or to the zero sentinel
          nz &= nz - 1u;
          w = __builtin_ctz(nz | (1u << kRealWinWords));
          cur = pm[w][lane];
        }
        // minimum image: u = x_i/L - x_j/L + 1/2 folded into [0, 1) by fract
        // (x - floor x), then d = (u - 1/2) L in [-L/2, L/2)
        const double dx = fma(__builtin_amdgcn_fract(hux + P.x), c.Lx, -hlx);
        const double dy = fma(__builtin_amdgcn_fract(huy + P.y), c.Ly, -hly);
        const double dz = fma(__builtin_amdgcn_fract(huz + P.z), c.Lz, -hlz);
        // (pairs with the home unit's lower ions are off the candidate words)
        bool ok = live;
        float p, f;
        if constexpr (KIND == kRealAnalytic32) {
          const double dr = fma(dx, dx, fma(dy, dy, dz * dz));
          ok = ok && dr <= c.rc2;
          // a discarded pair may give inf or NaN here; it is selected away
          pair32_fast(c, *reinterpret_cast<const PairRow*>(
                             reinterpret_cast<const char*>(hrow) + P.t), dr, p, f);
        } else {
          const double dr = dx * dx + dy * dy + dz * dz;
          ok = ok && dr <= c.rc2;
          int bin = (int)((sqrt(dr) - 0.5) * 1000.0);
          if (ok && (bin < 0 || bin >= c.nb)) { err |= kErrRange; ok = false; }
          bin = ok ? bin : 0;
          const int s = P.t;
          const float2* row =
              s == 0 ? h.t32[0] : (s == 1 ? h.t32[1] : (s == 2 ? h.t32[2] : h.t32[3]));
          const float2 pf = row[bin];
          p = pf.x; f = pf.y;
        }
        // the pair force f d once: +f d on the home ion, -f d on the partner
        // (fixed point); f/r clamped so that |f d| <= kPairForceMax (a
        // force far beyond any pair's) for the fixed point
        // (the selects stay in FP32, ahead of the widening: the asm keeps
        // the compiler from widening first and selecting both halves)
        float fsel = ok ? __builtin_amdgcn_fmed3f(f, -fsmax, fsmax) : 0.f;
        float psel = ok ? p : 0.f;
        __asm__ volatile("" : "+v"(fsel), "+v"(psel));
        const double fs = (double)fsel;
        const double gx = fs * dx, gy = fs * dy, gz = fs * dz;
        a.fx += gx; a.fy += gy; a.fz += gz;
        a.pot += (double)psel;
        a.vx = fma(gx, dx, a.vx); a.vy = fma(gy, dy, a.vy); a.vz = fma(gz, dz, a.vz);
        npw += (unsigned)__popcll(__ballot(ok));  // the wave's accepted pairs
        if (ok) {  // +f d here; the flush adds its negation to the partner
          atomicAdd(&pacc[0][slot], to_fix_near(gx, fx_two32, fx_magic));
          atomicAdd(&pacc[1][slot], to_fix_near(gy, fx_two32, fx_magic));
          atomicAdd(&pacc[2][slot], to_fix_near(gz, fx_two32, fx_magic));
        }
      } while (__ballot(cur != 0u));
      if (lane == 0) a.np += npw;
      __syncthreads();
      // the window's partner shares into the box's fixed-point sums
      for (int s = lane; s < cnt; s += kRealThreads) {
        const size_t id = (size_t)pr[s].id;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const unsigned long long v = pacc[q][s];
          if (v) atomicAdd(&B.fix[q * (size_t)N + id], 0ull - v);  // -f d, exact
        }
      }
      __syncthreads();
    }
    // the home ions' sums over this item's partners join the partner shares
    // in the fixed-point array (exact, so the items may land in any order)
    if (h.valid) {
      atomicAdd(&B.fix[h.i], to_fix64(a.fx));
      atomicAdd(&B.fix[(size_t)N + h.i], to_fix64(a.fy));
      atomicAdd(&B.fix[2 * (size_t)N + h.i], to_fix64(a.fz));
    }
    tot.pot += a.pot; tot.vx += a.vx; tot.vy += a.vy; tot.vz += a.vz;
    tot.np += a.np;
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
// Reciprocal space, pass A: partial structure factors over one ion chunk
//   S(k) = sum_j Z_j X_j[hx] Y_j[hy] W_j[hz],   X_j[h] = e^{2 pi i h x_j/Lx} ...
// (the sum of mdc/program/rcprcl.c:59-78 in separable form).  One wave per
// item; lane s owns one k segment (up to 8 consecutive hz of one column)
// and keeps its 8 sums in registers over the whole chunk: per (k, ion) one
// complex multiply-add, no cross-lane reduction.  Ion tiles of X, Y and
// Z*W rows are staged in LDS ([ion][h] layout, so lanes at different h
// read different banks).  Tile sums in R, chunk sums in FP64.
// ======================================================================
template <typename R>
__global__ __launch_bounds__(kSkThreads) void k_sk_seg(const BoxDesc* __restrict__ boxes,
                                                       const Item* __restrict__ items, int hs,
                                                       int hsw) {
  extern __shared__ unsigned char smem[];
  Cx<R>* Xt = reinterpret_cast<Cx<R>*>(smem);    // [kSkTile][hs]
  Cx<R>* Yt = Xt + kSkTile * hs;                  // [kSkTile][hs]
  Cx<R>* Wt = Yt + kSkTile * hs;                  // [kSkTile][hsw], hsw >= H + kSegLen - 1
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, H = B.hmax + 1, tid = threadIdx.x;
  const bool sv = tid < it.c;
  KSeg sg = {0, 0, 0, 0, 0, 0, {0, 0}};
  if (sv) {
    const cptr<KSeg> sp = as_const(B.segs) + it.b + tid;
    sg.hx = sp->hx; sg.hy = sp->hy; sg.hz0 = sp->hz0; sg.n = sp->n; sg.k0 = sp->k0;
    sg.qlo = sp->qlo;
  }
  const int ax = sg.hx < 0 ? -sg.hx : sg.hx, ay = sg.hy < 0 ? -sg.hy : sg.hy;
  const bool cxn = sg.hx < 0, cyn = sg.hy < 0;
  Cx<double> acc[kSegLen];
#pragma unroll
  for (int q = 0; q < kSegLen; ++q) acc[q] = {0.0, 0.0};
  const int jb = it.a * B.chunk, je = min(N, jb + B.chunk);
  for (int j0 = jb; j0 < je; j0 += kSkTile) {
    const int cnt = min(kSkTile, je - j0);
    // stage the tile: element e = h * kSkTile + u of each table, four
    // elements per thread in flight
    const int E = H * kSkTile;
    for (int e0 = tid; e0 < E; e0 += 4 * kSkThreads) {
      Cx<R> xv[4], yv[4], wv[4];
      int eu[4], eh[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + q * kSkThreads;
        eu[q] = e % kSkTile;
        eh[q] = e / kSkTile;
        const bool ok = e < E && eu[q] < cnt;
        const long g = (long)(ok ? eh[q] : 0) * N + j0 + (ok ? eu[q] : 0);
        xv[q] = ld_cx<R>(B.tabx, g);
        yv[q] = ld_cx<R>(B.taby, g);
        wv[q] = ld_cx<R>(B.tabw, g);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = e0 + q * kSkThreads;
        if (e < E && eu[q] < cnt) {
          const R zq = (R)B.Z[B.typ[j0 + eu[q]]];
          Xt[eu[q] * hs + eh[q]] = xv[q];
          Yt[eu[q] * hs + eh[q]] = yv[q];
          Wt[eu[q] * hsw + eh[q]] = {zq * wv[q].x, zq * wv[q].y};
        }
      }
    }
    __syncthreads();
    if (sv) {
      Cx<R> t[kSegLen];
#pragma unroll
      for (int q = 0; q < kSegLen; ++q) t[q] = {(R)0, (R)0};
      for (int u = 0; u < cnt; ++u) {
        const Cx<R> p = cmul(conj_if(Xt[u * hs + ax], cxn), conj_if(Yt[u * hs + ay], cyn));
        // hz0 and the row stride are multiples of kSegLen: aligned wide reads
        Cx<R> w[kSegLen];
        if constexpr (sizeof(R) == 4) {
          const float4* w4 = reinterpret_cast<const float4*>(Wt + u * hsw + sg.hz0);
#pragma unroll
          for (int q = 0; q < kSegLen / 2; ++q) {
            const float4 v = w4[q];
            w[2 * q] = {v.x, v.y};
            w[2 * q + 1] = {v.z, v.w};
          }
        } else {
#pragma unroll
          for (int q = 0; q < kSegLen; ++q) w[q] = Wt[u * hsw + sg.hz0 + q];
        }
#pragma unroll
        for (int q = 0; q < kSegLen; ++q) cmac(t[q], p, w[q]);
      }
#pragma unroll
      for (int q = 0; q < kSegLen; ++q) {
        acc[q].x += (double)t[q].x;
        acc[q].y += (double)t[q].y;
      }
    }
    __syncthreads();
  }
  if (sv) {
    double2* dst = B.skpart + (size_t)it.a * B.nk + sg.k0;
#pragma unroll
    for (int q = 0; q < kSegLen; ++q)
      if (q >= sg.qlo && q < sg.n) dst[q] = make_double2(acc[q].x, acc[q].y);
  }
}

// ======================================================================
// Reciprocal space, pass B: back-projection onto each ion
//   f_i += -(2/V) sum_k e^{-pi^2 m^2/alpha^2}/m^2 (c_i Ss - s_i Sc) m
// (mdc/program/rcprcl.c:84-92), c_i, s_i = Z_i cos, Z_i sin.  With
// T_k = cf_k (2Ss_k + i 2Sc_k) the per-ion sum over one column (hx, hy) is
//   sum_hz Re(X[hx] Y[hy] W[hz] T_k) = Re(Q U),  U = sum_hz W[hz] T_k,
// and the z component needs V = sum_hz hz W[hz] T_k.  Lane = ion; its
// W[0..HZ) row lives in registers; T_k comes through the scalar cache:
// two complex multiply-adds per (k, ion).  Column sums in R, sums over
// columns in FP64.
// ======================================================================
template <typename R, int HZ>
__global__ __launch_bounds__(kKfThreads) void k_kforce(const BoxDesc* __restrict__ boxes,
                                                       const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, H = B.hmax + 1;
  const int i = it.a + threadIdx.x;
  const bool valid = i < N;
  const int ii = valid ? i : 0;
  Cx<R> W[HZ];
#pragma unroll
  for (int h = 0; h < HZ; ++h) W[h] = h < H ? ld_cx<R>(B.tabw, (long)h * N + ii) : Cx<R>{0, 0};
  // Weight rows (4 columns x (A, B, hz A, hz B)) staged in LDS, up to
  // kKfRows rows at a time, copied contiguously; lanes then read each row
  // as a broadcast.  Per hz row one W[hz] register feeds four independent
  // complex accumulations.
  using R4 = typename std::conditional<sizeof(R) == 8, double4, float4>::type;
  __shared__ R4 Tl[kKfRows * 4];
  const R4* T;
  if constexpr (sizeof(R) == 8) T = reinterpret_cast<const R4*>(B.kab);
  else T = reinterpret_cast<const R4*>(B.kab32);
  const cptr<KQuad> quads = as_const(B.quads);
  double Fx = 0.0, Fy = 0.0, Fz = 0.0;
  const int qend = it.b + it.c;
  int q = it.b;
  while (q < qend) {
    const int r0 = quads[q].row0;
    int qe = q + 1;
    while (qe < qend && quads[qe].row0 + quads[qe].nrow - r0 <= kKfRows) ++qe;
    const int nslots = 4 * (quads[qe - 1].row0 + quads[qe - 1].nrow - r0);
    for (int e = threadIdx.x; e < nslots; e += kKfThreads) Tl[e] = T[4 * (size_t)r0 + e];
    __syncthreads();
    for (int qq = q; qq < qe; ++qq) {
      const int lo = quads[qq].lo, hi = lo + quads[qq].nrow, nc = quads[qq].ncol;
      int hxs[4], hys[4];
      Cx<R> Q[4], U[4], V[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        hxs[c] = quads[qq].hx[c];
        hys[c] = quads[qq].hy[c];
        U[c] = {0, 0};
        V[c] = {0, 0};
        Q[c] = {0, 0};
        if (c < nc) {
          const int ax = hxs[c] < 0 ? -hxs[c] : hxs[c], ay = hys[c] < 0 ? -hys[c] : hys[c];
          Q[c] = cmul(conj_if(ld_cx<R>(B.tabx, (long)ax * N + ii), hxs[c] < 0),
                      conj_if(ld_cx<R>(B.taby, (long)ay * N + ii), hys[c] < 0));
        }
      }
      const R4* t = Tl + 4 * (quads[qq].row0 - r0 - lo);
#pragma unroll
      for (int hz = 0; hz < HZ; ++hz) {
        if (hz >= lo && hz < hi) {
          const R4 t0 = t[4 * hz], t1 = t[4 * hz + 1], t2 = t[4 * hz + 2], t3 = t[4 * hz + 3];
          cmac(U[0], Cx<R>{t0.x, t0.y}, W[hz]); cmac(V[0], Cx<R>{t0.z, t0.w}, W[hz]);
          cmac(U[1], Cx<R>{t1.x, t1.y}, W[hz]); cmac(V[1], Cx<R>{t1.z, t1.w}, W[hz]);
          cmac(U[2], Cx<R>{t2.x, t2.y}, W[hz]); cmac(V[2], Cx<R>{t2.z, t2.w}, W[hz]);
          cmac(U[3], Cx<R>{t3.x, t3.y}, W[hz]); cmac(V[3], Cx<R>{t3.z, t3.w}, W[hz]);
        }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const R G = Q[c].x * U[c].x - Q[c].y * U[c].y;
        const R Gz = Q[c].x * V[c].x - Q[c].y * V[c].y;
        Fx += (double)hxs[c] * (double)G;
        Fy += (double)hys[c] * (double)G;
        Fz += (double)Gz;
      }
    }
    __syncthreads();
    q = qe;
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
