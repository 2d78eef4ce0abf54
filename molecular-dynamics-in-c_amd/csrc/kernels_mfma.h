// kernels_mfma.h -- the reciprocal-space passes of the mixed path on the
// matrix cores: the back-projection (this header comment) and the
// structure factor (its own section below).
//
// Per k column (hx, hy) and ion i the back-projection needs
//   U_i = sum_hz T_k W_i[hz],   V_i = sum_hz hz T_k W_i[hz]
// (T_k = cf_k (2Ss + i 2Sc), W_i[hz] = e^{2 pi i hz z_i / Lz}), then
// G = Re(X_i[hx] Y_i[hy] U_i) and Gz likewise from V_i (the sum of
// mdc/program/rcprcl.c:84-92 in separable form).  Over 8 columns and 32
// ions these are one real matrix product
//   C[(column, Re U | Im U | Re V | Im V)][ion] = sum_(hz, re|im) A B,
// done with v_mfma_f32_32x32x16_f16.  FP16 keeps 11 bits, so every operand
// is carried as hi + lo halves (hi = fp16(x), lo = fp16(x - hi)) and each
// product as hi*hi + hi*lo + lo*hi in FP32 accumulators: about 22 bits per
// product, the FP32 arithmetic of the VALU path.  The weights of a column
// are scaled by powers of two (largest |value| in [2^10, 2^11)) so that
// they sit in FP16's range; the scale is undone in the per-column epilogue
// weights.
//
// Operand layouts of v_mfma_f32_32x32x16_f16 (lane l, element e):
//   A[row = l % 32][k = 8 (l / 32) + e],   B[k = 8 (l / 32) + e][col = l % 32],
//   C[row = 8 (v / 4) + 4 (l / 32) + v % 4][col = l % 32] for result v.
// Here k = 2 hz' + (re|im), hz' = 8 kb + k / 2 within K block kb; a lane's
// 8 B elements are W[4 (l/32) .. 4 (l/32) + 3] (re, im) of its ion; its 16
// results are columns 2 j + l/32 (j = v / 4) with the four sums v % 4.
#pragma once

#include "kernels.h"

namespace mdg {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f16_t __attribute__((ext_vector_type(16)));
typedef unsigned int u4_t __attribute__((ext_vector_type(4)));
typedef float f4v_t __attribute__((ext_vector_type(4)));
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

// (re, im) as FP16 hi and lo pairs, hi = fp16(q), lo = fp16(q - hi): one
// packed conversion per pair and two exact FP32 residuals
__device__ __forceinline__ void split_f16x2(f2v_t q, h2_t& hi, h2_t& lo) {
  hi = __builtin_convertvector(q, h2_t);
  const f2v_t back = __builtin_convertvector(hi, f2v_t);
  lo = __builtin_convertvector(f2v_t{q.x - back.x, q.y - back.y}, h2_t);
}

// Byte pointer in the global address space: a load through it takes a
// uniform 64-bit base and a 32-bit per-lane byte offset.
typedef const __attribute__((address_space(1))) unsigned char* gbytes;
template <typename T>
__device__ __forceinline__ T gload(gbytes base, unsigned off) {
  return *reinterpret_cast<const __attribute__((address_space(1))) T*>(base + off);
}

// x with its sign flipped where s is negative
__device__ __forceinline__ float flip_sign(float x, float s) {
  return __uint_as_float(__float_as_uint(x) ^ (__float_as_uint(s) & 0x80000000u));
}

__device__ __forceinline__ void split_f16(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

__device__ __forceinline__ u4_t pack_h8(const _Float16 (&v)[8]) {
  h8_t h;
#pragma unroll
  for (int e = 0; e < 8; ++e) h[e] = v[e];
  return __builtin_bit_cast(u4_t, h);
}

// two FP16 halves in one 32-bit word, a in the low half
__device__ __forceinline__ unsigned pack2_f16(_Float16 a, _Float16 b) {
  return (unsigned)__builtin_bit_cast(unsigned short, a) |
         ((unsigned)__builtin_bit_cast(unsigned short, b) << 16);
}

// ----------------------------------------------------------------------
// Every per-step ion table of the mixed k-space, one thread per ion, each
// phase computed once (FP64 sincospi, rounded to FP32):
//  * X[h], Y[h] (FP32 complex, h <= hmax) for both passes;
//  * the back-projection's B operands: W[hz] as FP16 hi/lo records,
//    [hz/4][N] x 32 B (8 hi halves: re, im of hz = 4q .. 4q+3; then the 8
//    lo halves), zero past hmax;
//  * the structure factor's B operands: per group of 4 ions and hz, four
//    16-byte vectors (Re S hi, Re S lo, Im S hi, Im S lo) holding the 4
//    ions' (Z W_r, -Z W_i) or (Z W_i, Z W_r) halves, zero past hmax and for
//    the padding ions N .. npad-1.  A group's 4 threads exchange their
//    halves through LDS, so each vector is one 16-byte store.
// ----------------------------------------------------------------------
__global__ __launch_bounds__(kIonThreads) void k_mixed_tables(const BoxDesc* __restrict__ boxes,
                                                              const Item* __restrict__ items) {
  __shared__ unsigned sbx[2][4][kIonThreads];  // by hz parity: vector x thread
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, tid = threadIdx.x;
  const int i = it.a + tid;
  const bool ion = i < N;
  const bool rec = i < B.npad;  // threads that store structure-factor vectors
  double sx = 0.0, sy = 0.0, sz = 0.0;
  float zq = 0.f;
  if (ion) {
    sx = B.rx[i] * B.invL[0];
    sy = B.ry[i] * B.invL[1];
    sz = B.rz[i] * B.invL[2];
    zq = (float)B.Z[B.typ[i]];
    float2* X = reinterpret_cast<float2*>(B.tabx);
    float2* Y = reinterpret_cast<float2*>(B.taby);
    // e^{2 pi i h s} by FP64 angle addition from h = 1: the rounding drift
    // after h steps is ~h 1e-16, far below the FP32 entries stored
    double s1x, c1x, s1y, c1y;
    sincospi(2.0 * sx, &s1x, &c1x);
    sincospi(2.0 * sy, &s1y, &c1y);
    double cx = 1.0, snx = 0.0, cy = 1.0, sny = 0.0;
    for (int h = 0; h <= B.hmax; ++h) {
      X[(size_t)h * N + i] = make_float2((float)cx, (float)snx);
      Y[(size_t)h * N + i] = make_float2((float)cy, (float)sny);
      const double tx = cx * c1x - snx * s1x, ty = cy * c1y - sny * s1y;
      snx = cx * s1x + snx * c1x;
      sny = cy * s1y + sny * c1y;
      cx = tx;
      cy = ty;
    }
  }
  double s1z = 0.0, c1z = 1.0, cz = 1.0, snz = 0.0;
  if (ion) sincospi(2.0 * sz, &s1z, &c1z);
  u4_t* wb = reinterpret_cast<u4_t*>(B.wb);
  u4_t* sb = reinterpret_cast<u4_t*>(B.sb);
  const int HP = 16 * B.sknt, HB = 8 * B.kbk;
  const int q = tid >> 2, e2 = tid & 3;  // the group and vector this thread stores
  for (int q4 = 0; q4 < max(HB, HP) / 4; ++q4) {
    float c[4], s[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int hz = 4 * q4 + e;
      c[e] = s[e] = 0.f;
      if (ion && hz <= B.hmax) {
        c[e] = (float)cz;
        s[e] = (float)snz;
      }
      const double t = cz * c1z - snz * s1z;  // advance to hz + 1
      snz = cz * s1z + snz * c1z;
      cz = t;
    }
    if (ion && 4 * q4 < HB) {
      _Float16 hv[8], lv[8];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        split_f16(c[e], hv[2 * e], lv[2 * e]);
        split_f16(s[e], hv[2 * e + 1], lv[2 * e + 1]);
      }
      u4_t* dst = wb + ((size_t)q4 * N + i) * 2;
      dst[0] = pack_h8(hv);
      dst[1] = pack_h8(lv);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int hz = 4 * q4 + e;
      if (hz >= HP) break;  // uniform
      const float wr = zq * c[e], wi = zq * s[e];
      _Float16 a, al, b, bl, d, dl, f, fl;
      split_f16(wr, a, al);
      split_f16(-wi, b, bl);
      split_f16(wi, d, dl);
      split_f16(wr, f, fl);
      const int buf = hz & 1;
      sbx[buf][0][tid] = pack2_f16(a, b);    // Re S: (Z W_r, -Z W_i)
      sbx[buf][1][tid] = pack2_f16(al, bl);
      sbx[buf][2][tid] = pack2_f16(d, f);    // Im S: (Z W_i, Z W_r)
      sbx[buf][3][tid] = pack2_f16(dl, fl);
      __syncthreads();  // also orders the reads of hz - 1 before the next writes of hz + 1
      if (rec) {
        const u4_t v = {sbx[buf][e2][4 * q], sbx[buf][e2][4 * q + 1], sbx[buf][e2][4 * q + 2],
                        sbx[buf][e2][4 * q + 3]};
        sb[((size_t)(it.a / 4 + q) * HP + hz) * 4 + e2] = v;
      }
    }
  }
}

// ----------------------------------------------------------------------
// A operands of one tile (8 columns), one wave: lane l packs row l % 32 of
// every K block; rows are (column 2 (r/8) + (r/4) % 2, sum r % 4) so that
// the result layout gives each lane all four sums of its columns.  Also
// the per-column epilogue weights hx/sU, hy/sU, 1/sV.
// ----------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void k_kf_pack(const BoxDesc* __restrict__ boxes,
                                                   const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int t = it.a, l = threadIdx.x, r = l & 31, g = l >> 5;
  const KfTile tl = B.kft[t];
  const int ct = 2 * (r >> 3) + ((r >> 2) & 1), q = r & 3;
  const bool cv = ct < tl.nc;
  KCol col = {0, 0, 0, 0, 0, {0, 0, 0}};
  if (cv) col = B.cols[tl.c0 + ct];
  // column scales: largest |T| and |hz T| to [2^10, 2^11).  Eight lanes
  // scan each of the tile's columns (weights in k order), then combine.
  double mu = 0.0, mv = 0.0;
  {
    const int m = l >> 3, sub = l & 7;
    if (m < tl.nc) {
      const KCol cm = B.cols[tl.c0 + m];
      for (int j = sub; j < cm.len; j += 8) {
        const double* w = B.kab + 4 * (size_t)(cm.k0 + j);
        mu = fmax(mu, fmax(fabs(w[0]), fabs(w[1])));
        mv = fmax(mv, fmax(fabs(w[2]), fabs(w[3])));
      }
    }
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      mu = fmax(mu, __shfl_xor(mu, off));
      mv = fmax(mv, __shfl_xor(mv, off));
    }
    mu = __shfl(mu, 8 * ct);  // this lane's column
    mv = __shfl(mv, 8 * ct);
  }
  int eu = 0, ev = 0;
  (void)frexp(mu, &eu);
  (void)frexp(mv, &ev);
  const double su = mu > 0.0 ? ldexp(1.0, 11 - eu) : 1.0;
  const double sv = mv > 0.0 ? ldexp(1.0, 11 - ev) : 1.0;
  const bool isv = q >= 2;
  const double sc = isv ? sv : su;
  // the tile's 16-byte units: its 8 column records, then its K blocks
  u4_t* kfa = reinterpret_cast<u4_t*>(B.kfa) + (size_t)t * kf_tile_stride(B.kbk);
  // blocks past the tile's hz span are never staged: not written
  for (int kb = 0; kb < tl.kb_hi; ++kb) {
    _Float16 hv[8], lv[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int hz = 8 * kb + 4 * g + (e >> 1), p = e & 1;
      float v = 0.f;
      if (cv && hz >= col.hz0 && hz < col.hz0 + col.len) {
        const double* w = B.kab + 4 * (size_t)(col.k0 + hz - col.hz0) + (isv ? 2 : 0);
        const double tr = w[0] * sc, ti = w[1] * sc;
        // Re rows take (Tr, -Ti) against (Re W, Im W); Im rows (Ti, Tr)
        v = (float)((q & 1) == 0 ? (p == 0 ? tr : -ti) : (p == 0 ? ti : tr));
      }
      split_f16(v, hv[e], lv[e]);
    }
    u4_t* dst = kfa + kKfTileCols + (size_t)kb * 2 * kWave + l;
    dst[0] = pack_h8(hv);
    dst[kWave] = pack_h8(lv);
  }
  if (q == 0 && g == 0) {  // one lane per tile column ct, padding included
    const int bits = abs(col.hx) | (abs(col.hy) << 8);
    reinterpret_cast<float4*>(kfa)[(ct & 1) * 4 + (ct >> 1)] =
        cv ? make_float4((float)(col.hx / su), (float)(col.hy / su), (float)(1.0 / sv),
                         __int_as_float(bits))
           : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// ----------------------------------------------------------------------
// Back-projection.  A workgroup of kKfMfmaWaves waves owns 32 ions per
// wave (two lanes per ion, one per half of every tile's columns) and a run
// of tiles; the ions' W operands stay in registers.  The tiles' packed A
// operands and column records are the same for every wave, so each tile
// is copied once per workgroup into LDS, double-buffered: the copy of tile
// t+1 is in flight while tile t is computed, and the vector-memory pipe
// carries a tile once per workgroup instead of once per wave.  Per tile
// and K block three MFMAs, then G = Re(Q U) and Gz = Re(Q V) for the
// lane's 4 columns with Q = X[hx] Y[hy] gathered from the phase tables.
// Tile sums in FP32, sums over tiles in FP64.
// ----------------------------------------------------------------------
// 16-byte units of one LDS tile: the tile's column records, then KB blocks
// of (hi, lo) x 64 lane records
template <int KB>
constexpr int kf_tile_units() { return kKfTileCols + KB * 2 * kWave; }

// Workgroups per CU the back-projection is compiled for (its waves per
// SIMD, one wave of each workgroup per SIMD): up to four K blocks the
// registers fit four (left to itself the compiler used 132 VGPRs for three
// K blocks, three waves; capped, 106 without spills, and the 256-box step
// is 3% faster); wider boxes, whose W operands take 16 VGPRs per K block,
// keep their registers.
template <int KB>
constexpr int kf_min_waves() { return KB <= 4 ? 16 / kKfMfmaWaves : 1; }

template <int KB>
__global__ __launch_bounds__(kKfMfmaWaves* kWave, kf_min_waves<KB>()) void k_kforce_mfma(
    const BoxDesc* __restrict__ boxes, const Item* __restrict__ items) {
  constexpr int U = kf_tile_units<KB>();
  constexpr int NT = kKfMfmaWaves * kWave;
  constexpr int P = (U + NT - 1) / NT;  // staging loads per thread and tile
  __shared__ u4_t lds[2][U];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, kbk = B.kbk;
  const int tid = threadIdx.x;
  const int w = tid / kWave, l = tid % kWave, n = l & 31, g = l >> 5;
  const bool act = it.a + 32 * w < N;  // the wave has ions (uniform)
  const int i = it.a + 32 * w + n;
  const int ii = i < N ? i : 0;
  const u4_t* wb = reinterpret_cast<const u4_t*>(B.wb);
  h8_t bh[KB], bl[KB];
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    u4_t x = {0u, 0u, 0u, 0u}, y = {0u, 0u, 0u, 0u};
    if (i < N && kb < kbk) {
      const u4_t* p = wb + ((size_t)(2 * kb + g) * N + i) * 2;
      x = p[0];
      y = p[1];
    }
    bh[kb] = __builtin_bit_cast(h8_t, x);
    bl[kb] = __builtin_bit_cast(h8_t, y);
  }
  // Global-address-space views: uniform bases, 32-bit per-lane byte offsets
  const gbytes kfa = (gbytes)B.kfa;
  const gbytes X = (gbytes)B.tabx;
  const gbytes Y = (gbytes)B.taby;
  const cptr<KfTile> tiles = as_const(B.kft);
  const unsigned n8 = 8u * (unsigned)N, i8 = 8u * (unsigned)ii;
  const unsigned ut = 16u * (unsigned)kf_tile_stride(kbk);  // bytes per tile
  // Staging of tile t: its first na units (the column records, then K
  // blocks 0 .. kb_hi - 1) in one contiguous run.  Every thread issues
  // exactly P loads (units past the run reload its last unit and are not
  // stored), so the wait counts stay static.
  auto fetch = [&](int t, u4_t(&pf)[P]) {
    const int na = kKfTileCols + tiles[t].kb_hi * 2 * kWave;
#pragma unroll
    for (int m = 0; m < P; ++m)
      pf[m] = gload<u4_t>(kfa, (unsigned)t * ut + 16u * (unsigned)min(tid + NT * m, na - 1));
  };
  auto store = [&](int t, int buf, const u4_t(&pf)[P]) {
    const int na = kKfTileCols + tiles[t].kb_hi * 2 * kWave;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      const int u = tid + NT * m;
      if (u < na) lds[buf][u] = pf[m];
    }
  };
  double F0 = 0.0, F1 = 0.0, F2 = 0.0;
  const int tend = it.b + it.c;
  {
    u4_t pf[P];
    fetch(it.b, pf);
    store(it.b, 0, pf);
  }
  __syncthreads();
  for (int t = it.b; t < tend; ++t) {
    const int buf = (t - it.b) & 1;
    const bool more = t + 1 < tend;
    const u4_t* sa = lds[buf];
    f4v_t cw[4];
    f2v_t xv[4], yv[4];
    if (act) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        cw[j] = __builtin_bit_cast(f4v_t, sa[4 * g + j]);
        const unsigned bits = __float_as_uint(cw[j].w);
        // 24-bit products (8 N < 2^24): full-rate v_mad_u32_u24, not a
        // quarter-rate 64-bit multiply-add
        xv[j] = gload<f2v_t>(X, __umul24(bits & 0xffu, n8) + i8);
        yv[j] = gload<f2v_t>(Y, __umul24((bits >> 8) & 0xffu, n8) + i8);
      }
    }
    // the next tile's copy goes out after this tile's gathers, and always
    // (the last tile re-reads itself) so that the wait counts stay static
    __asm__ volatile("" ::: "memory");
    u4_t pf[P];
    fetch(more ? t + 1 : t, pf);
    if (act) {
      const int klo = tiles[t].kb_lo, khi = tiles[t].kb_hi;
      f16_t acc;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[v] = 0.f;
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        if (kb >= klo && kb < khi) {
          const h8_t ah = __builtin_bit_cast(h8_t, sa[kKfTileCols + kb * 2 * kWave + l]);
          const h8_t al = __builtin_bit_cast(h8_t, sa[kKfTileCols + kb * 2 * kWave + kWave + l]);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh[kb], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[kb], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[kb], acc, 0, 0, 0);
        }
      }
      float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // X[-h] = conj X[h]: the sign of hx (hy) is that of cw.x (cw.y)
        const float xr = xv[j].x, xi = flip_sign(xv[j].y, cw[j].x);
        const float yr = yv[j].x, yi = flip_sign(yv[j].y, cw[j].y);
        const float qr = xr * yr - xi * yi, qi = xr * yi + xi * yr;
        const float G = qr * acc[4 * j] - qi * acc[4 * j + 1];
        const float Gz = qr * acc[4 * j + 2] - qi * acc[4 * j + 3];
        fx += cw[j].x * G;
        fy += cw[j].y * G;
        fz += cw[j].z * Gz;
      }
      F0 += (double)fx;
      F1 += (double)fy;
      F2 += (double)fz;
    }
    if (more) store(t + 1, buf ^ 1, pf);
    __syncthreads();
  }
  // the two lanes of an ion hold the two halves of every tile's columns
  const double fx = F0 + __shfl_xor(F0, 32);
  const double fy = F1 + __shfl_xor(F1, 32);
  const double fz = F2 + __shfl_xor(F2, 32);
  if (act && g == 0 && i < N) {
    const double zq = B.Z[B.typ[i]];
    double* fp = B.fpart + (size_t)it.d * 3 * N;
    fp[i] = -zq * fx * B.invL[0];
    fp[N + i] = -zq * fy * B.invL[1];
    fp[2 * N + i] = -zq * fz * B.invL[2];
  }
}

}  // namespace mdg

namespace mdg {

// ======================================================================
// Structure factor on the matrix cores (mixed path).
//   S(k) = sum_i Z_i X_i[hx] Y_i[hy] W_i[hz]  (mdc/program/rcprcl.c:59-78)
// For 32 columns (hx, hy) and 16 hz this is one real matrix product over
// the ions:  C[column][(hz, Re|Im)] = sum_(ion, re|im) A B  with
//   A[column][(ion, p)] = Q = X[hx] Y[hy]  (p: Re Q, Im Q),
//   B[(ion, p)][(hz, out)] from Z W[hz]: Re S takes (ZWr, -ZWi), Im S
//   takes (ZWi, ZWr).
// Q is formed per lane (its column, 4 ions per K step) and split into FP16
// hi/lo; the B records come from k_mixed_tables.  Three MFMAs per K step and hz
// tile (hi*hi + hi*lo + lo*hi), FP32 accumulation over 64 ions, then FP64.
// ======================================================================

// One workgroup: 4 waves x 32 columns, one ion chunk, two hz tiles of 16
// (further hz tiles are further items).  Ion tiles of 32 are staged in LDS
// for the four waves: X and Y rows of every h (rows padded by two complex
// so that lanes reading different rows use different banks) and the
// tile's Z W records of the item's hz tiles.
constexpr int kSkmTile = 32;
constexpr int kSkmRow = kSkmTile + 2;  // complex per LDS row
constexpr int kSkmThreads = kSkmWaves * kWave;
// X/Y staging loads per thread in flight at once
constexpr int kSkmXYBatch = 4;
// X/Y elements per thread fetched one tile ahead (all of them for hmax < 24)
constexpr int kSkmPre = 6;

// LDS bytes of k_sk_mfma for a launch whose largest box has H = hmax + 1
// (X and Y rows h = 0 .. H - 1 plus a zero row each, and the Z W records)
__host__ __device__ inline int sk_mfma_lds_bytes(int H) {
  return 2 * (H + 1) * kSkmRow * 8 + 8 * 32 * 4 * 16;
}

// (four workgroups per CU: the compiler keeps the kernel in 128 registers, 116
// without spills, for four waves per SIMD instead of three: pass A alone
// 0.120 -> 0.109 ms on the 256-box batch)
__global__ __launch_bounds__(kSkmThreads, 4) void k_sk_mfma(const BoxDesc* __restrict__ boxes,
                                                         const Item* __restrict__ items,
                                                         int hl) {
  extern __shared__ u4_t smem4[];
  const Item it = items[blockIdx.x];
  const BoxDesc& B = boxes[it.box];
  const int N = B.N, ncol = B.ncol, HP = 16 * B.sknt, H = B.hmax + 1;
  const int tid = threadIdx.x, w = tid / kWave, l = tid % kWave, r = l & 31, g = l >> 5;
  const int t0 = it.c;                 // first hz tile of the item
  const int nt = min(2, B.sknt - t0);  // its hz tiles with k vectors (1 or 2)
  constexpr int kRow4 = kSkmRow / 2;   // LDS row in 16-byte units
  typedef __attribute__((address_space(3))) u4_t lds_u4;  // LDS stores, not flat
  lds_u4* xs = (lds_u4*)smem4;         // [hl + 1][kRow4]
  lds_u4* ys = xs + (hl + 1) * kRow4;  // [hl + 1][kRow4]
  lds_u4* ss = ys + (hl + 1) * kRow4;  // [8 groups][32 hz][4]
  typedef __attribute__((address_space(3))) f2v_t lds_f2w;
  lds_f2w* xsf = (lds_f2w*)xs;         // the same rows as complex elements
  lds_f2w* ysf = (lds_f2w*)ys;
  const int exf = H * kSkmTile;        // complex elements per table and tile
  constexpr int es = 8 * 32 * 4;
  const gbytes gxb = (gbytes)B.tabx;
  const gbytes gyb = (gbytes)B.taby;
  const gbytes gsb = (gbytes)B.sb;
  const int jb = it.a * B.chunk, je = min(N, jb + B.chunk);
  const int ngrp = B.npad / 4;
  // Row H of X and Y stays zero (staging writes rows < H): lanes without a
  // column read it and contribute nothing.
  if (tid < 2 * kRow4) (tid < kRow4 ? xs : ys)[H * kRow4 + tid % kRow4] = u4_t{0u, 0u, 0u, 0u};
  // this lane's column
  const int c0 = it.b + 32 * w;
  const int c = c0 + r;
  const bool cv = c < ncol;
  int ax = H, ay = H;
  float sx = 1.f, sy = 1.f;
  if (cv) {
    const int hx = B.cols[c].hx, hy = B.cols[c].hy;
    ax = hx < 0 ? -hx : hx;
    ay = hy < 0 ? -hy : hy;
    sx = hx < 0 ? -1.f : 1.f;
    sy = hy < 0 ? -1.f : 1.f;
  }
  typedef float f2_t __attribute__((ext_vector_type(2)));
  typedef __attribute__((address_space(3))) const f2_t lds_f2;
  lds_f2* xr = (lds_f2*)(xs + ax * kRow4);
  lds_f2* yr = (lds_f2*)(ys + ay * kRow4);
  // FP32 sums over the chunk (at most kSkmChunk ions); chunks add in FP64
  f16_t acc[2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
  // Staging of one ion tile: the records of its 8 ion groups for both hz
  // tiles (zero past the box's hz), and X and Y of one ion (h < H) per
  // element.  The records and the first kSkmPre elements of each thread
  // are fetched into registers one tile ahead, so their loads are in
  // flight while the MFMAs of the current tile run; elements beyond that
  // (large hmax) are loaded at the tile itself, kSkmXYBatch at a time.
  u4_t vss[es / kSkmThreads];
  f2v_t vpre[kSkmPre];
  // Each thread's staging addresses are the same for every tile but for a
  // uniform tile term, so they are formed once here: the record offsets
  // (and the last record group they may read), the prefetched X/Y
  // elements' offsets from the X table (the Y table lies a fixed distance
  // after it) and their LDS slots.
  constexpr int kRecs = es / kSkmThreads;
  unsigned rec_off[kRecs];
  int rec_grp[kRecs];
#pragma unroll
  for (int m = 0; m < kRecs; ++m) {
    const int q = tid + m * kSkmThreads;
    const int grp = q >> 7, rem = q & 127, hz = 16 * t0 + (rem >> 2);
    rec_off[m] = 16u * (((unsigned)grp * (unsigned)HP + 16u * (unsigned)t0) * 4u + (unsigned)rem);
    rec_grp[m] = hz < HP ? grp : 0x3fffffff;  // past the box's hz: never read
  }
  const unsigned xy_dist = (unsigned)((const char*)B.taby - (const char*)B.tabx);
  const int col = tid % kSkmTile;  // every element of this thread is ion column col
  unsigned xy_off[kSkmPre];
  int xy_lds[kSkmPre];  // index into the LDS X rows (the Y rows follow them); -1: none
#pragma unroll
  for (int m = 0; m < kSkmPre; ++m) {
    const int e = tid + m * kSkmThreads, tb = e >= exf, q = e - tb * exf;
    const int row = q / kSkmTile;
    xy_off[m] = (tb ? xy_dist : 0u) + 8u * ((unsigned)row * (unsigned)N + (unsigned)col);
    xy_lds[m] = e < 2 * exf ? (tb ? (hl + 1) * kSkmRow : 0) + row * kSkmRow + col : -1;
  }
  auto xy_elem = [&](int j0, int e) {
    const int tb = e >= exf, q = e - tb * exf, i = j0 + q % kSkmTile;
    f2v_t v = f2v_t{0.f, 0.f};
    if (e < 2 * exf && i < N)  // 24-bit row product (rows < 2^8, N < 2^24): full rate
      v = gload<f2v_t>(tb ? gyb : gxb,
                       8u * (__umul24((unsigned)(q / kSkmTile), (unsigned)N) + (unsigned)i));
    return v;
  };
  auto xy_store = [&](int e, f2v_t v) {
    const int tb = e >= exf, q = e - tb * exf;
    if (e < 2 * exf)
      (tb ? ysf : xsf)[__umul24((unsigned)(q / kSkmTile), (unsigned)kSkmRow) + q % kSkmTile] = v;
  };
  auto fetch = [&](int j0) {
    const int g0 = j0 >> 2;                                // the tile's first record group
    const unsigned rec_t = 64u * (unsigned)HP * (unsigned)g0;  // uniform
#pragma unroll
    for (int m = 0; m < kRecs; ++m) {
      vss[m] = u4_t{0u, 0u, 0u, 0u};
      if (g0 + rec_grp[m] < ngrp) vss[m] = gload<u4_t>(gsb, rec_off[m] + rec_t);
    }
    const bool ion_ok = j0 + col < N;
#pragma unroll
    for (int m = 0; m < kSkmPre; ++m) {
      vpre[m] = f2v_t{0.f, 0.f};
      if (ion_ok && xy_lds[m] >= 0) vpre[m] = gload<f2v_t>(gxb, xy_off[m] + 8u * (unsigned)j0);
    }
  };
  // the tile loop for NT hz tiles per item (a box with hmax < 16 has one:
  // no MFMAs on its all-zero second tile)
  auto tiles = [&](auto ntc) {
    constexpr int NT = decltype(ntc)::value;
    if (jb < je) fetch(jb);
    for (int j0 = jb; j0 < je; j0 += kSkmTile) {
      __syncthreads();  // the previous tile's reads are done
#pragma unroll
      for (int m = 0; m < es / kSkmThreads; ++m) ss[tid + m * kSkmThreads] = vss[m];
#pragma unroll
      for (int m = 0; m < kSkmPre; ++m)
        if (xy_lds[m] >= 0) xsf[xy_lds[m]] = vpre[m];
      for (int e0 = kSkmPre * kSkmThreads; e0 < 2 * exf; e0 += kSkmXYBatch * kSkmThreads) {
        f2v_t v[kSkmXYBatch];
#pragma unroll
        for (int m = 0; m < kSkmXYBatch; ++m) v[m] = xy_elem(j0, e0 + tid + m * kSkmThreads);
#pragma unroll
        for (int m = 0; m < kSkmXYBatch; ++m) xy_store(e0 + tid + m * kSkmThreads, v[m]);
      }
      __syncthreads();
      if (j0 + kSkmTile < je) fetch(j0 + kSkmTile);
#pragma unroll
      for (int s = 0; s < kSkmTile / 8; ++s) {
        const int u = 8 * s + 4 * g;  // tile-local first ion of this lane's K slice
        h2_t hi[4], lo[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const f2_t x = xr[u + e], y = yr[u + e];
          const float xi = sx * x.y, yi = sy * y.y;
          split_f16x2(f2v_t{x.x * y.x - xi * yi, x.x * yi + xi * y.x}, hi[e], lo[e]);
        }
        const h8_t ah = __builtin_shufflevector(
            __builtin_shufflevector(hi[0], hi[1], 0, 1, 2, 3),
            __builtin_shufflevector(hi[2], hi[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
        const h8_t al = __builtin_shufflevector(
            __builtin_shufflevector(lo[0], lo[1], 0, 1, 2, 3),
            __builtin_shufflevector(lo[2], lo[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
        const int grp = 2 * s + g;  // tile-local ion group of this K slice
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const lds_u4* rec = ss + (grp * 32 + 16 * t + (r >> 1)) * 4 + 2 * (r & 1);
          const h8_t bh = __builtin_bit_cast(h8_t, rec[0]);
          const h8_t bl = __builtin_bit_cast(h8_t, rec[1]);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, acc[t], 0, 0, 0);
        }
      }
    }
  };
  if (nt == 2)
    tiles(std::integral_constant<int, 2>{});
  else
    tiles(std::integral_constant<int, 1>{});
  if (c0 >= ncol) return;
  // result (row = column 8 (v/4) + 4 g + v%4, col = (hz r/2, out r%2)):
  // lanes r and r^1 hold Re S and Im S of the same (column, hz)
  double2* dst = B.skpart + (size_t)it.a * B.nk;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (t >= nt) continue;
    const int hz = 16 * (t0 + t) + (r >> 1);
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const double mine = (double)acc[t][v];
      const double other = __shfl_xor(mine, 1);
      const int cc = c0 + 8 * (v >> 2) + 4 * g + (v & 3);
      if ((r & 1) == 0 && cc < ncol) {
        const KCol col = B.cols[cc];
        if (hz >= col.hz0 && hz < col.hz0 + col.len)
          dst[col.k0 + hz - col.hz0] = make_double2(mine, other);
      }
    }
  }
}

}  // namespace mdg
