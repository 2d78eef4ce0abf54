// pr_dropin.hip -- libmdgpu_pr.so: pr-mdc's force routines (real_space,
// ewald1, ewald2, ewald3, calc_alpha) on gfx950 over pr-mdc's globals
// (include/prmdc_abi.h).
//
// pr-mdc (the reference's Parrinello-Rahman constant-pressure program) keeps
// a triclinic MD cell: positions are Cartesian, the cell matrix cell.hm
// holds the edge vectors a, b, c as rows, and every pair is folded into the
// cell by the minimum image in scaled coordinates (cell.c:305-329).  Its
// cutoff is half the smallest distance between opposite faces (cell.c:27-51),
// so every pair of the box is a candidate: the real-space kernels run over
// all pairs.
//
// Kernels (FP64 throughout; this unit is compiled with -ffp-contract=off so
// that every expression pr-mdc evaluates rounds the same way here):
//   k_pr_pairs<kPrTable>    real.c:11-77     BMH table pairs + virial tensor
//   k_pr_pairs<kPrEwald1>   ewald.c:11-68    Mori erfc pairs + virial tensor
//   k_pr_sk                 ewald.c:96-143   S(h) per reciprocal vector
//   k_pr_kforce             ewald.c:145-154  force back-projection
//   k_pr_ksum, k_pr_self    ewald.c:141-157, 162-173   energies
//   k_pr_fsum, k_pr_ssum    fixed-order sums of the per-split partials
//
// Pair kernels: one wave per (home chunk of 64 ions, split of the partner
// range).  Each lane gathers its own ion's force (no atomics); partners are
// staged 16 at a time in LDS, and the partner range splits at that grain so
// that a few thousand ions still give several waves per SIMD.  Every pair is evaluated in pr-mdc's
// orientation, d = r_lo - r_hi with lo < hi the ion indices, so both ions of
// a pair use the same image even where a scaled component sits exactly on
// the +-0.5 fold boundary; the higher ion subtracts as pr-mdc's j-side does.
// Energy and virial count each pair once, on the lower ion.  With one split
// a lane adds its partners in ascending index onto the force it was given,
// which is the order pr-mdc's i<j loop adds them: the table routine's forces
// are then bit-identical to pr-mdc's.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "mdgpu.h"
#include "prmdc_abi.h"

// The globals of pr-mdc/program/md.h:224-229 that the routines use.  A
// pr-mdc executable defines its own (md.h defines them in the header); ELF
// interposition then binds these references to the executable's objects.
calc_control ctl;
cell_parameters cell;
system_property sys;
pressure_temperature_set pt;
ion_property ion;
Parrinello_Rahman_set pr;

namespace {

// pr-mdc/program/md.h:7-15, as pr-mdc writes them
constexpr double kPI = 3.1415926535897932385;
constexpr double kPI2 = 6.283185307179587;
constexpr double kPIPI = 9.869604401089358619;
constexpr double kSQRTPI = 1.77245385090551602731;

constexpr int kW = 64;            // wavefront; pair and k-force blocks are one wave
constexpr int kPT = 16;           // partners staged per tile (the unit of a pair split)
constexpr int kSkThreads = 256;   // k_pr_sk block
constexpr int kPairWaves = 8192;  // pair launches: aim for this many waves
constexpr int kKWaves = 2048;     // k-force launches
constexpr int kMinKPerSplit = 8;  // reciprocal vectors per k-force split, at least
constexpr int kScal = 8;          // pair-kernel scalars: pot, 6 virial entries, bad pairs

enum { kPrTable = 0, kPrEwald1 = 1, kPrBoth = 2 };  // kPrBoth: the resident loop's fused pass

struct PrParams {
  int N, kinds, td, S;
};

struct PrK {
  int N, K, S;
};

// The cell and the scalars of a pr-mdc system, resident on the device: what
// the force kernels read (the drop-in routines upload it from pr-mdc's
// globals per call) and what mdg_pr_run's step loop updates in place.
// Matrices are pr-mdc's m[i][j] at 3i+j.
struct PrCell {
  double H[9], R[9], sig[9], G[9];         // cell.hm, hmr, sigma, g
  double vol, cutoff, cutoff2, a1, a1_sqrtpi, v2, vp;
  double Lx, Ly, Lz, ang_ab, ang_bc, ang_ca;
  double prv[9], prf[9], prf0[9], prmass, prdt;  // Parrinello-Rahman cell velocity, forces
  double HmF[9];                           // cell_force's (H^T)^-1 (dH/dt)^T
  double pset[9], ptens[9], vir[9], kint[9];
  double pot, kin, ptrace;
  double dt, dt2, e2t, temp_set, oldk, vscale;
  int step, count, t_step, skip_ewald, bad_cell, rec_base, rec_cap, pad;
};

// what the pair kernels read of the cell
struct Geo {
  double H[9], R[9], cutoff2, a1, a1_sqrtpi;
};

__device__ __forceinline__ Geo load_geo(const PrCell* __restrict__ c) {
  Geo q;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    q.H[k] = c->H[k];
    q.R[k] = c->R[k];
  }
  q.cutoff2 = c->cutoff2;
  q.a1 = c->a1;
  q.a1_sqrtpi = c->a1_sqrtpi;
  return q;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kW);
  return v;
}

// Mori's erfc as pr-mdc's erfcc evaluates it (pr-mdc/program/ext.c:73-104,
// the same constants as mdc's):
//   erfc(x) ~ (x/pi) e^{-x^2} S(x^2) - 2/(e^{4 pi x} - 1),
//   S(y) = sum_{n=0..13} w_n/(c_n + y), c_n = n^2/4, w_0 = 1/2, w_n = e^{-n^2/4}.
// pr-mdc divides once per term; here S is carried as one fraction
// num/den (num <- num (c_n + y) + w_n den, den <- den (c_n + y)), which is
// the same sum with one division instead of fourteen.  The result differs
// from the term-by-term sum by rounding only (a few 1e-16 relative); FP64
// division is a ten-instruction sequence on gfx950, and this is the bulk of
// the per-pair work of ewald1.
__device__ __forceinline__ double pr_erfc(double x) {
  const double w[13] = {
      0.7788007830714049,    0.3678794411714423,    0.1053992245618643,
      0.01831563888873417,   0.001930454136227709,  0.0001234098040866795,
      4.785117392129008e-6,  1.125351747192591e-7,  1.605228055185611e-9,
      1.388794386496401e-11, 7.287724095819693e-14, 2.319522830243569e-16,
      4.4777324417183e-19};
  const double y = x * x;
  double num = 0.5, den = y;
#pragma unroll
  for (int n = 1; n <= 13; ++n) {
    const double d = 0.25 * (double)(n * n) + y;
    num = num * d + w[n - 1] * den;
    den = den * d;
  }
  const double v = -2.0 / (-1.0 + exp(12.56637061435917 * x)) +
                   (0.3183098861837907 * x * (num / den)) / exp(y);
  return x >= 0.0 ? v : 2.0 - v;
}

// pr-mdc's cyclic_bc_dr (cell.c:305-329): to scaled coordinates with
// Hm^-1, fold each component into [-0.5, 0.5) ('>=' then '<'), back with Hm.
__device__ __forceinline__ void min_image(const Geo& p, double& dx, double& dy, double& dz) {
  double sdx = p.R[0] * dx + p.R[3] * dy + p.R[6] * dz;
  double sdy = p.R[1] * dx + p.R[4] * dy + p.R[7] * dz;
  double sdz = p.R[2] * dx + p.R[5] * dy + p.R[8] * dz;
  if (sdx >= 0.5) sdx -= 1.0;
  if (sdy >= 0.5) sdy -= 1.0;
  if (sdz >= 0.5) sdz -= 1.0;
  if (sdx < -0.5) sdx += 1.0;
  if (sdy < -0.5) sdy += 1.0;
  if (sdz < -0.5) sdz += 1.0;
  dx = p.H[0] * sdx + p.H[3] * sdy + p.H[6] * sdz;
  dy = p.H[1] * sdx + p.H[4] * sdy + p.H[7] * sdz;
  dz = p.H[2] * sdx + p.H[5] * sdy + p.H[8] * sdz;
}

// grid (ceil(N/64), S); r and f are SoA [x N][y N][z N].
// S == 1: f[i] += the pair forces, in place.  S > 1: the split's partial
// force goes to pf[(split*3 + c)*N + i] and k_pr_fsum adds them.
// ps[(split*gridDim.x + chunk)*kScal + c]: the wave's pot, virial, bad pairs.
template <int MODE>
__global__ __launch_bounds__(kW) void k_pr_pairs(PrParams p, const PrCell* __restrict__ cell_,
                                                 const double* __restrict__ r,
                                                 const int* __restrict__ typ,
                                                 const double2* __restrict__ tab,
                                                 const int* __restrict__ lookup,
                                                 const double* __restrict__ Z,
                                                 double* __restrict__ f, double* __restrict__ pf,
                                                 double* __restrict__ ps) {
  __shared__ double sx[kPT], sy[kPT], sz[kPT];
  __shared__ int st[kPT];
  const int N = p.N, lane = threadIdx.x;
  const int i = blockIdx.x * kW + lane;
  const bool valid = i < N;
  const Geo q = load_geo(cell_);
  const int ntile = (N + kPT - 1) / kPT;
  const int t0 = (int)((long long)ntile * blockIdx.y / p.S);
  const int t1 = (int)((long long)ntile * (blockIdx.y + 1) / p.S);
  double xi = 0, yi = 0, zi = 0;
  int ti = 0;
  if (valid) {
    xi = r[i];
    yi = r[N + i];
    zi = r[2 * N + i];
    ti = typ[i];
  }
  double fx = 0, fy = 0, fz = 0;
  if (p.S == 1 && valid) {
    fx = f[i];
    fy = f[N + i];
    fz = f[2 * N + i];
  }
  const double Zi = Z[ti];
  double pot = 0, v00 = 0, v01 = 0, v02 = 0, v11 = 0, v12 = 0, v22 = 0, bad = 0;
  for (int t = t0; t < t1; ++t) {
    const int j0 = t * kPT, jl = j0 + lane;
    __syncthreads();  // the previous tile's reads are done
    if (lane < kPT && jl < N) {
      sx[lane] = r[jl];
      sy[lane] = r[N + jl];
      sz[lane] = r[2 * N + jl];
      st[lane] = typ[jl];
    }
    __syncthreads();
    if (!valid) continue;  // every lane still reaches both barriers of each tile
    const int nj = min(kPT, N - j0);
    for (int k = 0; k < nj; ++k) {
      const int j = j0 + k;
      if (j == i) continue;
      const bool lo = i < j;
      double dx, dy, dz;
      if (lo) {
        dx = xi - sx[k];
        dy = yi - sy[k];
        dz = zi - sz[k];
      } else {
        dx = sx[k] - xi;
        dy = sy[k] - yi;
        dz = sz[k] - zi;
      }
      min_image(q, dx, dy, dz);
      double dr = dx * dx + dy * dy + dz * dz;
      if (dr > q.cutoff2) continue;
      double fij = 0.0, pe = 0.0;
      const double rr = sqrt(dr);
      if constexpr (MODE != kPrEwald1) {
        // real.c:43-56: the 0.001 A bin of r from 0.5 A
        const int tl = lo ? ti : st[k], th = lo ? st[k] : ti;
        const int ddr = (int)((rr - 0.5) * 1000.0);
        if (ddr < 0 || ddr > p.td) {  // inside the table's first bin: pr-mdc reads out of bounds
          bad += 1.0;
          continue;
        }
        const double2 e = tab[(size_t)lookup[tl * p.kinds + th] * (p.td + 1) + ddr];
        fij = e.y;
        pe = e.x;
      }
      if constexpr (MODE != kPrTable) {
        // ewald.c:32-43
        const double ZZ = lo ? Zi * Z[st[k]] : Z[st[k]] * Zi;
        const double adr = q.a1 * rr;
        const double e = pr_erfc(adr) / rr;
        const double fe = ZZ * (2.0 * q.a1_sqrtpi * exp(-adr * adr) + e) / (rr * rr);
        if constexpr (MODE == kPrBoth) {
          pe += ZZ * e;
          fij += fe;
        } else {
          pe = ZZ * e;
          fij = fe;
        }
      }
      if (lo) pot += pe;
      if (lo) {
        fx += fij * dx;
        fy += fij * dy;
        fz += fij * dz;
        v00 += fij * dx * dx;
        v01 += fij * dx * dy;
        v02 += fij * dx * dz;
        v11 += fij * dy * dy;
        v12 += fij * dy * dz;
        v22 += fij * dz * dz;
      } else {
        fx -= fij * dx;
        fy -= fij * dy;
        fz -= fij * dz;
      }
    }
  }
  if (valid) {
    double* o = p.S == 1 ? f : pf + (size_t)blockIdx.y * 3 * N;
    o[i] = fx;
    o[N + i] = fy;
    o[2 * N + i] = fz;
  }
  double v[kScal] = {pot, v00, v01, v02, v11, v12, v22, bad};
#pragma unroll
  for (int c = 0; c < kScal; ++c) v[c] = wsum(v[c]);
  if (lane == 0) {
    double* o = ps + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * kScal;
#pragma unroll
    for (int c = 0; c < kScal; ++c) o[c] = v[c];
  }
}

// sum_{k < n} p[k*stride] with eight independent accumulators (k mod 8)
// combined in a fixed tree: the loads overlap instead of waiting on one add
// chain, and the order is fixed for a given n.
__device__ __forceinline__ double sum_strided(const double* __restrict__ p, int n, size_t stride) {
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int k = 0;
  for (; k + 8 <= n; k += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += p[(size_t)(k + u) * stride];
  }
  for (int u = 0; k < n; ++k, ++u) a[u] += p[(size_t)k * stride];
  return ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

// f[c*N + i] += the sum over splits s of pf[(s*3 + c)*N + i]
__global__ __launch_bounds__(256) void k_pr_fsum(int n3, int S, const double* __restrict__ pf,
                                                 double* __restrict__ f) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= n3) return;
  f[idx] = f[idx] + sum_strided(pf + idx, S, (size_t)n3);
}

// out[c] = sum over rows w = 0..nw-1 of ps[w*kScal + c]: 64 groups of
// threads each sum every 64th row, then the 64 group sums are added in
// group order (a fixed order for a given nw).
constexpr int kSsumThreads = 512;
constexpr int kSsumGroups = kSsumThreads / kScal;
__global__ __launch_bounds__(kSsumThreads) void k_pr_ssum(int nw, const double* __restrict__ ps,
                                                          double* __restrict__ out) {
  __shared__ double part[kSsumGroups][kScal];
  const int c = threadIdx.x % kScal, grp = threadIdx.x / kScal;
  const int rows = nw > grp ? (nw - grp + kSsumGroups - 1) / kSsumGroups : 0;
  part[grp][c] = sum_strided(ps + (size_t)grp * kScal + c, rows, (size_t)kSsumGroups * kScal);
  __syncthreads();
  if (threadIdx.x < kScal) {
    double t = 0.0;
    for (int k = 0; k < kSsumGroups; ++k) t += part[k][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

// One block per reciprocal vector h (ewald.c:96-143).  kc[k*8 + .]:
// {2C, 2S, v2 e/|h|^2, hhx, hhy, hhz, vp e/|h|^2 ((2C)^2 + (2S)^2), 0}
// with C, S = sum_j Z_j (cos, sin)(2 pi hh . r_j).
__global__ __launch_bounds__(kSkThreads) void k_pr_sk(PrK p, const PrCell* __restrict__ cl,
                                                      const int4* __restrict__ hv,
                                                      const double* __restrict__ r,
                                                      const int* __restrict__ typ,
                                                      const double* __restrict__ Z,
                                                      double* __restrict__ kc) {
  __shared__ double red[2][kSkThreads / kW];
  const int k = blockIdx.x, N = p.N, tid = threadIdx.x;
  const int4 h = hv[k];
  const double hx = (double)h.x, hy = (double)h.y, hz = (double)h.z;
  const double* G = cl->G;
  const double hhx = hx * G[0] + hy * G[3] + hz * G[6];
  const double hhy = hx * G[1] + hy * G[4] + hz * G[7];
  const double hhz = hx * G[2] + hy * G[5] + hz * G[8];
  const double hhx1 = kPI2 * hhx, hhy1 = kPI2 * hhy, hhz1 = kPI2 * hhz;
  double c = 0.0, s = 0.0;
  for (int j = tid; j < N; j += kSkThreads) {
    const double ti = hhx1 * r[j] + hhy1 * r[N + j] + hhz1 * r[2 * N + j];
    const double zj = Z[typ[j]];
    double sn, cs;
    sincos(ti, &sn, &cs);
    c += zj * cs;
    s += zj * sn;
  }
  c = wsum(c);
  s = wsum(s);
  if ((tid & (kW - 1)) == 0) {
    red[0][tid / kW] = c;
    red[1][tid / kW] = s;
  }
  __syncthreads();
  if (tid == 0) {
    c = 0.0;
    s = 0.0;
    for (int w = 0; w < kSkThreads / kW; ++w) {
      c += red[0][w];
      s += red[1][w];
    }
    const double hh2 = hhx * hhx + hhy * hhy + hhz * hhz;
    const double ex = exp(-hh2 * (kPIPI / (cl->a1 * cl->a1)));
    const double vpe = cl->vp * (ex / hh2), v2e = cl->v2 * (ex / hh2);
    const double C2 = c * 2.0, S2 = s * 2.0;
    double* o = kc + (size_t)k * 8;
    o[0] = C2;
    o[1] = S2;
    o[2] = v2e;
    o[3] = hhx;
    o[4] = hhy;
    o[5] = hhz;
    o[6] = vpe * (C2 * C2 + S2 * S2);
    o[7] = 0.0;
  }
}

// grid (ceil(N/64), S): lane i adds the back-projection of its split's
// reciprocal vectors (ewald.c:145-154); S == 1 in place, else partials.
__global__ __launch_bounds__(kW) void k_pr_kforce(PrK p, const double* __restrict__ r,
                                                  const int* __restrict__ typ,
                                                  const double* __restrict__ Z,
                                                  const double* __restrict__ kc,
                                                  double* __restrict__ f, double* __restrict__ pf) {
  const int N = p.N, i = blockIdx.x * kW + threadIdx.x;
  if (i >= N) return;  // no barriers below
  const int k0 = (int)((long long)p.K * blockIdx.y / p.S);
  const int k1 = (int)((long long)p.K * (blockIdx.y + 1) / p.S);
  const double x = r[i], y = r[N + i], z = r[2 * N + i], zi = Z[typ[i]];
  double fx = 0, fy = 0, fz = 0;
  if (p.S == 1) {
    fx = f[i];
    fy = f[N + i];
    fz = f[2 * N + i];
  }
  for (int k = k0; k < k1; ++k) {
    const double* q = kc + (size_t)k * 8;
    const double hhx = q[3], hhy = q[4], hhz = q[5];
    const double ti = (kPI2 * hhx) * x + (kPI2 * hhy) * y + (kPI2 * hhz) * z;
    double sn, cs;
    sincos(ti, &sn, &cs);
    const double ci = zi * cs, si = zi * sn;
    const double foc = -q[2] * (ci * q[1] - si * q[0]);
    fx += foc * hhx;
    fy += foc * hhy;
    fz += foc * hhz;
  }
  double* o = p.S == 1 ? f : pf + (size_t)blockIdx.y * 3 * N;
  o[i] = fx;
  o[N + i] = fy;
  o[2 * N + i] = fz;
}

// out[0] = sum over k of the vectors' energies (ewald.c:141-157): thread t
// sums every 256th vector from t, then the 256 sums are added in order.
__global__ __launch_bounds__(256) void k_pr_ksum(int K, const double* __restrict__ kc,
                                                 double* __restrict__ out) {
  __shared__ double part[256];
  const int t = threadIdx.x;
  const int n = K > t ? (K - t + 255) / 256 : 0;
  part[t] = sum_strided(kc + (size_t)t * 8 + 6, n, (size_t)256 * 8);
  __syncthreads();
  if (t == 0) {
    double a = 0.0;
    for (int k = 0; k < 256; ++k) a += part[k];
    out[0] = a;
  }
}

// out[0] = -sum_i Z_i^2 * a1/sqrt(pi) (ewald.c:162-173)
__global__ __launch_bounds__(256) void k_pr_self(int N, const PrCell* __restrict__ c,
                                                 const int* __restrict__ typ,
                                                 const double* __restrict__ Z,
                                                 double* __restrict__ out) {
  __shared__ double red[256 / kW];
  double a = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) {
    const double z = Z[typ[i]];
    a += z * z;
  }
  a = wsum(a);
  if ((threadIdx.x & (kW - 1)) == 0) red[threadIdx.x / kW] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double z2 = 0.0;
    for (int w = 0; w < 256 / kW; ++w) z2 += red[w];
    out[0] = -z2 * c->a1_sqrtpi;
  }
}

// ---------------------------------------------------------------- step loop
// pr-mdc's step (main.c:100-111 without output and MSD), resident:
// next_rv_verlet (rv.c:12-17: calc_foc, next_v, next_r), control_press_PR
// (pt.c:27-63), control_temp (pt.c:124-150), calc_kin (rv.c:68-105),
// calc_press (pt.c:8-24).  The cell matrices and scalars live in PrCell and
// are updated by single-thread kernels; per-ion work runs one lane per ion.

// out slots: [0, kScal) pair sums (the drop-in's pair routine, or the step
// loop's fused pass), kOutK reciprocal energy, kOutSelf self term
enum { kOutK = kScal, kOutSelf = kScal + 1, kOutLen = 16 };
constexpr int kRec = 13;  // per-step record: pot, kin, pressure, volume, H

// cofactor inverse as pr-mdc writes it (cell.c:228-240, 80-90):
// o = M^-1 with m[i][j] at 3i+j; returns det M
__device__ double cof_inverse(const double* M, double* o) {
  const double det = M[0] * (M[4] * M[8] - M[7] * M[5]) + M[3] * (M[7] * M[2] - M[1] * M[8]) +
                     M[6] * (M[1] * M[5] - M[4] * M[2]);
  o[0] = (-(M[5] * M[7]) + M[4] * M[8]) / det;
  o[3] = (M[5] * M[6] - M[3] * M[8]) / det;
  o[6] = (-(M[4] * M[6]) + M[3] * M[7]) / det;
  o[1] = (M[2] * M[7] - M[1] * M[8]) / det;
  o[4] = (-(M[2] * M[6]) + M[0] * M[8]) / det;
  o[7] = (M[1] * M[6] - M[0] * M[7]) / det;
  o[2] = (-(M[2] * M[4]) + M[1] * M[5]) / det;
  o[5] = (M[2] * M[3] - M[0] * M[5]) / det;
  o[8] = (-(M[1] * M[3]) + M[0] * M[4]) / det;
  return det;
}

// calc_foc's prologue (rv.c:26-33): calc_sigma (cell.c:148-175), set_cutoff
// (cell.c:27-51), and cell_force's (H^T)^-1 (dH/dt)^T (cell.c:58-95).
__device__ void cell_pre(PrCell* __restrict__ c) {
  c->step += 1;
  const double* H = c->H;
  double* sg = c->sig;
  sg[0] = (H[4] * H[8] - H[5] * H[7]);
  sg[1] = (H[5] * H[6] - H[3] * H[8]);
  sg[2] = (H[3] * H[7] - H[4] * H[6]);
  sg[3] = (H[7] * H[2] - H[8] * H[1]);
  sg[4] = (H[8] * H[0] - H[6] * H[2]);
  sg[5] = (H[6] * H[1] - H[7] * H[0]);
  sg[6] = (H[1] * H[5] - H[2] * H[4]);
  sg[7] = (H[2] * H[3] - H[0] * H[5]);
  sg[8] = (H[0] * H[4] - H[1] * H[3]);
  for (int k = 0; k < 9; ++k) c->G[k] = sg[k] / c->vol;
  const double* q = c->G;
  const double h1 = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  const double h2 = 1.0 / sqrt(q[3] * q[3] + q[4] * q[4] + q[5] * q[5]);
  const double h3 = 1.0 / sqrt(q[6] * q[6] + q[7] * q[7] + q[8] * q[8]);
  double mn = 0.0;
  if ((h1 <= h2) && (h1 <= h3)) mn = h1;
  if ((h2 <= h1) && (h2 <= h3)) mn = h2;
  if ((h3 <= h1) && (h3 <= h2)) mn = h3;
  if (c->step <= 10) {
    c->cutoff = mn / 2.0;
    c->cutoff2 = c->cutoff * c->cutoff;
    if (c->skip_ewald != 1) c->a1 = 3.8 / mn;
  }
  c->a1_sqrtpi = c->a1 / kSQRTPI;
  c->v2 = 2.0 / c->vol;
  c->vp = 1.0 / (kPI * c->vol);
  double T[9], Tr[9], VT[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      T[3 * i + j] = H[3 * j + i];
      VT[3 * i + j] = c->prv[3 * j + i];
    }
  cof_inverse(T, Tr);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double a = 0.0;
      for (int k = 0; k < 3; k++) a += Tr[3 * i + k] * VT[3 * k + j];
      c->HmF[3 * i + j] = a;
    }
}

// the step's first kernel: the prologue above (block 0, thread 0) and
// clear_foc (init.c:223-235): f0 <- f, f <- 0
__global__ __launch_bounds__(256) void k_pr_begin(PrCell* __restrict__ c, int n3,
                                                  double* __restrict__ f,
                                                  double* __restrict__ f0) {
  if (blockIdx.x == 0 && threadIdx.x == 0) cell_pre(c);
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n3) return;
  f0[k] = f[k];
  f[k] = 0.0;
}

// cell_force's per-ion part (cell.c:100-111)
__global__ __launch_bounds__(256) void k_pr_cell_force(int N, const PrCell* __restrict__ c,
                                                       const int* __restrict__ typ,
                                                       const double* __restrict__ m,
                                                       const double* __restrict__ v,
                                                       double* __restrict__ f) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const double* F = c->HmF;
  const double mass = m[typ[i]], vx = v[i], vy = v[N + i], vz = v[2 * N + i];
  f[i] -= mass * (F[0] * vx + F[3] * vy + F[6] * vz);
  f[N + i] -= mass * (F[1] * vx + F[4] * vy + F[7] * vz);
  f[2 * N + i] -= mass * (F[2] * vx + F[5] * vy + F[8] * vz);
}

// the step's potential and virial tensor from the routines' sums
// (real.c:20-25, 73-76; ewald.c:36, 50-67, 157, 172)
// out[0, kScal): the fused pair pass (table + ewald1), or the table pass
// alone when Ewald is skipped
__device__ void scal_step(PrCell* __restrict__ c, const double* __restrict__ out, int has_k) {
  double pot = out[0];
  const bool ew = c->skip_ewald != 1;
  if (ew) pot = (pot + (has_k ? out[kOutK] : 0.0)) + out[kOutSelf];
  c->pot = pot;
  const int up[6] = {0, 1, 2, 4, 5, 8};
  for (int k = 0; k < 6; ++k) c->vir[up[k]] = out[1 + k];
  c->vir[3] = c->vir[1];
  c->vir[6] = c->vir[2];
  c->vir[7] = c->vir[5];
  if (out[7] != 0.0) c->bad_cell |= 2;
}

// cyclic_bc (cell.c:255-297): to scaled coordinates, into [0, 1), back
__device__ __forceinline__ void wrap_ion(const PrCell* __restrict__ c, double& x, double& y,
                                         double& z, double& sx, double& sy, double& sz) {
  const double* Rm = c->R;
  const double* H = c->H;
  sx = Rm[0] * x + Rm[3] * y + Rm[6] * z;
  sy = Rm[1] * x + Rm[4] * y + Rm[7] * z;
  sz = Rm[2] * x + Rm[5] * y + Rm[8] * z;
  if (sx >= 1.0) sx -= 1.0;
  if (sy >= 1.0) sy -= 1.0;
  if (sz >= 1.0) sz -= 1.0;
  if (sx < 0.0) sx += 1.0;
  if (sy < 0.0) sy += 1.0;
  if (sz < 0.0) sz += 1.0;
  x = H[0] * sx + H[3] * sy + H[6] * sz;
  y = H[1] * sx + H[4] * sy + H[7] * sz;
  z = H[2] * sx + H[5] * sy + H[8] * sz;
}

// next_v then next_r with cyclic_bc (rv.c:44-66)
__global__ __launch_bounds__(256) void k_pr_next_vr(int N, PrCell* __restrict__ c,
                                                    const double* __restrict__ out, int has_k,
                                                    const int* __restrict__ typ,
                                                    const double* __restrict__ m,
                                                    double* __restrict__ v,
                                                    const double* __restrict__ f,
                                                    const double* __restrict__ f0,
                                                    double* __restrict__ r, double* __restrict__ sc) {
  // the step's potential and virial (block 0, thread 0); this kernel reads
  // none of what that writes
  if (blockIdx.x == 0 && threadIdx.x == 0) scal_step(c, out, has_k);
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const double m2 = 2.0 * m[typ[i]], dt = c->dt, dt2 = c->dt2;
  double p[3], w[3];
  for (int d = 0; d < 3; ++d) {
    double vv = v[d * N + i];
    vv += dt * (f[d * N + i] + f0[d * N + i]) / m2;
    v[d * N + i] = vv;
    p[d] = r[d * N + i] + (dt * vv + dt2 * f[d * N + i] / m2);
  }
  wrap_ion(c, p[0], p[1], p[2], w[0], w[1], w[2]);
  for (int d = 0; d < 3; ++d) {
    r[d * N + i] = p[d];
    sc[d * N + i] = w[d];
  }
}

// control_press_PR's cell update (pt.c:27-57) and calc_cell (cell.c:214-252)
__global__ void k_pr_cell_post(PrCell* __restrict__ c) {
  if (threadIdx.x != 0) return;
  double ff[9];
  for (int k = 0; k < 9; ++k) ff[k] = c->ptens[k] - c->pset[k];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      c->prf0[3 * i + j] = c->prf[3 * i + j];
      double a = 0.0;
      for (int k = 0; k < 3; k++) a += ff[3 * i + k] * c->sig[3 * k + j];
      c->prf[3 * i + j] = a;
    }
  for (int k = 0; k < 9; ++k) c->prv[k] += c->prdt * (c->prf[k] + c->prf0[k]) / (2.0 * c->prmass);
  for (int k = 0; k < 9; ++k)
    c->H[k] += c->prdt * c->prv[k] + (c->prdt * c->prdt) * c->prf[k] / (2.0 * c->prmass);
  const double* H = c->H;
  const double det = cof_inverse(H, c->R);
  if (det == 0.0) c->bad_cell |= 1;  // pr-mdc prints and exits here
  c->vol = fabs(det);
  c->Lx = sqrt(H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
  c->Ly = sqrt(H[3] * H[3] + H[4] * H[4] + H[5] * H[5]);
  c->Lz = sqrt(H[6] * H[6] + H[7] * H[7] + H[8] * H[8]);
  const double a_b = H[0] * H[3] + H[1] * H[4] + H[2] * H[5];
  const double b_c = H[3] * H[6] + H[4] * H[7] + H[5] * H[8];
  const double c_a = H[6] * H[0] + H[7] * H[1] + H[8] * H[2];
  c->ang_ab = (acos(a_b / (c->Lx * c->Ly))) * 180.0 / kPI;
  c->ang_bc = (acos(b_c / (c->Ly * c->Lz))) * 180.0 / kPI;
  c->ang_ca = (acos(c_a / (c->Lz * c->Lx))) * 180.0 / kPI;
}

// control_press_PR's closing cyclic_bc of every ion (pt.c:59-61)
__global__ __launch_bounds__(256) void k_pr_wrap(int N, const PrCell* __restrict__ c,
                                                 double* __restrict__ r, double* __restrict__ sc) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double x = r[i], y = r[N + i], z = r[2 * N + i], a, b, d;
  wrap_ion(c, x, y, z, a, b, d);
  r[i] = x;
  r[N + i] = y;
  r[2 * N + i] = z;
  sc[i] = a;
  sc[N + i] = b;
  sc[2 * N + i] = d;
}

// control_temp's bookkeeping (pt.c:124-140): the kinetic energy it averages
// is the previous step's calc_kin
__device__ void temp_step(PrCell* __restrict__ c) {
  c->count += 1;
  c->oldk += c->kin;
  c->vscale = 0.0;  // 0: no scaling this step
  if (c->t_step > 0 && c->step % c->t_step == 0) {
    const double averaged_K = c->oldk / (double)c->count;
    c->oldk = 0.0;
    c->count = 0;
    const double tempK = c->temp_set / c->e2t;
    c->vscale = sqrt(tempK / averaged_K);
  }
}

// control_temp's scaling (pt.c:141-145), calc_kin (rv.c:68-105) and
// calc_press (pt.c:8-24); writes the step's record row.  One block: each
// thread sums every 256th ion, then a fixed tree.
__global__ __launch_bounds__(256) void k_pr_kin(int N, PrCell* __restrict__ c,
                                                const int* __restrict__ typ,
                                                const double* __restrict__ m,
                                                double* __restrict__ v, double* __restrict__ rec) {
  // rec: the run's record array; this step's row is step - rec_base - 1
  __shared__ double part[6][256];
  __shared__ double sc_sh;
  const int t = threadIdx.x;
  if (t == 0) {
    temp_step(c);
    sc_sh = c->vscale;
  }
  __syncthreads();
  const double sc = sc_sh;
  double k6[6] = {0, 0, 0, 0, 0, 0};
  for (int i = t; i < N; i += 256) {
    double vx = v[i], vy = v[N + i], vz = v[2 * N + i];
    if (sc != 0.0) {
      vx *= sc;
      vy *= sc;
      vz *= sc;
      v[i] = vx;
      v[N + i] = vy;
      v[2 * N + i] = vz;
    }
    const double mm = m[typ[i]];
    k6[0] += mm * vx * vx;
    k6[1] += mm * vx * vy;
    k6[2] += mm * vx * vz;
    k6[3] += mm * vy * vy;
    k6[4] += mm * vy * vz;
    k6[5] += mm * vz * vz;
  }
  for (int k = 0; k < 6; ++k) part[k][t] = k6[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o)
      for (int k = 0; k < 6; ++k) part[k][t] += part[k][t + o];
    __syncthreads();
  }
  if (t == 0) {
    double* K = c->kint;
    K[0] = part[0][0];
    K[1] = part[1][0];
    K[2] = part[2][0];
    K[3] = part[1][0];
    K[4] = part[3][0];
    K[5] = part[4][0];
    K[6] = part[2][0];
    K[7] = part[4][0];
    K[8] = part[5][0];
    c->kin = (K[0] + K[4] + K[8]) / 2.0;
    double* P = c->ptens;
    for (int k = 0; k < 9; ++k) P[k] = (2.0 * K[k] + c->vir[k]) / c->vol;
    c->ptrace = (P[0] + P[4] + P[8]) / 3.0;
    const int row = c->step - c->rec_base - 1;
    if (rec && row >= 0 && row < c->rec_cap) {
      double* o = rec + (size_t)row * kRec;
      o[0] = c->pot;
      o[1] = c->kin;
      o[2] = c->ptrace;
      o[3] = c->vol;
      for (int k = 0; k < 9; ++k) o[4 + k] = c->H[k];
    }
  }
}

// ---------------------------------------------------------------- host

struct State {
  bool ready = false;
  hipStream_t s = nullptr;
  int status = MDG_OK;
  bool reported = false;
  int pair_splits = 0, k_splits = 0;  // 0: chosen by size
  // device buffers and their capacities (elements)
  double *r = nullptr, *f = nullptr, *pf = nullptr, *ps = nullptr, *kc = nullptr, *Z = nullptr,
         *out = nullptr;
  int *typ = nullptr, *lookup = nullptr;
  double2* tab = nullptr;
  int4* hv = nullptr;
  size_t capN = 0, capF = 0, capT = 0, capZ = 0, capL = 0, capH = 0, capPf = 0, capPs = 0,
         capK = 0, capTab = 0;
  // what the device holds
  int N = 0, kinds = 0, td = -1, nint = 0, hm = -1, hm_sqrt = -1;
  std::vector<int> h_typ, h_lookup, h_hv;
  std::vector<double> h_Z, h_pe, h_fe;
  // shapes of the last launches (for mdg_pr_profile)
  PrCell* cell = nullptr;  // device
  double *v = nullptr, *f0 = nullptr, *sc = nullptr, *m = nullptr, *rec = nullptr;
  size_t capV = 0, capF0 = 0, capS = 0, capM = 0, capRec = 0;
  PrCell h_cell{};          // host staging of cell (stable until the call's sync)
  PrParams last_pair{};
  PrK last_k{};
  bool have_pair = false, have_k = false;
  double h_out[kOutLen];
};

State g;

const char* err_name(int code) {
  switch (code) {
    case MDG_E_NODEVICE: return "no HIP device";
    case MDG_E_HIP: return "HIP runtime error";
    case MDG_E_ARG: return "invalid argument";
    case MDG_E_NOMEM: return "device allocation failed";
    case MDG_E_RANGE: return "pair inside the table's 0.5 A start, or cutoff past its end";
    default: return "error";
  }
}

// pr-mdc exits on a fatal error (cell.c:218-221, files.c) and its main.c
// checks no status, so a failing routine exits by default; a host that
// checks mdg_pr_status opts out (mdg_pr_set_error_mode, or
// MDGPU_DROPIN_ERRORS=status in the environment).
int error_mode = -1;  // -1: not chosen yet (the environment decides)

int current_error_mode() {
  if (error_mode < 0) {
    const char* e = getenv("MDGPU_DROPIN_ERRORS");
    error_mode = (e && strcmp(e, "status") == 0) ? MDG_DROPIN_STATUS : MDG_DROPIN_EXIT;
  }
  return error_mode;
}

void fail(int code, const char* what) {
  if (g.status == MDG_OK) g.status = code;
  if (!g.reported) {
    fprintf(stderr, "libmdgpu_pr: %s failed: %s\n", what, err_name(code));
    g.reported = true;
  }
  if (current_error_mode() == MDG_DROPIN_EXIT) {
    fflush(stdout);
    exit(EXIT_FAILURE);
  }
}

#define PR_HIP(x)                            \
  do {                                       \
    if ((x) != hipSuccess) return MDG_E_HIP; \
  } while (0)

template <typename T>
int grow(T** p, size_t* cap, size_t n) {
  if (n <= *cap && *p) return MDG_OK;
  if (*p) PR_HIP(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  if (hipMalloc((void**)p, (n ? n : 1) * sizeof(T)) != hipSuccess) return MDG_E_NOMEM;
  *cap = n;
  return MDG_OK;
}

int init_device() {
  if (g.ready) return MDG_OK;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess || c < 1) return MDG_E_NODEVICE;
  PR_HIP(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
  if (hipMalloc((void**)&g.cell, sizeof(PrCell)) != hipSuccess) return MDG_E_NOMEM;
  if (hipMalloc((void**)&g.out, kOutLen * sizeof(double)) != hipSuccess) return MDG_E_NOMEM;
  g.ready = true;
  return MDG_OK;
}

void mat9(double** m, double* o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[3 * i + j] = m[i][j];
}

// The cell as the force routines read it, from pr-mdc's globals: H, H^-1
// and g as pr-mdc's calc_cell / calc_sigma left them, the cutoff and alpha of
// set_cutoff, and the prefactors ewald1 and ewald2 form (ewald.c:38-41, 90-91).
void fill_cell(PrCell& c) {
  mat9(cell.hm, c.H);
  mat9(cell.hmr, c.R);
  mat9(cell.g, c.G);
  c.vol = cell.vol;
  c.cutoff = sys.cutoff;
  c.cutoff2 = sys.cutoff2;
  c.a1 = sys.a1;
  c.a1_sqrtpi = sys.a1 / kSQRTPI;
  c.v2 = 2.0 / cell.vol;
  c.vp = 1.0 / (kPI * cell.vol);
}

int upload_cell() {
  fill_cell(g.h_cell);
  PR_HIP(hipMemcpyAsync(g.cell, &g.h_cell, sizeof(PrCell), hipMemcpyHostToDevice, g.s));
  return MDG_OK;
}

int splits_for(int nchunk, int limit, int fixed, int target) {
  if (fixed > 0) return fixed < limit ? fixed : limit;
  int s = (target + nchunk - 1) / nchunk;
  if (s > limit) s = limit;
  return s < 1 ? 1 : s;
}

// Uploads positions, forces, species and charges (and, for the table
// routine, the pair table) for the current globals.
int upload_state(bool table) {
  const int N = sys.N;
  int rc;
  if ((rc = grow(&g.r, &g.capN, (size_t)3 * N)) != MDG_OK) return rc;
  if ((rc = grow(&g.f, &g.capF, (size_t)3 * N)) != MDG_OK) return rc;
  if ((rc = grow(&g.typ, &g.capT, (size_t)N)) != MDG_OK) return rc;
  const size_t row = (size_t)N * sizeof(double);
  PR_HIP(hipMemcpyAsync(g.r, cell.rx, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.r + N, cell.ry, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.r + 2 * N, cell.rz, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.f, cell.fx, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.f + N, cell.fy, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.f + 2 * N, cell.fz, row, hipMemcpyHostToDevice, g.s));
  const int kinds = ctl.kinds_of_ions;
  if (g.h_typ.size() != (size_t)N || g.kinds != kinds ||
      memcmp(g.h_typ.data(), sys.ion, (size_t)N * sizeof(int)) != 0) {
    for (int i = 0; i < N; ++i)  // the kernels index the charges and the table by species
      if (sys.ion[i] < 0 || sys.ion[i] >= kinds) return MDG_E_ARG;
    g.h_typ.assign(sys.ion, sys.ion + N);
    PR_HIP(hipMemcpyAsync(g.typ, g.h_typ.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice,
                          g.s));
  }
  if (g.kinds != kinds || g.h_Z.size() != (size_t)kinds ||
      memcmp(g.h_Z.data(), ion.Z, (size_t)kinds * sizeof(double)) != 0) {
    if ((rc = grow(&g.Z, &g.capZ, (size_t)kinds)) != MDG_OK) return rc;
    g.h_Z.assign(ion.Z, ion.Z + kinds);
    PR_HIP(hipMemcpyAsync(g.Z, g.h_Z.data(), (size_t)kinds * sizeof(double),
                          hipMemcpyHostToDevice, g.s));
  }
  if (table) {
    const int td = sys.table_division, nint = ctl.kinds_of_interaction;
    const size_t n = (size_t)nint * (td + 1);
    bool same = g.td == td && g.nint == nint && g.kinds == kinds && g.h_pe.size() == n &&
                memcmp(g.h_pe.data(), sys.pe1r1, n * sizeof(double)) == 0 &&
                memcmp(g.h_fe.data(), sys.fe1r1, n * sizeof(double)) == 0 &&
                g.h_lookup.size() == (size_t)kinds * kinds &&
                memcmp(g.h_lookup.data(), sys.lookup, (size_t)kinds * kinds * sizeof(int)) == 0;
    if (!same) {
      if (td < 0 || nint < 1) return MDG_E_ARG;
      for (int e = 0; e < kinds * kinds; ++e)
        if (sys.lookup[e] < 0 || sys.lookup[e] >= nint) return MDG_E_ARG;
      if ((rc = grow(&g.tab, &g.capTab, n)) != MDG_OK) return rc;
      if ((rc = grow(&g.lookup, &g.capL, (size_t)kinds * kinds)) != MDG_OK) return rc;
      g.h_pe.assign(sys.pe1r1, sys.pe1r1 + n);
      g.h_fe.assign(sys.fe1r1, sys.fe1r1 + n);
      g.h_lookup.assign(sys.lookup, sys.lookup + kinds * kinds);
      std::vector<double2> t(n);
      for (size_t e = 0; e < n; ++e) t[e] = make_double2(g.h_pe[e], g.h_fe[e]);
      PR_HIP(hipMemcpyAsync(g.tab, t.data(), n * sizeof(double2), hipMemcpyHostToDevice, g.s));
      PR_HIP(hipMemcpyAsync(g.lookup, g.h_lookup.data(), (size_t)kinds * kinds * sizeof(int),
                            hipMemcpyHostToDevice, g.s));
      PR_HIP(hipStreamSynchronize(g.s));  // t is a local
      g.td = td;
      g.nint = nint;
    }
  }
  g.N = N;
  g.kinds = kinds;
  return MDG_OK;
}

int download_forces() {
  const int N = sys.N;
  const size_t row = (size_t)N * sizeof(double);
  PR_HIP(hipMemcpyAsync(cell.fx, g.f, row, hipMemcpyDeviceToHost, g.s));
  PR_HIP(hipMemcpyAsync(cell.fy, g.f + N, row, hipMemcpyDeviceToHost, g.s));
  PR_HIP(hipMemcpyAsync(cell.fz, g.f + 2 * N, row, hipMemcpyDeviceToHost, g.s));
  PR_HIP(hipMemcpyAsync(g.h_out, g.out, kOutLen * sizeof(double), hipMemcpyDeviceToHost, g.s));
  PR_HIP(hipStreamSynchronize(g.s));
  return MDG_OK;
}

int enqueue_pairs(int mode, const PrParams& p, int out_at = 0) {
  const int nchunk = (p.N + kW - 1) / kW;
  int rc;
  if ((rc = grow(&g.ps, &g.capPs, (size_t)nchunk * p.S * kScal)) != MDG_OK) return rc;
  if (p.S > 1 && (rc = grow(&g.pf, &g.capPf, (size_t)p.S * 3 * p.N)) != MDG_OK) return rc;
  const dim3 grid(nchunk, p.S);
  if (mode == kPrTable)
    hipLaunchKernelGGL(k_pr_pairs<kPrTable>, grid, dim3(kW), 0, g.s, p, g.cell, g.r, g.typ, g.tab,
                       g.lookup, g.Z, g.f, g.pf, g.ps);
  else if (mode == kPrEwald1)
    hipLaunchKernelGGL(k_pr_pairs<kPrEwald1>, grid, dim3(kW), 0, g.s, p, g.cell, g.r, g.typ, g.tab,
                       g.lookup, g.Z, g.f, g.pf, g.ps);
  else
    hipLaunchKernelGGL(k_pr_pairs<kPrBoth>, grid, dim3(kW), 0, g.s, p, g.cell, g.r, g.typ, g.tab,
                       g.lookup, g.Z, g.f, g.pf, g.ps);
  PR_HIP(hipGetLastError());
  if (p.S > 1) {
    hipLaunchKernelGGL(k_pr_fsum, dim3((3 * p.N + 255) / 256), dim3(256), 0, g.s, 3 * p.N, p.S,
                       g.pf, g.f);
    PR_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_pr_ssum, dim3(1), dim3(kSsumThreads), 0, g.s, nchunk * p.S, g.ps,
                     g.out + out_at);
  PR_HIP(hipGetLastError());
  return MDG_OK;
}

int enqueue_kspace(const PrK& p) {
  int rc;
  if ((rc = grow(&g.kc, &g.capK, (size_t)p.K * 8)) != MDG_OK) return rc;
  if (p.S > 1 && (rc = grow(&g.pf, &g.capPf, (size_t)p.S * 3 * p.N)) != MDG_OK) return rc;
  hipLaunchKernelGGL(k_pr_sk, dim3(p.K), dim3(kSkThreads), 0, g.s, p, g.cell, g.hv, g.r, g.typ, g.Z,
                     g.kc);
  PR_HIP(hipGetLastError());
  const int nchunk = (p.N + kW - 1) / kW;
  hipLaunchKernelGGL(k_pr_kforce, dim3(nchunk, p.S), dim3(kW), 0, g.s, p, g.r, g.typ, g.Z, g.kc,
                     g.f, g.pf);
  PR_HIP(hipGetLastError());
  if (p.S > 1) {
    hipLaunchKernelGGL(k_pr_fsum, dim3((3 * p.N + 255) / 256), dim3(256), 0, g.s, 3 * p.N, p.S,
                       g.pf, g.f);
    PR_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_pr_ksum, dim3(1), dim3(256), 0, g.s, p.K, g.kc, g.out + kOutK);
  PR_HIP(hipGetLastError());
  return MDG_OK;
}

int enqueue_self(int N) {
  hipLaunchKernelGGL(k_pr_self, dim3(1), dim3(256), 0, g.s, N, g.cell, g.typ, g.Z,
                     g.out + kOutSelf);
  PR_HIP(hipGetLastError());
  return MDG_OK;
}

PrParams pair_params() {
  PrParams p{};
  p.N = sys.N;
  p.kinds = ctl.kinds_of_ions;
  p.td = sys.table_division;
  const int nchunk = (p.N + kW - 1) / kW, ntile = (p.N + kPT - 1) / kPT;
  p.S = splits_for(nchunk, ntile, g.pair_splits, kPairWaves);
  return p;
}

// The half-sphere vectors of ewald.c:96-105, in pr-mdc's loop order.
int kvectors() {
  if (g.hm == sys.hm && g.hm_sqrt == sys.hm_sqrt) return MDG_OK;
  std::vector<int> hv;
  const int hs = sys.hm_sqrt;
  for (int hx = -hs; hx <= hs; hx++)
    for (int hy = -hs; hy <= hs; hy++)
      for (int hz = 0; hz <= hs; hz++) {
        if (hz == 0 && (hy < 0 || (hy == 0 && hx <= 0))) continue;
        const int h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > sys.hm) continue;
        hv.push_back(hx);
        hv.push_back(hy);
        hv.push_back(hz);
        hv.push_back(0);
      }
  const size_t K = hv.size() / 4;
  int rc;
  if ((rc = grow(&g.hv, &g.capH, K)) != MDG_OK) return rc;
  if (K) {
    PR_HIP(hipMemcpyAsync(g.hv, hv.data(), K * sizeof(int4), hipMemcpyHostToDevice, g.s));
    PR_HIP(hipStreamSynchronize(g.s));
  }
  g.h_hv.swap(hv);
  g.hm = sys.hm;
  g.hm_sqrt = sys.hm_sqrt;
  return MDG_OK;
}

PrK k_params() {
  PrK p{};
  p.N = sys.N;
  p.K = (int)(g.h_hv.size() / 4);
  const int nchunk = (p.N + kW - 1) / kW;
  const int klimit = p.K / kMinKPerSplit > 0 ? p.K / kMinKPerSplit : 1;
  p.S = g.k_splits > 0 ? splits_for(nchunk, p.K > 0 ? p.K : 1, g.k_splits, kKWaves)
                       : splits_for(nchunk, klimit, 0, kKWaves);
  return p;
}

bool enter(const char* what) {
  if (g.status != MDG_OK) return false;
  const int rc = init_device();
  if (rc != MDG_OK) {
    fail(rc, what);
    return false;
  }
  if (sys.N < 0 || ctl.kinds_of_ions < 1 || ctl.kinds_of_ions > 16) {
    fail(MDG_E_ARG, what);
    return false;
  }
  return true;
}

// real.c:11-77 and ewald.c:11-68: the two pair routines
void pair_routine(int mode, const char* what) {
  if (mode == kPrTable) {
    sys.pot = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) pt.virial[i][j] = 0.0;
  }
  if (sys.N > 0) {
    if (!enter(what)) return;
    if (mode == kPrTable && sqrt(sys.cutoff2) > 0.5 + 0.001 * sys.table_division) {
      fail(MDG_E_RANGE, what);  // the cutoff reaches past the table's last bin
      return;
    }
    int rc = upload_state(mode == kPrTable);
    if (rc == MDG_OK) rc = upload_cell();
    const PrParams p = pair_params();
    if (rc == MDG_OK) rc = enqueue_pairs(mode, p);
    if (rc == MDG_OK) rc = download_forces();
    if (rc != MDG_OK) {
      fail(rc, what);
      return;
    }
    g.last_pair = p;
    g.have_pair = true;
    const double* o = g.h_out;
    if (o[7] != 0.0) fail(MDG_E_RANGE, what);  // a pair inside the table's 0.5 A start
    sys.pot += o[0];
    pt.virial[0][0] += o[1];
    pt.virial[0][1] += o[2];
    pt.virial[0][2] += o[3];
    pt.virial[1][1] += o[4];
    pt.virial[1][2] += o[5];
    pt.virial[2][2] += o[6];
  }
  pt.virial[1][0] = pt.virial[0][1];
  pt.virial[2][0] = pt.virial[0][2];
  pt.virial[2][1] = pt.virial[1][2];
}

}  // namespace

extern "C" {

void real_space(void) { pair_routine(kPrTable, "real_space"); }

void ewald1(void) { pair_routine(kPrEwald1, "ewald1"); }

// ewald.c:75-159
void ewald2(void) {
  if (sys.N <= 0) return;
  if (!enter("ewald2")) return;
  int rc = kvectors();
  if (rc == MDG_OK) rc = upload_state(false);
  if (rc == MDG_OK) rc = upload_cell();
  const PrK p = k_params();
  if (rc == MDG_OK && p.K > 0) {
    rc = enqueue_kspace(p);
    if (rc == MDG_OK) rc = download_forces();
  }
  if (rc != MDG_OK) {
    fail(rc, "ewald2");
    return;
  }
  if (p.K > 0) {
    g.last_k = p;
    g.have_k = true;
    sys.pot += g.h_out[kOutK];
  }
}

// ewald.c:162-173
void ewald3(void) {
  if (sys.N <= 0) {
    sys.pot += -0.0 * (sys.a1 / kSQRTPI);
    return;
  }
  if (!enter("ewald3")) return;
  int rc = upload_state(false);
  if (rc == MDG_OK) rc = upload_cell();
  if (rc == MDG_OK) rc = enqueue_self(sys.N);
  if (rc == MDG_OK) {
    if (hipMemcpyAsync(g.h_out, g.out, kOutLen * sizeof(double), hipMemcpyDeviceToHost, g.s) !=
            hipSuccess ||
        hipStreamSynchronize(g.s) != hipSuccess)
      rc = MDG_E_HIP;
  }
  if (rc != MDG_OK) {
    fail(rc, "ewald3");
    return;
  }
  sys.pot += g.h_out[kOutSelf];
}

// ewald.c:178-193
void calc_alpha(void) {
  sys.a1 = 0.01;
  for (int i = 0; i < 200; i++) {
    sys.pot = 0.0;
    ewald1();
    ewald2();
    ewald3();
    printf("%f  %f\n", sys.a1, sys.pot);
    sys.a1 += 0.01;
  }
  exit(0);
}

// main.c:100-111 (pr-mdc's step loop without its output and MSD), resident
int mdg_pr_run(int nsteps, double* record) {
  if (nsteps < 0) return MDG_E_ARG;
  // an entry point with a return code: it reports, it never exits
  if (g.status != MDG_OK) return g.status;
  if (const int rc0 = init_device(); rc0 != MDG_OK) return rc0;
  if (sys.N < 0 || ctl.kinds_of_ions < 1 || ctl.kinds_of_ions > 16) return MDG_E_ARG;
  const int N = sys.N;
  if (N < 1 || !pt.pres_set || !pt.pres_tensor || !pt.virial || !pt.kin_tensor || !pr.v ||
      !pr.f || !pr.f0 || !cell.sigma || !cell.vx || !cell.fx0 || !ion.m || pr.mass == 0.0)
    return MDG_E_ARG;
  if (sqrt(sys.cutoff2) > 0.5 + 0.001 * sys.table_division) return MDG_E_RANGE;
  int rc = kvectors();
  if (rc == MDG_OK) rc = upload_state(true);
  const int kinds = ctl.kinds_of_ions;
  const size_t row = (size_t)N * sizeof(double);
  if (rc == MDG_OK) rc = grow(&g.v, &g.capV, (size_t)3 * N);
  if (rc == MDG_OK) rc = grow(&g.f0, &g.capF0, (size_t)3 * N);
  if (rc == MDG_OK) rc = grow(&g.sc, &g.capS, (size_t)3 * N);
  if (rc == MDG_OK) rc = grow(&g.m, &g.capM, (size_t)kinds);
  if (rc == MDG_OK) rc = grow(&g.rec, &g.capRec, (size_t)(nsteps > 0 ? nsteps : 1) * kRec);
  if (rc != MDG_OK) return rc;
  PR_HIP(hipMemcpyAsync(g.v, cell.vx, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.v + N, cell.vy, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.v + 2 * N, cell.vz, row, hipMemcpyHostToDevice, g.s));
  PR_HIP(hipMemcpyAsync(g.m, ion.m, (size_t)kinds * sizeof(double), hipMemcpyHostToDevice, g.s));
  PrCell& c = g.h_cell;
  c = PrCell{};
  fill_cell(c);
  mat9(cell.sigma, c.sig);
  mat9(pr.v, c.prv);
  mat9(pr.f, c.prf);
  mat9(pr.f0, c.prf0);
  c.prmass = pr.mass;
  c.prdt = pr.dt;
  mat9(pt.pres_set, c.pset);
  mat9(pt.pres_tensor, c.ptens);
  mat9(pt.virial, c.vir);
  mat9(pt.kin_tensor, c.kint);
  c.pot = sys.pot;
  c.kin = sys.kin;
  c.ptrace = pt.pres_trace;
  c.Lx = cell.Lx;
  c.Ly = cell.Ly;
  c.Lz = cell.Lz;
  c.ang_ab = cell.angle_ab;
  c.ang_bc = cell.angle_bc;
  c.ang_ca = cell.angle_ca;
  c.dt = sys.dt;
  c.dt2 = sys.dt2;
  c.e2t = sys.e2t;
  c.temp_set = pt.temp_set;
  c.t_step = pt.t_control_step;
  c.skip_ewald = ctl.skip_ewald;
  c.step = sys.step;
  c.rec_base = sys.step;
  c.rec_cap = nsteps;
  PR_HIP(hipMemcpyAsync(g.cell, &c, sizeof(PrCell), hipMemcpyHostToDevice, g.s));
  const PrParams pp = pair_params();
  const PrK pk = k_params();
  const int nb = (N + 255) / 256, n3b = (3 * N + 255) / 256;
  // one step's launches (every buffer they use is allocated by the first,
  // eager step, so that a captured block allocates nothing)
  auto step = [&]() -> int {
    hipLaunchKernelGGL(k_pr_begin, dim3(n3b), dim3(256), 0, g.s, g.cell, 3 * N, g.f, g.f0);
    // real_space and ewald1 in one pair pass (cell_force then follows both)
    int e = enqueue_pairs(ctl.skip_ewald != 1 ? kPrBoth : kPrTable, pp, 0);
    if (e != MDG_OK) return e;
    hipLaunchKernelGGL(k_pr_cell_force, dim3(nb), dim3(256), 0, g.s, N, g.cell, g.typ, g.m, g.v,
                       g.f);
    if (ctl.skip_ewald != 1) {
      if (pk.K > 0) e = enqueue_kspace(pk);
      if (e == MDG_OK) e = enqueue_self(N);
      if (e != MDG_OK) return e;
    }
    hipLaunchKernelGGL(k_pr_next_vr, dim3(nb), dim3(256), 0, g.s, N, g.cell, g.out, pk.K > 0,
                       g.typ, g.m, g.v, g.f, g.f0, g.r, g.sc);
    hipLaunchKernelGGL(k_pr_cell_post, dim3(1), dim3(kW), 0, g.s, g.cell);
    hipLaunchKernelGGL(k_pr_wrap, dim3(nb), dim3(256), 0, g.s, N, g.cell, g.r, g.sc);
    hipLaunchKernelGGL(k_pr_kin, dim3(1), dim3(256), 0, g.s, N, g.cell, g.typ, g.m, g.v, g.rec);
    return hipGetLastError() == hipSuccess ? MDG_OK : MDG_E_HIP;
  };
  int done = 0;
  if (nsteps > 0) {
    rc = step();
    done = 1;
  }
  // blocks of kGraphSteps steps as one hipGraph (about 25 launches a step)
  constexpr int kGraphSteps = 16;
  if (rc == MDG_OK && nsteps - done >= 2 * kGraphSteps) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ok = hipStreamBeginCapture(g.s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    int e = MDG_OK;
    if (ok) {
      for (int k = 0; k < kGraphSteps && e == MDG_OK; ++k) e = step();
      ok = hipStreamEndCapture(g.s, &graph) == hipSuccess && e == MDG_OK;
    }
    if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (ok) {
      while (rc == MDG_OK && nsteps - done >= kGraphSteps) {
        if (hipGraphLaunch(exec, g.s) != hipSuccess) rc = MDG_E_HIP;
        done += kGraphSteps;
      }
    } else {
      (void)hipGetLastError();  // capture refused: the same launches, one by one
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  for (; done < nsteps && rc == MDG_OK; ++done) rc = step();
  if (rc != MDG_OK) return rc;
  // write back what pr-mdc's routines write
  double* hr[3] = {cell.rx, cell.ry, cell.rz};
  double* hv[3] = {cell.vx, cell.vy, cell.vz};
  double* hf[3] = {cell.fx, cell.fy, cell.fz};
  double* hf0[3] = {cell.fx0, cell.fy0, cell.fz0};
  double* hs[3] = {cell.sx, cell.sy, cell.sz};
  for (int d = 0; d < 3; ++d) {
    PR_HIP(hipMemcpyAsync(hr[d], g.r + (size_t)d * N, row, hipMemcpyDeviceToHost, g.s));
    PR_HIP(hipMemcpyAsync(hv[d], g.v + (size_t)d * N, row, hipMemcpyDeviceToHost, g.s));
    PR_HIP(hipMemcpyAsync(hf[d], g.f + (size_t)d * N, row, hipMemcpyDeviceToHost, g.s));
    PR_HIP(hipMemcpyAsync(hf0[d], g.f0 + (size_t)d * N, row, hipMemcpyDeviceToHost, g.s));
    if (nsteps > 0 && hs[d])
      PR_HIP(hipMemcpyAsync(hs[d], g.sc + (size_t)d * N, row, hipMemcpyDeviceToHost, g.s));
  }
  PR_HIP(hipMemcpyAsync(&c, g.cell, sizeof(PrCell), hipMemcpyDeviceToHost, g.s));
  if (record && nsteps > 0)
    PR_HIP(hipMemcpyAsync(record, g.rec, (size_t)nsteps * kRec * sizeof(double),
                          hipMemcpyDeviceToHost, g.s));
  PR_HIP(hipStreamSynchronize(g.s));
  auto put9 = [](const double* a, double** m) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m[i][j] = a[3 * i + j];
  };
  put9(c.H, cell.hm);
  put9(c.R, cell.hmr);
  put9(c.sig, cell.sigma);
  put9(c.G, cell.g);
  put9(c.prv, pr.v);
  put9(c.prf, pr.f);
  put9(c.prf0, pr.f0);
  put9(c.ptens, pt.pres_tensor);
  put9(c.vir, pt.virial);
  put9(c.kint, pt.kin_tensor);
  cell.vol = c.vol;
  cell.Lx = c.Lx;
  cell.Ly = c.Ly;
  cell.Lz = c.Lz;
  cell.angle_ab = c.ang_ab;
  cell.angle_bc = c.ang_bc;
  cell.angle_ca = c.ang_ca;
  sys.pot = c.pot;
  sys.kin = c.kin;
  sys.a1 = c.a1;
  sys.cutoff = c.cutoff;
  sys.cutoff2 = c.cutoff2;
  sys.step = c.step;
  pt.pres_trace = c.ptrace;
  // the device's f now differs from what the drop-in's next upload brings; it
  // re-uploads per call, so nothing is stale
  if (c.bad_cell & 1) return MDG_E_STATE;
  if (c.bad_cell & 2) return MDG_E_RANGE;
  return MDG_OK;
}

int mdg_pr_set_error_mode(int mode) {
  if (mode != MDG_DROPIN_EXIT && mode != MDG_DROPIN_STATUS) return MDG_E_ARG;
  error_mode = mode;
  return MDG_OK;
}

int mdg_pr_status(void) { return g.status; }

void mdg_pr_clear_status(void) {
  g.status = MDG_OK;
  g.reported = false;
}

int mdg_pr_set_splits(int pair, int kspace) {
  if (pair < 0 || kspace < 0) return MDG_E_ARG;
  g.pair_splits = pair;
  g.k_splits = kspace;
  return MDG_OK;
}

int mdg_pr_profile(int reps, double* ms_eval, double* ms_kernel) {
  if (reps < 1 || !ms_eval) return MDG_E_ARG;
  if (!g.ready || !g.have_pair || g.td < 0 || g.N < 1) return MDG_E_STATE;
  PrParams p = g.last_pair;
  if (p.N != g.N) return MDG_E_STATE;
  const bool kspace = g.have_k && g.last_k.N == g.N;
  hipEvent_t ev[5];
  for (auto& e : ev) PR_HIP(hipEventCreate(&e));
  double acc[4] = {0, 0, 0, 0};
  int rc = MDG_OK;
  for (int it = 0; it <= reps && rc == MDG_OK; ++it) {  // iteration 0 warms up
    (void)hipEventRecord(ev[0], g.s);
    rc = enqueue_pairs(kPrTable, p);
    (void)hipEventRecord(ev[1], g.s);
    if (rc == MDG_OK) rc = enqueue_pairs(kPrEwald1, p);
    (void)hipEventRecord(ev[2], g.s);
    if (rc == MDG_OK && kspace) rc = enqueue_kspace(g.last_k);
    (void)hipEventRecord(ev[3], g.s);
    if (rc == MDG_OK) rc = enqueue_self(p.N);
    (void)hipEventRecord(ev[4], g.s);
    if (rc == MDG_OK && hipEventSynchronize(ev[4]) != hipSuccess) rc = MDG_E_HIP;
    if (rc == MDG_OK && it > 0)
      for (int k = 0; k < 4; ++k) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
        acc[k] += ms;
      }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc != MDG_OK) return rc;
  double tot = 0;
  for (int k = 0; k < 4; ++k) {
    acc[k] /= reps;
    tot += acc[k];
    if (ms_kernel) ms_kernel[k] = acc[k];
  }
  *ms_eval = tot;
  return MDG_OK;
}

}  // extern "C"
