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

enum { kPrTable = 0, kPrEwald1 = 1 };

struct PrParams {
  int N, kinds, td, S;
  double cutoff2, a1, a1_sqrtpi;
  double H[9];  // cell.hm[i][j] at 3i+j
  double R[9];  // cell.hmr[i][j] at 3i+j
};

struct PrK {
  int N, K, S;
  double g[9];  // cell.g[i][j] at 3i+j
  double a1, v2, vp;
};

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
__device__ __forceinline__ void min_image(const PrParams& p, double& dx, double& dy, double& dz) {
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
__global__ __launch_bounds__(kW) void k_pr_pairs(PrParams p, const double* __restrict__ r,
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
      min_image(p, dx, dy, dz);
      double dr = dx * dx + dy * dy + dz * dz;
      if (dr > p.cutoff2) continue;
      double fij;
      if constexpr (MODE == kPrTable) {
        // real.c:43-56: the 0.001 A bin of r from 0.5 A
        const int tl = lo ? ti : st[k], th = lo ? st[k] : ti;
        const int ddr = (int)((sqrt(dr) - 0.5) * 1000.0);
        if (ddr < 0 || ddr > p.td) {  // inside the table's first bin: pr-mdc reads out of bounds
          bad += 1.0;
          continue;
        }
        const double2 e = tab[(size_t)lookup[tl * p.kinds + th] * (p.td + 1) + ddr];
        fij = e.y;
        if (lo) pot += e.x;
      } else {
        // ewald.c:32-43
        dr = sqrt(dr);
        const double ZZ = lo ? Zi * Z[st[k]] : Z[st[k]] * Zi;
        const double adr = p.a1 * dr;
        const double e = pr_erfc(adr) / dr;
        if (lo) pot += ZZ * e;
        fij = ZZ * (2.0 * p.a1_sqrtpi * exp(-adr * adr) + e) / (dr * dr);
      }
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
__global__ __launch_bounds__(kSkThreads) void k_pr_sk(PrK p, const int4* __restrict__ hv,
                                                      const double* __restrict__ r,
                                                      const int* __restrict__ typ,
                                                      const double* __restrict__ Z,
                                                      double* __restrict__ kc) {
  __shared__ double red[2][kSkThreads / kW];
  const int k = blockIdx.x, N = p.N, tid = threadIdx.x;
  const int4 h = hv[k];
  const double hx = (double)h.x, hy = (double)h.y, hz = (double)h.z;
  const double hhx = hx * p.g[0] + hy * p.g[3] + hz * p.g[6];
  const double hhy = hx * p.g[1] + hy * p.g[4] + hz * p.g[7];
  const double hhz = hx * p.g[2] + hy * p.g[5] + hz * p.g[8];
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
    const double ex = exp(-hh2 * (kPIPI / (p.a1 * p.a1)));
    const double vpe = p.vp * (ex / hh2), v2e = p.v2 * (ex / hh2);
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
__global__ __launch_bounds__(256) void k_pr_self(int N, double a1_sqrtpi,
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
    out[0] = -z2 * a1_sqrtpi;
  }
}

// ---------------------------------------------------------------- host
// out slots: [0, kScal) pair-kernel sums, kOutK reciprocal energy, kOutSelf
enum { kOutK = kScal, kOutSelf = kScal + 1, kOutLen = 16 };

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

void fail(int code, const char* what) {
  if (g.status == MDG_OK) g.status = code;
  if (!g.reported) {
    fprintf(stderr, "libmdgpu_pr: %s failed: %s\n", what, err_name(code));
    g.reported = true;
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
  if (hipMalloc((void**)&g.out, kOutLen * sizeof(double)) != hipSuccess) return MDG_E_NOMEM;
  g.ready = true;
  return MDG_OK;
}

void mat9(double** m, double* o) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[3 * i + j] = m[i][j];
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

int enqueue_pairs(int mode, const PrParams& p) {
  const int nchunk = (p.N + kW - 1) / kW;
  int rc;
  if ((rc = grow(&g.ps, &g.capPs, (size_t)nchunk * p.S * kScal)) != MDG_OK) return rc;
  if (p.S > 1 && (rc = grow(&g.pf, &g.capPf, (size_t)p.S * 3 * p.N)) != MDG_OK) return rc;
  const dim3 grid(nchunk, p.S);
  if (mode == kPrTable)
    hipLaunchKernelGGL(k_pr_pairs<kPrTable>, grid, dim3(kW), 0, g.s, p, g.r, g.typ, g.tab,
                       g.lookup, g.Z, g.f, g.pf, g.ps);
  else
    hipLaunchKernelGGL(k_pr_pairs<kPrEwald1>, grid, dim3(kW), 0, g.s, p, g.r, g.typ, g.tab,
                       g.lookup, g.Z, g.f, g.pf, g.ps);
  PR_HIP(hipGetLastError());
  if (p.S > 1) {
    hipLaunchKernelGGL(k_pr_fsum, dim3((3 * p.N + 255) / 256), dim3(256), 0, g.s, 3 * p.N, p.S,
                       g.pf, g.f);
    PR_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_pr_ssum, dim3(1), dim3(kSsumThreads), 0, g.s, nchunk * p.S, g.ps, g.out);
  PR_HIP(hipGetLastError());
  return MDG_OK;
}

int enqueue_kspace(const PrK& p) {
  int rc;
  if ((rc = grow(&g.kc, &g.capK, (size_t)p.K * 8)) != MDG_OK) return rc;
  if (p.S > 1 && (rc = grow(&g.pf, &g.capPf, (size_t)p.S * 3 * p.N)) != MDG_OK) return rc;
  hipLaunchKernelGGL(k_pr_sk, dim3(p.K), dim3(kSkThreads), 0, g.s, p, g.hv, g.r, g.typ, g.Z, g.kc);
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

int enqueue_self(int N, double a1_sqrtpi) {
  hipLaunchKernelGGL(k_pr_self, dim3(1), dim3(256), 0, g.s, N, a1_sqrtpi, g.typ, g.Z,
                     g.out + kOutSelf);
  PR_HIP(hipGetLastError());
  return MDG_OK;
}

PrParams pair_params() {
  PrParams p{};
  p.N = sys.N;
  p.kinds = ctl.kinds_of_ions;
  p.td = sys.table_division;
  p.cutoff2 = sys.cutoff2;
  p.a1 = sys.a1;
  p.a1_sqrtpi = sys.a1 / kSQRTPI;
  mat9(cell.hm, p.H);
  mat9(cell.hmr, p.R);
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
  mat9(cell.g, p.g);
  p.a1 = sys.a1;
  p.v2 = 2.0 / cell.vol;
  p.vp = 1.0 / (kPI * cell.vol);
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
  if (rc == MDG_OK) rc = enqueue_self(sys.N, sys.a1 / kSQRTPI);
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
    if (rc == MDG_OK) rc = enqueue_self(p.N, p.a1_sqrtpi);
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
