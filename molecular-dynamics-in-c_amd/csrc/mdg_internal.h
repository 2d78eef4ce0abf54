// mdg_internal.h -- types shared by the host orchestration and the
// CDNA4 kernels of libmdgpu.
//
// Data layout in HBM.  Every box owns one device allocation carved into
// SoA arrays (positions, velocities, forces and previous forces as FP64
// x[N], y[N], z[N]; species as int32), its k-space data (k columns and
// segments, per-k coefficients, per-ion phase tables, partial sums) and
// its cell list.  A BoxDesc holds the device pointers and constants of one
// box; a launch over many boxes reads one BoxDesc per work item, so boxes
// of any size share a launch without repacking their arrays.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdg {

constexpr int kWave = 64;           // CDNA wavefront
constexpr int kRealThreads = 64;    // one wave per real-space work item
constexpr int kSkThreads = 256;     // structure factor: one k segment per thread, shared ion tiles
constexpr int kSkTile = 32;         // ions staged in LDS per structure-factor tile
constexpr int kSegLen = 8;          // k vectors per segment (one column, consecutive hz)
constexpr int kKfThreads = 64;      // back-projection: one wave, one ion per lane
constexpr int kKfRows = 64;         // weight rows (4 columns each) staged in LDS at a time
constexpr int kIonThreads = 256;    // per-ion streaming kernels
constexpr int kCellThreads = 1024;  // cell counts and starts, one workgroup per box
constexpr int kCellLds = 4096;      // cells whose counters live in LDS
constexpr int kCellFusedN = 8192;   // boxes up to this size build their cell list in one workgroup
// Ions per box at most.  With hmax <= 63 the kernels' 32-bit byte offsets
// hold: the mixed k-space's Z.W records (16 B x up to 64 hz per ion,
// < 2^31 in all) and its X and Y tables addressed from X (16 B x up to 65
// rows per ion, < 2^32); and index products fit 24-bit multiplies (8 N < 2^24).
constexpr int kMaxIons = 1 << 21;

// A box whose whole cell list is built by k_cell_count's one workgroup (bin,
// count, scan, scatter and placement in LDS): few enough ions for an LDS
// slot array and few enough cells for LDS counters.
__host__ __device__ constexpr bool cell_fused(int n, int ncell) {
  return n <= kCellFusedN && ncell <= kCellLds;
}
constexpr int kScalThreads = 256;   // per-box scalar reduction

// error bits written by kernels into BoxDesc::err
constexpr int kErrState = 1;        // non-finite / out-of-box coordinate
constexpr int kErrRange = 2;        // pair closer than the table start
// scale of the fixed-point real-space partner sums (BoxDesc::fix)
constexpr double kFixScale = 4294967296.0;  // 2^32

// One column of the half-sphere k set: fixed (hx, hy), hz = hz0 .. hz0+len-1.
// Columns are ordered by hx, then hy, as mdc's loops visit them
// (mdc/program/rcprcl.c:32-38), so the k index order equals mdc's.
struct KCol {
  int hx, hy, hz0, len;
  int k0;            // index of the column's first k in the per-k arrays
  int pad[3];
};

// kSegLen consecutive hz of one column starting at hz0 (a multiple of
// kSegLen); slots qlo .. n-1 are k vectors k0 + q of the column (k0 may
// precede the column's first k when qlo > 0).
struct KSeg {
  int hx, hy, hz0, n;
  int k0, qlo, pad[2];
};

// Four consecutive k columns, back-projected together.  Their weights sit
// in rows hz = lo .. lo+nrow-1 of a zero-padded [row][4 columns] layout
// starting at row0; a column shorter than the quad's hz span has zeros.
struct KQuad {
  int hx[4], hy[4];
  int lo, nrow, row0, ncol;
};

// Eight consecutive k columns back-projected by one MFMA tile (mixed
// path): columns c0 .. c0+nc-1, their weights nonzero only in K blocks
// (8 hz each) kb_lo .. kb_hi-1.
struct KfTile {
  int c0, nc, kb_lo, kb_hi;
};
constexpr int kKfTileCols = 8;
// 16-byte units of one tile in the packed operand buffer: its column
// records, then kbk K blocks of (hi, lo) x 64 lane records.  A tile's K
// blocks start at block 0 (every column's hz range starts at 0 or 1), so the
// blocks it uses follow its records in one run.
__host__ __device__ inline int kf_tile_stride(int kbk) { return kKfTileCols + kbk * 2 * 64; }
constexpr int kKfMfmaWaves = 4;     // waves per back-projection workgroup
constexpr int kSkmWaves = 4;        // waves per MFMA structure-factor workgroup (32 columns each)
constexpr int kSkmChunk = 512;      // mixed structure factor: ions per FP32 partial sum

// Work item of a launch over many boxes.
struct Item {
  int box;   // index into the BoxDesc array of the launch
  int a;     // first ion / cell / chunk / k, by kernel
  int b;     // partial-sum slot, first segment or first column
  int c;     // split index / segment count / column count
  int d;     // back-projection partial slot
};

// (a s) mod n for 0 <= a, s < 2^21: the product is exact in FP64 and so is
// the remainder, an integer, formed by one FMA
__device__ __forceinline__ int mul_mod(int a, int st, int n, double inv_n) {
  const double x = (double)a * (double)st;
  const double q = floor(x * inv_n);
  double r = fma(-q, (double)n, x);
  r = r < 0.0 ? r + n : (r >= (double)n ? r - n : r);
  return (int)r;
}

struct BoxDesc {
  // ---- ion state (N each) ----
  double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz, *fx0, *fy0, *fz0;
  int *typ;
  int N;
  int ewald;          // 0: Ewald skipped (alpha == 0 or alpha > 10)
  int kspace;         // 1: ewald and at least one k vector (reciprocal work)
  int real_mode;      // 0 table, 1 analytic
  int all_pairs;      // 1: O(N^2) pair loop in mdc's order; 0: cell list
  int rstride;        // mixed all pairs: the real space visits ions in the order
                      // p -> (p * rstride) mod N (coprime with N; 1: index order)
  int rstride_inv;    // its inverse mod N: ion i sits at position i * rstride_inv
                      // mod N, the index of its real-space force in `fix`
  int kinds;          // species count (table lookup stride)
  int lookup[16];     // pair type of (ti, tj): lookup[ti*kinds + tj]
  // ---- geometry and Ewald constants ----
  double L[3], halfL[3], invL[3];
  double rc2;
  double alpha, a3;   // alpha, alpha/sqrt(pi)
  double dt, dt2;
  double Z[4], m[4];
  double self_pot;    // -a3 * sum Z^2
  double vol;
  // ---- BMH (analytic mode) per pair type ----
  double cb[3], sig[3], rho, Cd[3], Dd[3];
  // ---- pair table (table mode): {pe, fe} per 0.001 A bin ----
  const double2 *table;
  const float2 *table32;
  int table_nb;       // bins per pair type (table_division + 1)
  // ---- real-space split: partners divided over rsplit items per home unit ----
  int rsplit;         // 1: forces written in place (mdc's order); >1: partials
  double *rpf;        // [rsplit][3][N] partial real-space forces (FP64, rsplit > 1)
  // mixed path: the real-space force as 64-bit fixed point (x 2^32, two's
  // complement), [3][N]: each pair is evaluated once, the partner's share
  // and each item's home sums land here exactly and in any order;
  // k_force_vv reads it and clears it
  unsigned long long *fix;
  // ---- cell list ----
  int nc[3], ncell;
  int cell_hb;        // home blocks per cell item stride (item d: first block)
  double cellscale[3];  // nc / L
  int *cell_of;       // [N]
  int *cell_count;    // [ncell] counts, then the scatter's fill cursors
  int *cell_tmp;      // [N] ions grouped by cell in the scatter's (unordered) slots
  int *cell_start;    // [ncell + 1]
  int *sorted;        // [N] ion index in cell order (ascending within a cell)
  double *sx, *sy, *sz;
  int *styp;
  // ---- k space ----
  int nk, ncol, nseg, nquad, hmax;
  const KCol *cols;
  const KSeg *segs;
  const KQuad *quads;
  const int *khz;     // per k: hz
  const int *kslot;   // per k: slot (row * 4 + column in quad) of its weights
  const double *kce;  // per k: vp * exp(-pi^2 m^2/alpha^2) / m^2  (energy)
  const double *kcf;  // per k: v2 * exp(...) / m^2                  (force)
  void *tabx, *taby, *tabw;  // e^{2 pi i h r_a/L_a}, h = 0..hmax: [(hmax+1)][N] complex
  double2 *skpart;    // [nchunk][nk] partial structure factors
  int nchunk, chunk;  // ion chunks of the structure-factor pass
  double *kab;        // [slot][4]: cf*2Ss, cf*2Sc, hz*cf*2Ss, hz*cf*2Sc  (FP64)
  float *kab32;       // same, FP32 (mixed)
  int nslot;          // 4 * total quad rows (padding slots stay zero)
  double *kpot;       // [nk] energy terms
  double *fpart;      // [ngroup][3][N] back-projection partial forces
  int ngroup;
  // ---- back-projection on MFMA (mixed path) ----
  // W[hz] of each ion as FP16 hi + lo pairs in the B-operand order of
  // v_mfma_f32_32x32x16_f16: [hz/4][N] records of 8 hi then 8 lo halves
  // (re, im of 4 consecutive hz), hz < 8 * kbk, zero beyond hmax.
  void *wb;
  // Per tile (kf_tile_stride(kbk) 16-byte units): its epilogue records in
  // lane order [half g][j], column 2 j + g as (hx/sU, hy/sU, 1/sV,
  // |hx| | |hy| << 8), all zero for the padding columns of a short tile (the
  // signs of hx and hy are those of the first two fields); then its packed
  // A operands per K block, [kbk][hi, lo][64 lanes] 16-byte lane records,
  // rows (column, {Re U, Im U, Re V, Im V}).
  void *kfa;
  const KfTile *kft;  // [ntile]
  int kbk, ntile;
  // ---- structure factor on MFMA (mixed path) ----
  // Z*W[hz] as B operands of v_mfma_f32_32x32x16_f16: per group of 4 ions,
  // per hz < 16 * sknt, per output (Re S, Im S) a 32-byte record of 8 hi
  // then 8 lo halves, (p = re, im of Q) x 4 ions: for Re S (ZWr, -ZWi),
  // for Im S (ZWi, ZWr).  Ions padded to a multiple of 8 with zeros.
  void *sb;
  int sknt;           // hz tiles of 16
  int npad;           // N rounded up to a multiple of 8
  // ---- reductions ----
  double *rpart;      // [nreal][4]: pot, vir x, y, z per real-space item
  unsigned long long *rpairs;  // [nreal] accepted unordered pairs per item
  int nreal;
  double *spart;      // [nspart][9] per-block partial sums of the box scalars
  int nspart;         // ceil(N / kIonThreads)
  double *scal;       // kScal* layout below
  int *err;
  // ---- the per-step observables and thermostat of mdc's step loop
  //      (main.c:44-60): calc_msd, control_temp, moment_correction ----
  double *msd_old;    // [3][N] positions at the previous sample (msd.rx_old, ry_old, rz_old)
  double *msd_d;      // [3][N] summed minimum-image displacements (msd.dx, dy, dz)
  double *obs;        // kObs* layout below
  int msd_on;         // calc_msd every step
  int t_control_step; // control_temp period (ctl.t_control_step); 0: off
  long long t_control_until;  // control_temp only while step < this (main.c:57: 10000)
  double t_control_ke;        // target kinetic energy temp / e2t (pt.c:99)
  int moment_every;   // moment_correction when step % moment_every == 0; 0: off
};

// per-box observable state (device double array)
enum {
  kObsMsd = 0,        // [4] MSD per species (msd.value)
  kObsMsdIons = 4,    // [4] ions per species (msd.number_of_ion)
  kObsOldK = 8,       // control_temp's running kinetic-energy sum (its static oldk)
  kObsCount = 9,      // control_temp's sample count (its static count)
  kObsMsdInit = 10,   // 1 once the MSD origin is recorded
  kObsScalings = 11,  // velocity scalings done
  kObsLen = 12
};

// per-box scalar layout (device double array)
enum {
  kScalPot = 0, kScalPotReal, kScalPotRecip, kScalPotSelf, kScalEtot,
  kScalVirX, kScalVirY, kScalVirZ, kScalKin, kScalKinX, kScalKinY, kScalKinZ,
  kScalPres, kScalPresX, kScalPresY, kScalPresZ, kScalStep, kScalPairs,
  kScalCount
};

}  // namespace mdg
